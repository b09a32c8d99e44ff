// Gmsh MSH 4.0 / 4.1 ASCII reader. See msh.hpp.
#include "msh.hpp"

#include <algorithm>
#include <fstream>
#include <set>
#include <sstream>
#include <stdexcept>

namespace dg {

bool gmshElementInfo(int type, int& dim, int& order, int& numNodes) {
    switch (type) {
        case 15: dim = 0; order = 0; numNodes = 1; return true;
        case 1: dim = 1; order = 1; numNodes = 2; return true;
        case 8: dim = 1; order = 2; numNodes = 3; return true;
        case 26: dim = 1; order = 3; numNodes = 4; return true;
        case 27: dim = 1; order = 4; numNodes = 5; return true;
        case 28: dim = 1; order = 5; numNodes = 6; return true;
        case 2: dim = 2; order = 1; numNodes = 3; return true;
        case 9: dim = 2; order = 2; numNodes = 6; return true;
        case 21: dim = 2; order = 3; numNodes = 10; return true;
        case 23: dim = 2; order = 4; numNodes = 15; return true;
        case 25: dim = 2; order = 5; numNodes = 21; return true;
        case 4: dim = 3; order = 1; numNodes = 4; return true;
        case 11: dim = 3; order = 2; numNodes = 10; return true;
        case 29: dim = 3; order = 3; numNodes = 20; return true;
        case 30: dim = 3; order = 4; numNodes = 35; return true;
        case 31: dim = 3; order = 5; numNodes = 56; return true;
        default: return false;
    }
}

void MshMesh::solverArrays(std::vector<double>& coords, std::vector<int64_t>& verts) const {
    const int64_t K = numElements();
    coords.assign(K * nodesPerElement * 3, 0.0);
    verts.assign(K * (dim + 1), 0);
    for (int64_t e = 0; e < K; ++e) {
        for (int a = 0; a < nodesPerElement; ++a) {
            const int64_t tag = elementNodes[e * nodesPerElement + a];
            auto it = nodes.find(tag);
            if (it == nodes.end()) throw std::runtime_error("element node tag " + std::to_string(tag) + " undefined");
            for (int c = 0; c < 3; ++c) coords[(e * nodesPerElement + a) * 3 + c] = it->second[c];
        }
        for (int v = 0; v <= dim; ++v) verts[e * (dim + 1) + v] = elementNodes[e * nodesPerElement + v];
    }
}

MshMesh readMsh(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("Mesh file '" + path + "' not found or corrupted");
    MshMesh m;
    double version = 0;
    std::map<std::pair<int, int>, std::string> names;             // (dim, tag) -> name
    std::map<std::pair<int, int>, std::vector<int>> entityPhys;   // (dim, entity) -> physical tags
    struct Block {
        int dim, entity, type;
        std::vector<int64_t> tags, nodes;
        int nn;
    };
    std::vector<Block> blocks;
    std::string tok;
    while (in >> tok) {
        if (tok == "$MeshFormat") {
            int ft, ds;
            in >> version >> ft >> ds;
            if (ft != 0) throw std::runtime_error("binary MSH files are not supported: " + path);
            if (version < 4.0 || version >= 5.0)
                throw std::runtime_error("MSH version " + std::to_string(version) + " not supported (4.0/4.1)");
        } else if (tok == "$PhysicalNames") {
            int n;
            in >> n;
            std::string line;
            std::getline(in, line);
            for (int i = 0; i < n; ++i) {
                std::getline(in, line);
                std::istringstream ls(line);
                int d, t;
                ls >> d >> t;
                std::string rest;
                std::getline(ls, rest);
                const size_t q0 = rest.find('"'), q1 = rest.rfind('"');
                names[{d, t}] = (q0 != std::string::npos && q1 > q0) ? rest.substr(q0 + 1, q1 - q0 - 1) : rest;
            }
        } else if (tok == "$Entities") {
            int64_t np, nc, ns, nv;
            in >> np >> nc >> ns >> nv;
            const int64_t counts[4] = {np, nc, ns, nv};
            for (int d = 0; d < 4; ++d)
                for (int64_t i = 0; i < counts[d]; ++i) {
                    int tag;
                    double x;
                    in >> tag;
                    const int nbox = (d == 0 && version >= 4.1) ? 3 : 6;
                    for (int k = 0; k < nbox; ++k) in >> x;
                    int nphys;
                    in >> nphys;
                    std::vector<int> ph(nphys);
                    for (int k = 0; k < nphys; ++k) in >> ph[k];
                    entityPhys[{d, tag}] = ph;
                    if (d > 0) {
                        int nb, b;
                        in >> nb;
                        for (int k = 0; k < nb; ++k) in >> b;
                    }
                }
        } else if (tok == "$Nodes") {
            int64_t nblocks, nnodes;
            in >> nblocks >> nnodes;
            if (version >= 4.1) {
                int64_t mn, mx;
                in >> mn >> mx;
            }
            for (int64_t b = 0; b < nblocks; ++b) {
                int a1, a2, param;
                int64_t n;
                in >> a1 >> a2 >> param >> n;
                if (param) throw std::runtime_error("parametric node coordinates are not supported");
                if (version >= 4.1) {
                    std::vector<int64_t> tags(n);
                    for (auto& t : tags) in >> t;
                    for (int64_t i = 0; i < n; ++i) {
                        std::array<double, 3> x;
                        in >> x[0] >> x[1] >> x[2];
                        m.nodes[tags[i]] = x;
                    }
                } else {
                    for (int64_t i = 0; i < n; ++i) {
                        int64_t t;
                        std::array<double, 3> x;
                        in >> t >> x[0] >> x[1] >> x[2];
                        m.nodes[t] = x;
                    }
                }
            }
        } else if (tok == "$Elements") {
            int64_t nblocks, nel;
            in >> nblocks >> nel;
            if (version >= 4.1) {
                int64_t mn, mx;
                in >> mn >> mx;
            }
            for (int64_t b = 0; b < nblocks; ++b) {
                Block bl;
                int64_t n;
                if (version >= 4.1)
                    in >> bl.dim >> bl.entity >> bl.type >> n;
                else
                    in >> bl.entity >> bl.dim >> bl.type >> n;
                int d, o, nn;
                if (!gmshElementInfo(bl.type, d, o, nn))
                    throw std::runtime_error("element type " + std::to_string(bl.type) + " not supported");
                bl.nn = nn;
                bl.tags.resize(n);
                bl.nodes.resize(n * nn);
                for (int64_t i = 0; i < n; ++i) {
                    in >> bl.tags[i];
                    for (int k = 0; k < nn; ++k) in >> bl.nodes[i * nn + k];
                }
                blocks.push_back(std::move(bl));
            }
        }
    }
    if (!in.eof() && in.fail()) throw std::runtime_error("malformed MSH file: " + path);
    // solver elements: the first element type of the top dimension
    int top = 0;
    for (const auto& bl : blocks) top = std::max(top, bl.dim);
    if (top < 2) throw std::runtime_error("mesh has no 2D or 3D elements: " + path);
    int firstType = -1;
    for (const auto& bl : blocks)
        if (bl.dim == top) {
            firstType = bl.type;
            break;
        }
    int d, o, nn;
    gmshElementInfo(firstType, d, o, nn);
    m.dim = top;
    m.elementType = firstType;
    m.order = o;
    m.nodesPerElement = nn;
    for (const auto& bl : blocks)
        if (bl.dim == top && bl.type == firstType) {
            m.elementTags.insert(m.elementTags.end(), bl.tags.begin(), bl.tags.end());
            m.elementNodes.insert(m.elementNodes.end(), bl.nodes.begin(), bl.nodes.end());
        }
    // physical groups (sorted by (dim, tag)) and the node sets of their elements
    std::map<std::pair<int, int>, std::set<int64_t>> groupNodes;
    for (const auto& kv : entityPhys)
        for (int ph : kv.second) groupNodes[{kv.first.first, ph}];
    for (const auto& bl : blocks) {
        auto it = entityPhys.find({bl.dim, bl.entity});
        if (it == entityPhys.end()) continue;
        for (int ph : it->second) {
            auto& set = groupNodes[{bl.dim, ph}];
            set.insert(bl.nodes.begin(), bl.nodes.end());
        }
    }
    for (const auto& kv : groupNodes) {
        PhysicalGroup g;
        g.dim = kv.first.first;
        g.tag = kv.first.second;
        auto nm = names.find(kv.first);
        g.name = nm != names.end() ? nm->second : std::string();
        g.nodeTags.assign(kv.second.begin(), kv.second.end());
        m.groups.push_back(std::move(g));
    }
    return m;
}

}  // namespace dg
