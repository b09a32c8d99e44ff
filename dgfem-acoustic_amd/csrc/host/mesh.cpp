// Host-side mesh generation, ingest and FP64 precompute. See mesh.hpp.
#include "mesh.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <map>
#include <numeric>
#include <stdexcept>
#include <string>
#include <unordered_set>

namespace dg {

namespace {

// Gmsh-style high-order node order on a simplex of order p (as barycentric tuples):
// vertices, then edge nodes along each local edge, then face-interior nodes (3D), then
// interior nodes, the interior recursing on the order p-(dim+1) sub-simplex.
void gmshOrderRec(int dim, int p, const std::array<int, 4>& shift, std::vector<std::array<int, 4>>& out) {
    if (p < 0) return;
    auto add = [&](std::array<int, 4> b) {
        for (int k = 0; k <= dim; ++k) b[k] += shift[k];
        out.push_back(b);
    };
    if (p == 0) {
        add({0, 0, 0, 0});
        return;
    }
    for (int v = 0; v <= dim; ++v) {
        std::array<int, 4> b = {0, 0, 0, 0};
        b[v] = p;
        add(b);
    }
    static const int e2[3][2] = {{0, 1}, {1, 2}, {2, 0}};
    static const int e3[6][2] = {{0, 1}, {1, 2}, {2, 0}, {3, 0}, {3, 2}, {3, 1}};
    int ne = (dim == 2) ? 3 : 6;
    for (int e = 0; e < ne; ++e) {
        int a = (dim == 2) ? e2[e][0] : e3[e][0];
        int b = (dim == 2) ? e2[e][1] : e3[e][1];
        for (int s = 1; s < p; ++s) {
            std::array<int, 4> t = {0, 0, 0, 0};
            t[a] = p - s;
            t[b] = s;
            add(t);
        }
    }
    if (dim == 2) {
        if (p >= 3) gmshOrderRec(2, p - 3, {shift[0] + 1, shift[1] + 1, shift[2] + 1, shift[3]}, out);
        return;
    }
    // face-interior nodes: order p-3 triangles on each face (Gmsh face table order)
    if (p >= 3) {
        for (int f = 0; f < 4; ++f) {
            const int* fv = kFaceVerts3D[f];
            std::vector<std::array<int, 4>> tri;
            gmshOrderRec(2, p - 3, {1, 1, 1, 0}, tri);
            for (auto& t : tri) {
                std::array<int, 4> b = {0, 0, 0, 0};
                for (int k = 0; k < 3; ++k) b[fv[k]] = t[k];
                add(b);
            }
        }
    }
    if (p >= 4) gmshOrderRec(3, p - 4, {shift[0] + 1, shift[1] + 1, shift[2] + 1, shift[3] + 1}, out);
}

inline int64_t baryKey(const std::array<int, 4>& b, int p) {
    return ((static_cast<int64_t>(b[0]) * (p + 1) + b[1]) * (p + 1) + b[2]) * (p + 1) + b[3];
}

}  // namespace

SimplexMesh generateBoxMesh(int dim, int n, int order, double jitter, uint64_t seed, double lo, double hi) {
    if ((dim != 2 && dim != 3) || n < 1 || order < 1) throw std::invalid_argument("generateBoxMesh: bad arguments");
    SimplexMesh m;
    m.dim = dim;
    m.order = order;
    const RefElement& R = refElement(dim, order);
    m.Np = R.Np;
    const double h = (hi - lo) / n;
    SplitMix64 rng(seed);
    const int n1 = n + 1;
    int64_t nv = (dim == 2) ? int64_t(n1) * n1 : int64_t(n1) * n1 * n1;
    m.vertexCoords.assign(nv * 3, 0.0);
    for (int64_t v = 0; v < nv; ++v) {
        int ijk[3] = {int(v % n1), int((v / n1) % n1), int(v / (int64_t(n1) * n1))};
        bool interior = true;
        for (int k = 0; k < dim; ++k) {
            m.vertexCoords[v * 3 + k] = lo + (hi - lo) * ijk[k] / n;
            if (ijk[k] == 0 || ijk[k] == n) interior = false;
        }
        if (interior && jitter > 0)
            for (int k = 0; k < dim; ++k) m.vertexCoords[v * 3 + k] += rng.uniform(-jitter * h, jitter * h);
    }
    auto vid2 = [&](int i, int j) { return int64_t(j) * n1 + i; };
    auto vid3 = [&](int i, int j, int k) { return (int64_t(k) * n1 + j) * n1 + i; };
    if (dim == 2) {
        m.K = int64_t(2) * n * n;
        m.elemVerts.reserve(m.K * 3);
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                int64_t a = vid2(i, j), b = vid2(i + 1, j), c = vid2(i + 1, j + 1), d = vid2(i, j + 1);
                bool diagBD = rng.uniform() < 0.5;
                int64_t t[2][3];
                if (!diagBD) {
                    int64_t t0[3] = {a, b, c}, t1[3] = {a, c, d};
                    std::copy(t0, t0 + 3, t[0]);
                    std::copy(t1, t1 + 3, t[1]);
                } else {
                    int64_t t0[3] = {a, b, d}, t1[3] = {b, c, d};
                    std::copy(t0, t0 + 3, t[0]);
                    std::copy(t1, t1 + 3, t[1]);
                }
                for (auto& tt : t) m.elemVerts.insert(m.elemVerts.end(), tt, tt + 3);
            }
    } else {
        m.K = int64_t(6) * n * n * n;
        m.elemVerts.reserve(m.K * 4);
        static const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
        static const int parity[6] = {+1, -1, -1, +1, +1, -1};
        for (int k = 0; k < n; ++k)
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    for (int q = 0; q < 6; ++q) {
                        int c[3] = {i, j, k};
                        int64_t v0 = vid3(c[0], c[1], c[2]);
                        c[perms[q][0]]++;
                        int64_t v1 = vid3(c[0], c[1], c[2]);
                        c[perms[q][1]]++;
                        int64_t v2 = vid3(c[0], c[1], c[2]);
                        int64_t v3 = vid3(i + 1, j + 1, k + 1);
                        if (parity[q] < 0) std::swap(v1, v2);
                        int64_t tt[4] = {v0, v1, v2, v3};
                        m.elemVerts.insert(m.elemVerts.end(), tt, tt + 4);
                    }
                }
    }
    // DG nodes in Gmsh-style order (vertices first).
    std::vector<std::array<int, 4>> go;
    gmshOrderRec(dim, order, {0, 0, 0, 0}, go);
    if (static_cast<int>(go.size()) != R.Np) throw std::logic_error("gmsh order size mismatch");
    const int nvpe = dim + 1;
    m.nodeCoords.assign(m.K * R.Np * 3, 0.0);
    m.faceSide.assign(m.K * nvpe, -1);
    for (int64_t e = 0; e < m.K; ++e) {
        const int64_t* ev = &m.elemVerts[e * nvpe];
        for (int a = 0; a < R.Np; ++a) {
            double x[3] = {0, 0, 0};
            for (int v = 0; v < nvpe; ++v) {
                double w = static_cast<double>(go[a][v]) / order;
                if (w == 0.0) continue;
                for (int c = 0; c < 3; ++c) x[c] += w * m.vertexCoords[ev[v] * 3 + c];
            }
            // exact copies at vertices
            for (int v = 0; v < nvpe; ++v)
                if (go[a][v] == order)
                    for (int c = 0; c < 3; ++c) x[c] = m.vertexCoords[ev[v] * 3 + c];
            std::copy(x, x + 3, &m.nodeCoords[(e * R.Np + a) * 3]);
        }
        for (int f = 0; f < nvpe; ++f) {
            int nfv = dim;
            for (int side = 0; side < 2 * dim; ++side) {
                int axis = side / 2;
                int target = (side % 2 == 0) ? 0 : n;
                bool all = true;
                for (int k = 0; k < nfv; ++k) {
                    int lv = (dim == 2) ? kFaceVerts2D[f][k] : kFaceVerts3D[f][k];
                    int64_t v = ev[lv];
                    int ijk[3] = {int(v % n1), int((v / n1) % n1), int(v / (int64_t(n1) * n1))};
                    if (ijk[axis] != target) { all = false; break; }
                }
                if (all) { m.faceSide[e * nvpe + f] = static_cast<int8_t>(side); break; }
            }
        }
    }
    return m;
}

std::vector<int8_t> resolveBoundaryTypes(const SimplexMesh& m, const std::vector<uint32_t>& groupSideMask,
                                         const std::vector<int8_t>& groupType) {
    const int nvpe = m.dim + 1;
    std::vector<std::vector<int64_t>> nodes(groupSideMask.size());
    for (size_t g = 0; g < groupSideMask.size(); ++g) {
        for (int64_t e = 0; e < m.K; ++e)
            for (int f = 0; f < nvpe; ++f) {
                int s = m.faceSide[e * nvpe + f];
                if (s < 0 || !(groupSideMask[g] & (1u << s))) continue;
                for (int k = 0; k < m.dim; ++k) {
                    int lv = (m.dim == 2) ? kFaceVerts2D[f][k] : kFaceVerts3D[f][k];
                    nodes[g].push_back(m.elemVerts[e * nvpe + lv]);
                }
            }
        std::sort(nodes[g].begin(), nodes[g].end());
        nodes[g].erase(std::unique(nodes[g].begin(), nodes[g].end()), nodes[g].end());
    }
    std::vector<uint8_t> isB(m.K * nvpe);
    for (int64_t i = 0; i < m.K * nvpe; ++i) isB[i] = m.faceSide[i] >= 0;
    return resolveBoundaryTypesFromNodeSets(m.dim, m.K, m.elemVerts.data(), isB, nodes, groupType);
}

std::vector<int8_t> resolveBoundaryTypesFromNodeSets(int dim, int64_t K, const int64_t* elemVerts,
                                                     const std::vector<uint8_t>& isBoundaryFace,
                                                     const std::vector<std::vector<int64_t>>& groupNodes,
                                                     const std::vector<int8_t>& groupType) {
    const int nvpe = dim + 1;
    std::vector<int8_t> bc(K * nvpe, 0);
    for (int64_t e = 0; e < K; ++e)
        for (int f = 0; f < nvpe; ++f) {
            if (!isBoundaryFace[e * nvpe + f]) continue;
            int lv0 = (dim == 2) ? kFaceVerts2D[f][0] : kFaceVerts3D[f][0];
            int64_t first = elemVerts[e * nvpe + lv0];
            int8_t t = 0;
            for (size_t g = 0; g < groupNodes.size(); ++g)
                if (std::binary_search(groupNodes[g].begin(), groupNodes[g].end(), first)) t = groupType[g];
            bc[e * nvpe + f] = t;
        }
    return bc;
}

std::vector<uint8_t> boundaryFaceFlags(int dim, int64_t K, const int64_t* elemVerts) {
    const int nvpe = dim + 1;
    std::vector<std::pair<std::array<int64_t, 3>, int64_t>> keys(K * nvpe);
    for (int64_t e = 0; e < K; ++e)
        for (int f = 0; f < nvpe; ++f) {
            std::array<int64_t, 3> k = {-1, -1, -1};
            for (int v = 0; v < dim; ++v) k[v] = elemVerts[e * nvpe + ((dim == 2) ? kFaceVerts2D[f][v] : kFaceVerts3D[f][v])];
            std::sort(k.begin(), k.begin() + dim);
            keys[e * nvpe + f] = {k, e * nvpe + f};
        }
    std::sort(keys.begin(), keys.end());
    std::vector<uint8_t> flags(K * nvpe, 0);
    for (size_t i = 0; i < keys.size();) {
        size_t j = i + 1;
        while (j < keys.size() && keys[j].first == keys[i].first) ++j;
        if (j - i == 1) flags[keys[i].second] = 1;
        i = j;
    }
    return flags;
}

DGMesh buildDGMesh(int dim, int order, int64_t K, const double* nodeCoords, const int64_t* elemVerts,
                   const int8_t* faceBC) {
    if (dim != 2 && dim != 3) throw std::runtime_error("dimension must be 2 or 3");
    if (K <= 0) throw std::runtime_error("mesh has no elements");
    const RefElement& R = refElement(dim, order);
    DGMesh m;
    m.dim = dim;
    m.p = order;
    m.Np = R.Np;
    m.Nfp = R.Nfp;
    m.Nfaces = R.Nfaces;
    m.K = K;
    const int Np = R.Np, nvpe = dim + 1, p = order;
    std::map<int64_t, int> canon;
    for (int n = 0; n < Np; ++n) canon[baryKey(R.bary[n], p)] = n;

    m.nodes.assign(K * Np * 3, 0.0);
    m.toCanon.assign(K * Np, -1);
    m.Jinv.assign(K * 9, 0.0);
    m.detJ.assign(K, 0.0);
    m.inradius.assign(K, 0.0);
    std::string err;
    for (int64_t e = 0; e < K; ++e) {
        const double* X = nodeCoords + e * Np * 3;
        double J[3][3] = {{0}};
        for (int r = 0; r < dim; ++r)
            for (int c = 0; c < 3; ++c) J[c][r] = X[(r + 1) * 3 + c] - X[c];
        double det, Ji[3][3] = {{0}};
        if (dim == 2) {
            det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
            Ji[0][0] = J[1][1] / det;
            Ji[0][1] = -J[0][1] / det;
            Ji[1][0] = -J[1][0] / det;
            Ji[1][1] = J[0][0] / det;
        } else {
            det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                  J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
            Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
            Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
            Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
            Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
            Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
            Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
            Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
            Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
            Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
        }
        if (!(std::fabs(det) > 0.0) || !std::isfinite(det)) {
            err = "degenerate element " + std::to_string(e);
            continue;
        }
        m.detJ[e] = det;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) m.Jinv[e * 9 + r * 3 + c] = Ji[r][c];
        // Identify each node on the equispaced lattice through the inverse affine map.
        for (int a = 0; a < Np; ++a) {
            double d[3] = {X[a * 3] - X[0], X[a * 3 + 1] - X[1], X[a * 3 + 2] - X[2]};
            double lam[4] = {1, 0, 0, 0};
            std::array<int, 4> b = {0, 0, 0, 0};
            int sum = 0;
            bool ok = true;
            for (int r = 0; r < dim; ++r) {
                double xi = 0;
                for (int c = 0; c < dim; ++c) xi += Ji[r][c] * d[c];
                lam[r + 1] = xi;
                lam[0] -= xi;
            }
            for (int k = 0; k <= dim; ++k) {
                double s = lam[k] * p;
                int ri = static_cast<int>(std::lround(s));
                if (std::fabs(s - ri) > 1e-6 || ri < 0 || ri > p) ok = false;
                b[k] = ri;
                sum += ri;
            }
            auto it = canon.find(baryKey(b, p));
            if (!ok || sum != p || it == canon.end()) {
                err = "element " + std::to_string(e) + " node " + std::to_string(a) +
                      " is not on the straight-sided equispaced lattice";
                break;
            }
            m.toCanon[e * Np + a] = it->second;
            std::copy(X + a * 3, X + a * 3 + 3, &m.nodes[(e * Np + it->second) * 3]);
        }
        // inradius
        double meas = std::fabs(det) / (dim == 2 ? 2.0 : 6.0), bnd = 0;
        for (int f = 0; f < nvpe; ++f) {
            double nn = 0;
            for (int c = 0; c < dim; ++c) {
                double s = 0;
                for (int r = 0; r < dim; ++r) s += Ji[r][c] * R.Nref[f * 3 + r];
                nn += s * s;
            }
            bnd += std::sqrt(nn) * std::fabs(det) * R.faceParamMeasure;
        }
        m.inradius[e] = dim * meas / bnd;
    }
    if (!err.empty()) throw std::runtime_error(err);
    for (int64_t e = 0; e < K; ++e) {
        std::vector<char> seen(Np, 0);
        for (int a = 0; a < Np; ++a) {
            int c = m.toCanon[e * Np + a];
            if (c < 0 || seen[c]) throw std::runtime_error("element " + std::to_string(e) + " has repeated nodes");
            seen[c] = 1;
        }
    }

    // Face pairing through sorted vertex keys.
    struct FK {
        std::array<int64_t, 3> key;
        int64_t ef;
    };
    std::vector<FK> keys(K * nvpe);
    for (int64_t e = 0; e < K; ++e)
        for (int f = 0; f < nvpe; ++f) {
            std::array<int64_t, 3> k = {-1, -1, -1};
            for (int v = 0; v < dim; ++v) {
                int lv = (dim == 2) ? kFaceVerts2D[f][v] : kFaceVerts3D[f][v];
                k[v] = elemVerts[e * nvpe + lv];
            }
            std::sort(k.begin(), k.begin() + dim);
            keys[e * nvpe + f] = {k, e * nvpe + f};
        }
    std::sort(keys.begin(), keys.end(), [](const FK& a, const FK& b) {
        return a.key != b.key ? a.key < b.key : a.ef < b.ef;
    });
    m.nbr.assign(K * nvpe, -1);
    m.nbrCode.assign(K * nvpe, 0);
    m.isBoundary.assign(K * nvpe, 0);
    auto faceVerts = [&](int64_t e, int f, int64_t out[3]) {
        for (int v = 0; v < dim; ++v) {
            int lv = (dim == 2) ? kFaceVerts2D[f][v] : kFaceVerts3D[f][v];
            out[v] = elemVerts[e * nvpe + lv];
        }
    };
    int64_t nf = 0, nb = 0;
    for (size_t i = 0; i < keys.size();) {
        size_t j = i + 1;
        while (j < keys.size() && keys[j].key == keys[i].key) ++j;
        if (j - i > 2) throw std::runtime_error("non-manifold face shared by more than two elements");
        ++nf;
        if (j - i == 1) {
            int64_t e = keys[i].ef / nvpe;
            int f = int(keys[i].ef % nvpe);
            int8_t bc = faceBC ? faceBC[e * nvpe + f] : 0;
            if (bc < 0 || bc > 2)
                throw std::runtime_error("boundary type must be 0 (absorbing), 1 (reflecting) or 2 (learned)");
            m.nbr[e * nvpe + f] = -1 - bc;
            m.isBoundary[e * nvpe + f] = 1;
            ++nb;
        } else {
            for (int side = 0; side < 2; ++side) {
                int64_t ea = keys[i + side].ef / nvpe, eb = keys[i + 1 - side].ef / nvpe;
                int fa = int(keys[i + side].ef % nvpe), fb = int(keys[i + 1 - side].ef % nvpe);
                int64_t A[3] = {0, 0, 0}, B[3] = {0, 0, 0};
                faceVerts(ea, fa, A);
                faceVerts(eb, fb, B);
                int code = -1;
                for (int c = 0; c < R.nFacePerms && code < 0; ++c) {
                    bool match = true;
                    for (int k = 0; k < dim; ++k)
                        if (B[k] != A[R.permVerts[c][k]]) match = false;
                    if (match) code = c;
                }
                if (code < 0) throw std::runtime_error("face vertex permutation not found");
                m.nbr[ea * nvpe + fa] = static_cast<int32_t>(eb);
                m.nbrCode[ea * nvpe + fa] = static_cast<uint8_t>(fb | (code << 2));
            }
        }
        i = j;
    }
    m.numFaces = nf;
    m.numBoundaryFaces = nb;
    // learned-closure faces, numbered in (element, local face) order: nbr = -3 - index
    for (int64_t ef = 0; ef < K * nvpe; ++ef)
        if (m.nbr[ef] == -3) {
            m.nbr[ef] = -3 - static_cast<int32_t>(m.learnedFaces.size());
            m.learnedFaces.push_back(ef);
        }
    return m;
}

void faceGeometry(const DGMesh& m, int64_t e, int f, double n[3], double& fscale) {
    const RefElement& R = refElement(m.dim, m.p);
    double nn = 0;
    for (int c = 0; c < 3; ++c) {
        double s = 0;
        if (c < m.dim)
            for (int r = 0; r < m.dim; ++r) s += m.Jinv[e * 9 + r * 3 + c] * R.Nref[f * 3 + r];
        n[c] = s;
        nn += s * s;
    }
    fscale = std::sqrt(nn);
    for (int c = 0; c < 3; ++c) n[c] /= fscale;
}

}  // namespace dg
