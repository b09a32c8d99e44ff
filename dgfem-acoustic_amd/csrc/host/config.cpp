// Configuration front-end. See config.hpp.
#include "config.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <fstream>
#include <sstream>
#include <stdexcept>

#include "formula.hpp"
#include "json.hpp"
#include "mesh.hpp"
#include "sources_io.hpp"

namespace dg {

bool SourceDef::valueAt(double t, double& v) const {
    switch (kind) {
        case Monopole:
            if (!(t < duration)) return false;
            v = amp * std::sin(2 * M_PI * freq * t + phase);
            return true;
        case Formula:
            if (!(t < duration)) return false;
            v = evalFormula(formula, t);
            return true;
        case Table:
            v = interpolateTable(table, t);
            return true;
    }
    return false;
}

namespace {

std::vector<std::string> split(const std::string& s, char d) {
    std::vector<std::string> out;
    std::stringstream ss(s);
    std::string tok;
    while (std::getline(ss, tok, d)) out.push_back(tok);
    return out;
}

std::string extension(const std::string& path) {
    const size_t dot = path.find_last_of('.');
    return dot == std::string::npos ? std::string() : path.substr(dot + 1);
}

std::string unquote(std::string s) {
    if (!s.empty() && s.front() == '"') s.erase(0, 1);
    if (!s.empty() && s.back() == '"') s.pop_back();
    return s;
}

double toDouble(const std::string& v, const std::string& key) {
    try {
        size_t pos = 0;
        double d = std::stod(v, &pos);
        return d;
    } catch (const std::exception&) {
        throw std::runtime_error("configuration value of '" + key + "' is not a number: '" + v + "'");
    }
}

// monopole / dipole / quadrupole expansion (src/configParser.cpp:76-116)
void addPoleSources(Config& c, const std::string& type, double x, double y, double z, double size, double amp,
                    double freq, double phase, double duration) {
    auto mk = [&](double pole, double xx, double yy, double sz, double ph) {
        SourceDef s;
        s.kind = SourceDef::Monopole;
        s.pole = pole;
        s.x = xx;
        s.y = yy;
        s.z = z;
        s.size = sz;
        s.amp = amp;
        s.freq = freq;
        s.phase = ph;
        s.duration = duration;
        c.sources.push_back(s);
    };
    if (type == "dipole") {
        mk(1, x - size, y, size / 2., phase);
        mk(1, x + size, y, size / 2., phase + M_PI);
    } else if (type == "quadrupole") {
        mk(2, x - size, y, size / 2., phase);
        mk(2, x + size, y, size / 2., phase);
        mk(2, x, y - size, size / 2., phase + M_PI);
        mk(2, x, y + size, size / 2., phase + M_PI);
    } else {
        mk(0, x, y, size, phase);
    }
}

SourceDef fileSource(const std::string& filename, double x, double y, double z, double size) {
    SourceDef s;
    s.kind = SourceDef::Table;
    s.pole = -1;
    s.x = x;
    s.y = y;
    s.z = z;
    s.size = size;
    const std::string ext = extension(filename);
    if (ext == "csv")
        s.table = parseCsvTable(filename, ';');
    else if (ext == "wav")
        s.table = parseWavTable(filename);
    else
        throw std::runtime_error("Not supported data");
    return s;
}

SourceDef formulaSource(const std::string& expr, double x, double y, double z, double size, double duration) {
    SourceDef s;
    s.kind = SourceDef::Formula;
    s.pole = -1;
    s.formula = expr;
    s.x = x;
    s.y = y;
    s.z = z;
    s.size = size;
    s.duration = duration;
    return s;
}

MshMesh openMesh(const std::string& path) {
    std::ifstream probe(path);
    if (!probe) throw std::runtime_error("Mesh file '" + path + "' not found or corrupted");
    return readMsh(path);
}

}  // namespace

Config parseConf(const std::string& path, MshMesh* meshOut) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("Project file '" + path + "' not found or corrupted");
    std::map<std::string, std::string> kv;  // ordered: sources/observers are visited in key order
    std::string line;
    while (std::getline(in, line)) {
        line.erase(std::remove_if(line.begin(), line.end(), [](unsigned char ch) { return std::isspace(ch); }),
                   line.end());
        if (line.empty() || line[0] == '#') continue;
        const size_t eq = line.find('=');
        kv[line.substr(0, eq)] = line.substr(eq == std::string::npos ? 0 : eq + 1);
    }
    auto get = [&](const std::string& k) -> std::string {
        auto it = kv.find(k);
        if (it == kv.end()) throw std::runtime_error("configuration key '" + k + "' missing in " + path);
        return it->second;
    };
    Config c;
    c.meshFileName = get("meshFileName");
    c.timeStart = toDouble(get("timeStart"), "timeStart");
    c.timeEnd = toDouble(get("timeEnd"), "timeEnd");
    c.timeStep = toDouble(get("timeStep"), "timeStep");
    c.timeRate = toDouble(get("timeRate"), "timeRate");
    c.elementType = get("elementType");
    c.timeIntMethod = get("timeIntMethod");
    c.numThreads = static_cast<int>(toDouble(get("numThreads"), "numThreads"));
    if (c.numThreads == 1) c.numThreads = 0;
    c.v0[0] = toDouble(get("v0_x"), "v0_x");
    c.v0[1] = toDouble(get("v0_y"), "v0_y");
    c.v0[2] = toDouble(get("v0_z"), "v0_z");
    c.rho0 = toDouble(get("rho0"), "rho0");
    c.c0 = toDouble(get("c0"), "c0");
    for (const auto& e : kv) {
        const std::string& key = e.first;
        if (key.find("source") == 0) {
            std::vector<std::string> s = split(e.second, ',');
            if (s.empty()) throw std::runtime_error("empty source definition '" + key + "'");
            auto num = [&](size_t i) {
                if (i >= s.size()) throw std::runtime_error("source '" + key + "' has too few fields");
                return toDouble(s[i], key);
            };
            if (s[0] != "formula" && s[0] != "file") {
                addPoleSources(c, s[0], num(1), num(2), num(3), num(4), num(5), num(6), num(7), num(8));
            } else if (s[0] == "formula") {
                if (s.size() < 2) throw std::runtime_error("source '" + key + "' has too few fields");
                c.sources.push_back(formulaSource(unquote(s[1]), num(2), num(3), num(4), num(5), num(6)));
            } else {
                if (s.size() < 2) throw std::runtime_error("source '" + key + "' has too few fields");
                c.sources.push_back(fileSource(unquote(s[1]), num(2), num(3), num(4), num(5)));
            }
        } else if (key.find("observer") == 0) {
            std::vector<std::string> s = split(e.second, ',');
            if (s.size() < 4) throw std::runtime_error("observer '" + key + "' has too few fields");
            c.observers.push_back({toDouble(s[0], key), toDouble(s[1], key), toDouble(s[2], key), toDouble(s[3], key)});
        } else if (key.find("initialCondtition") == 0) {
            std::vector<std::string> s = split(e.second, ',');
            if (s.size() < 6) throw std::runtime_error("initial condition '" + key + "' has too few fields");
            c.initConditions.push_back({0, toDouble(s[1], key), toDouble(s[2], key), toDouble(s[3], key),
                                        toDouble(s[4], key), toDouble(s[5], key)});
        }
    }
    MshMesh mesh = openMesh(c.meshFileName);
    for (const PhysicalGroup& g : mesh.groups) {
        if (g.dim != mesh.dim - 1) continue;
        auto it = kv.find(g.name);
        const std::string v = it == kv.end() ? std::string() : it->second;
        if (v.find("Absorbing") == 0)
            c.physBCs[g.tag] = {"Absorbing", 0};
        else if (v.find("Reflecting") == 0)
            c.physBCs[g.tag] = {"Reflecting", 0};
    }
    if (meshOut) *meshOut = std::move(mesh);
    return c;
}

Config parseJsonConfig(const std::string& path, MshMesh* meshOut) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("Project file '" + path + "' not found or corrupted");
    std::stringstream ss;
    ss << in.rdbuf();
    const JsonValue root = parseJson(ss.str());
    Config c;
    const JsonValue& meshJ = root.at("mesh");
    c.meshFileName = meshJ.at("File").string();
    const JsonValue& bcJ = meshJ.at("BC");
    (void)bcJ.at("number").number();
    MshMesh mesh = openMesh(c.meshFileName);
    int p = 0;
    for (const PhysicalGroup& g : mesh.groups) {
        if (g.dim != mesh.dim - 1) continue;
        ++p;
        const JsonValue* b = bcJ.find("boundary" + std::to_string(p));
        if (!b) continue;
        const JsonValue* name = b->find("name");
        const JsonValue* type = b->find("type");
        if (!name || !type || name->kind != JsonValue::String || name->str != g.name) continue;
        if (type->str == "Absorbing")
            c.physBCs[g.tag] = {"Absorbing", 0};
        else if (type->str == "Reflecting")
            c.physBCs[g.tag] = {"Reflecting", 0};
    }
    const JsonValue& solver = root.at("solver");
    const JsonValue& time = solver.at("time");
    c.timeStart = time.at("start").number();
    c.timeEnd = time.at("end").number();
    c.timeStep = time.at("step").number();
    c.timeRate = time.at("rate").number();
    c.elementType = solver.at("elementType").string();
    c.timeIntMethod = solver.at("timeIntMethod").string();
    c.numThreads = static_cast<int>(solver.at("numThreads").number());
    if (c.numThreads == 1) c.numThreads = 0;
    const JsonValue& init = root.at("initialization");
    const JsonValue& mf = init.at("meanFlow");
    c.v0[0] = mf.at("vx").number();
    c.v0[1] = mf.at("vy").number();
    c.v0[2] = mf.at("vz").number();
    c.rho0 = mf.at("rho").number();
    c.c0 = mf.at("c").number();
    const int nInit = static_cast<int>(init.at("number").number());
    for (int i = 0; i < nInit; ++i) {
        const JsonValue& ic = init.at("initialCondition" + std::to_string(i + 1));
        (void)ic.at("type").string();
        c.initConditions.push_back({0, ic.at("x").number(), ic.at("y").number(), ic.at("z").number(),
                                    ic.at("size").number(), ic.at("amplitude").number()});
    }
    const JsonValue& obs = root.at("observers");
    const int nObs = static_cast<int>(obs.at("number").number());
    for (int i = 0; i < nObs; ++i) {
        const JsonValue& o = obs.at("observer" + std::to_string(i + 1));
        c.observers.push_back({o.at("x").number(), o.at("y").number(), o.at("z").number(), o.at("size").number()});
    }
    const JsonValue& src = root.at("sources");
    const int nSrc = static_cast<int>(src.at("number").number());
    for (int i = 0; i < nSrc; ++i) {
        const JsonValue& s = src.at("source" + std::to_string(i + 1));
        const std::string type = s.at("type").string();
        const std::string fct = s.at("fct").string();
        const double x = s.at("x").number(), y = s.at("y").number(), z = s.at("z").number();
        const double size = s.at("size").number(), amp = s.at("amplitude").number();
        const double freq = s.at("frequency").number(), phase = s.at("phase").number();
        const double duration = s.at("duration").number();
        if (type != "formula" && type != "file") {
            addPoleSources(c, type, x, y, z, size, amp, freq, phase, duration);
        } else if (type == "formula") {
            std::string expr = fct;
            expr.erase(std::remove_if(expr.begin(), expr.end(), [](unsigned char ch) { return std::isspace(ch); }),
                       expr.end());
            c.sources.push_back(formulaSource(unquote(expr), x, y, z, size, duration));
        } else {
            c.sources.push_back(fileSource(unquote(fct), x, y, z, size));
        }
    }
    if (meshOut) *meshOut = std::move(mesh);
    return c;
}

Case loadCase(const std::string& configPath) {
    Case cs;
    const std::string ext = extension(configPath);
    if (ext == "conf")
        cs.config = parseConf(configPath, &cs.mesh);
    else if (ext == "json")
        cs.config = parseJsonConfig(configPath, &cs.mesh);
    else
        throw std::runtime_error("configuration file must end in .conf or .json: " + configPath);
    return cs;
}

std::vector<int8_t> boundaryTypes(const Case& c) {
    const MshMesh& m = c.mesh;
    const int dim = m.dim, nv = dim + 1;
    const int64_t K = m.numElements();
    std::vector<int64_t> verts(K * nv);
    for (int64_t e = 0; e < K; ++e)
        for (int v = 0; v < nv; ++v) verts[e * nv + v] = m.elementNodes[e * m.nodesPerElement + v];
    const std::vector<uint8_t> isB = boundaryFaceFlags(dim, K, verts.data());
    std::vector<std::vector<int64_t>> groupNodes;
    std::vector<int8_t> groupType;
    for (const auto& kv : c.config.physBCs) {  // ascending tag order
        for (const PhysicalGroup& g : m.groups)
            if (g.dim == dim - 1 && g.tag == kv.first) {
                groupNodes.push_back(g.nodeTags);
                groupType.push_back(kv.second.first == "Reflecting" ? 1 : 0);
            }
    }
    return resolveBoundaryTypesFromNodeSets(dim, K, verts.data(), isB, groupNodes, groupType);
}

}  // namespace dg
