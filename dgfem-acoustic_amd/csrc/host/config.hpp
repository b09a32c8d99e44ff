// Solver configuration front-end: the reference's `.conf` and JSON files
// (src/configParser.cpp:33-450, include/configParser.h:17-115).
#pragma once
#include <array>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "msh.hpp"

namespace dg {

struct SourceDef {
    enum Kind { Monopole, Formula, Table } kind = Monopole;
    double pole = 0;  // 0 monopole, 1 dipole part, 2 quadrupole part, -1 formula/file
    double x = 0, y = 0, z = 0, size = 0;
    double amp = 0, freq = 0, phase = 0, duration = 0;
    std::string formula;
    std::vector<std::vector<double>> table;  // (time, value) rows
    // Value at t and whether the pressure is overwritten (src/solver.cpp:456-482).
    bool valueAt(double t, double& v) const;
};

struct Config {
    std::string meshFileName;
    double timeStart = 0, timeEnd = 1, timeStep = 0.1, timeRate = 0.1;
    std::string elementType = "Lagrange";
    std::string timeIntMethod = "Euler1";
    std::map<int, std::pair<std::string, double>> physBCs;  // physical tag -> (type, value)
    double v0[3] = {0, 0, 0};
    double rho0 = 1, c0 = 1;
    int numThreads = 1;
    std::vector<SourceDef> sources;
    std::vector<std::array<double, 4>> observers;       // x, y, z, size
    std::vector<std::array<double, 6>> initConditions;  // kind, x, y, z, size, amplitude
};

struct Case {
    Config config;
    MshMesh mesh;
};

// Parse a `.conf` (key = value) or `.json` configuration and read the mesh it names.
// Paths inside the configuration are taken relative to the current directory, as the
// reference does. Throws std::runtime_error with the reference's messages for missing files.
Case loadCase(const std::string& configPath);
Config parseConf(const std::string& path, MshMesh* meshOut);
Config parseJsonConfig(const std::string& path, MshMesh* meshOut);

// BC type per (element, local face) of the solver mesh from the physical groups of
// dimension dim-1 and config.physBCs (src/Mesh.cpp:579-605); absorbing by default.
std::vector<int8_t> boundaryTypes(const Case& c);

}  // namespace dg
