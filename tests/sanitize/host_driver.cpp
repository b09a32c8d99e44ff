// Test infrastructure: the host side of the library (configuration and mesh readers, the BC
// rule, source tables, the formula evaluator, box meshes and the DG precompute of every
// (dim, order)) run under AddressSanitizer and UndefinedBehaviorSanitizer, no device code.
//   host_driver --formulas FILE [--table PATH]... [--config PATH]...
// FILE holds one "t<TAB>expression" per line. Exit status 0 when every call returned.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "dgfem_amd.h"

static int fails = 0;

static void check(bool ok, const std::string& what) {
    if (!ok) {
        std::fprintf(stderr, "FAIL %s\n", what.c_str());
        ++fails;
    }
}

int main(int argc, char** argv) {
    std::vector<std::string> configs, tables;
    std::string formulas;
    for (int i = 1; i + 1 < argc; i += 2) {
        const std::string k = argv[i];
        if (k == "--config") configs.push_back(argv[i + 1]);
        else if (k == "--table") tables.push_back(argv[i + 1]);
        else if (k == "--formulas") formulas = argv[i + 1];
    }
    // formulas
    int nform = 0;
    if (!formulas.empty()) {
        std::ifstream in(formulas);
        std::string line;
        while (std::getline(in, line)) {
            const size_t tab = line.find('\t');
            if (tab == std::string::npos) continue;
            (void)dg_eval_formula(line.substr(tab + 1).c_str(), std::stod(line.substr(0, tab)));
            ++nform;
        }
    }
    // source tables
    for (const std::string& p : tables) {
        int64_t n = 0;
        check(dg_read_source_table(p.c_str(), nullptr, 0, &n) == DG_OK && n > 0, "table size " + p);
        std::vector<double> rows(2 * n);
        int64_t m = 0;
        check(dg_read_source_table(p.c_str(), rows.data(), n, &m) == DG_OK && m == n, "table " + p);
    }
    // configuration files (reader, mesh, BC rule, sources, observers)
    std::vector<char> buf(1 << 20);
    for (const std::string& p : configs)
        check(dg_case_summary(p.c_str(), buf.data(), static_cast<int64_t>(buf.size())) == DG_OK, "config " + p);
    check(dg_case_summary("does/not/exist.conf", buf.data(), static_cast<int64_t>(buf.size())) != DG_OK,
          "missing config is an error");
    // reference operators, box meshes and the DG precompute
    for (int dim = 2; dim <= 3; ++dim)
        for (int p = 1; p <= 5; ++p) {
            int64_t K = 0;
            int32_t Np = 0;
            const int n = dim == 2 ? 3 : 2;
            check(dg_box_mesh_size(dim, n, p, &K, &Np) == DG_OK, "box size");
            std::vector<double> X(static_cast<size_t>(K) * Np * 3);
            std::vector<int64_t> V(static_cast<size_t>(K) * (dim + 1));
            std::vector<int8_t> side(static_cast<size_t>(K) * (dim + 1)), bc(side.size(), DG_BC_ABSORBING);
            check(dg_box_mesh_fill(dim, n, p, 0.15, 7, -1.0, 1.0, X.data(), V.data(), side.data()) == DG_OK, "box");
            for (size_t i = 0; i < side.size(); ++i)
                if (side[i] == 0) bc[i] = DG_BC_REFLECTING;
            const int Nfp = dim == 2 ? p + 1 : (p + 1) * (p + 2) / 2;
            std::vector<double> Dw(static_cast<size_t>(dim) * Np * Np), L(static_cast<size_t>(Np) * (dim + 1) * Nfp),
                Nr((dim + 1) * 3), M(static_cast<size_t>(Np) * Np);
            std::vector<int32_t> fm((dim + 1) * Nfp), lat(Np * 4);
            check(dg_ref_operators(dim, p, Dw.data(), L.data(), fm.data(), Nr.data(), lat.data()) == DG_OK, "ops");
            check(dg_ref_mass(dim, p, M.data()) == DG_OK, "mass");
            dg_mesh_desc md{dim, p, K, Np, X.data(), V.data(), bc.data()};
            std::vector<double> Jinv(static_cast<size_t>(K) * 9);
            std::vector<int32_t> nbr(static_cast<size_t>(K) * (dim + 1)), perm(6 * 64), canon(static_cast<size_t>(K) * Np);
            std::vector<uint8_t> code(nbr.size());
            check(dg_mesh_precompute(&md, Jinv.data(), nbr.data(), code.data(), perm.data(), canon.data()) == DG_OK,
                  "precompute dim " + std::to_string(dim) + " p " + std::to_string(p));
        }
    std::printf("host driver: %d formulas, %zu tables, %zu configs, %d failures\n", nform, tables.size(),
                configs.size(), fails);
    return fails == 0 ? 0 : 1;
}
