// dgalerkin: command-line driver with the reference's interface (src/dgalerkin.cpp:14-85):
//   dgalerkin <config.conf|config.json>
// Reads the configuration and mesh, applies the Gaussian initial conditions, runs the
// configured time integrator on the GPU through the C ABI, and writes the reference's
// outputs: residuals.csv, results/observers{i}.txt, results/result{step}.vtu at save
// steps, results/observer_{i}.wav / _fft.txt / _psd.txt, and results.pvd.
// DGFEM_PRECISION=mixed selects the mixed-precision path (default FP64).
#include <sys/stat.h>

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "../../../include/dgfem_amd.h"
#include "../host/config.hpp"
#include "../host/outputs.hpp"
#include "../host/sources_io.hpp"

namespace {

struct RunState {
    const dg::Case* cs = nullptr;
    std::ofstream residuals;
    std::vector<std::ofstream> observers;
    std::vector<std::vector<float>> series;
    double c0 = 1.0;
    bool euler = false;
};

int onStep(void* user, const dg_step_record* r) {
    auto* st = static_cast<RunState*>(user);
    st->residuals << dg::fmt6(r->t) << ";";
    for (int k = 0; k < 5; ++k) st->residuals << dg::fmt6(r->residuals[k]) << ";";
    st->residuals << dg::fmt6(r->step_seconds) << "\n";
    for (int o = 0; o < r->num_observers; ++o) {
        const double* ob = r->observers + 4 * o;
        const double p = ob[0], rho = p / (st->c0 * st->c0);
        st->observers[o] << dg::fmt6(r->t) << ";" << dg::fmt6(rho) << ";" << dg::fmt6(p) << ";" << dg::fmt6(ob[1]) << ";"
                         << dg::fmt6(ob[2]) << ";" << dg::fmt6(ob[3]) << "\n";
        st->series[o].push_back(static_cast<float>(p));
    }
    return 0;
}

int onSave(void* user, int64_t step, double t, const double* state) {
    auto* st = static_cast<RunState*>(user);
    std::cout << "[" << t << "/" << st->cs->config.timeEnd << "s] Step number : " << step << std::endl;
    const std::string name = "results/result" + std::to_string(step) + ".vtu";
    return dg::writeVtu(name, st->cs->mesh, state, st->c0) ? 0 : 1;
}

int tableSource(void* user, double t, double* value) {
    const auto* s = static_cast<const dg::SourceDef*>(user);
    return s->valueAt(t, *value) ? 1 : 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 2) return E2BIG;
    dg::Case cs;
    try {
        cs = dg::loadCase(argv[1]);
    } catch (const std::exception& ex) {
        std::cerr << "Fatal error: " << ex.what() << std::endl;
        return 1;
    }
    const dg::Config& c = cs.config;
    int integrator;
    if (c.timeIntMethod == "Euler1")
        integrator = DG_EULER;
    else if (c.timeIntMethod == "Runge-Kutta")
        integrator = DG_RK4;
    else {
        std::cerr << "Fatal error: Time integration method error" << std::endl;
        return 1;
    }
    std::vector<double> coords;
    std::vector<int64_t> verts;
    std::vector<int8_t> bc;
    try {
        cs.mesh.solverArrays(coords, verts);
        bc = dg::boundaryTypes(cs);
    } catch (const std::exception& ex) {
        std::cerr << "Fatal error: " << ex.what() << std::endl;
        return 1;
    }
    dg_mesh_desc desc{cs.mesh.dim, cs.mesh.order, cs.mesh.numElements(), cs.mesh.nodesPerElement, coords.data(),
                      verts.data(), bc.data()};
    dg_physics ph{{c.v0[0], c.v0[1], c.v0[2]}, c.rho0, c.c0, DG_FIELDS_REF4};
    const char* prec = std::getenv("DGFEM_PRECISION");
    dg_options opt{(prec && std::string(prec) == "mixed") ? DG_MIXED : DG_FP64, 1, 20, 0};
    dg_context* ctx = nullptr;
    if (dg_create(&desc, &ph, &opt, &ctx) != DG_OK) {
        std::cerr << "Fatal error: " << dg_last_error(nullptr) << std::endl;
        return 1;
    }
    for (const auto& ic : c.initConditions) dg_add_gaussian(ctx, ic[1], ic[2], ic[3], ic[4], ic[5]);
    for (const auto& s : c.sources) {
        int64_t n = 0;
        if (s.kind == dg::SourceDef::Monopole)
            dg_add_source_monopole(ctx, s.x, s.y, s.z, s.size, s.amp, s.freq, s.phase, s.duration, &n);
        else
            dg_add_source_table(ctx, s.x, s.y, s.z, s.size, tableSource, const_cast<dg::SourceDef*>(&s), &n);
    }
    for (const auto& o : c.observers) {
        int64_t n = 0;
        dg_add_observer(ctx, o[0], o[1], o[2], o[3], &n);
        if (n == 0) std::cerr << "warning: observer at (" << o[0] << "," << o[1] << "," << o[2] << ") has no nodes\n";
    }
    mkdir("results", 0755);
    RunState st;
    st.cs = &cs;
    st.c0 = c.c0;
    st.euler = integrator == DG_EULER;
    st.residuals.open("residuals.csv");
    st.residuals << "time;res_p;res_rho;res_vx;res_vy;res_vz;elapsed_time\n";
    st.observers.resize(c.observers.size());
    st.series.resize(c.observers.size());
    for (size_t o = 0; o < c.observers.size(); ++o) {
        st.observers[o].open("results/observers" + std::to_string(o + 1) + ".txt");
        st.observers[o] << (st.euler ? "time;pressure;density;velocity_x;velocity_y;velocity_z\n"
                                     : "time;density;pressure;velocity_x;velocity_y;velocity_z\n");
    }
    dg_run_params prm{c.timeStart, c.timeEnd, c.timeStep, c.timeRate, integrator, 1, -1};
    int64_t steps = 0;
    const dg_status rs = dg_run(ctx, &prm, onStep, onSave, &st, &steps);
    if (rs != DG_OK) std::cerr << "error: " << dg_last_error(ctx) << std::endl;
    for (size_t o = 0; o < c.observers.size(); ++o) {
        const std::string base = "results/observer_" + std::to_string(o + 1);
        dg::writeWav16(base + ".wav", st.series[o], static_cast<uint32_t>(1.0 / c.timeStep));
        dg::writeSpectra(st.series[o], c.timeStep, base);
    }
    dg::writePvd("results.pvd", c.timeStart, c.timeEnd, c.timeStep, c.timeRate);
    dg_destroy(ctx);
    std::cout << "steps: " << steps << std::endl;
    return rs == DG_OK ? EXIT_SUCCESS : 1;
}
