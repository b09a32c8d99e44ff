/*
 * CPU oracle for the DG right-hand side and RK4 / forward-Euler time loop of the
 * reference solver (src/solver.cpp:38-57, :66-302, :311-584; src/Mesh.cpp:20-975).
 *
 * TEST INFRASTRUCTURE ONLY. This code is the checker that the device path is compared
 * against; it is linked only by the test suite, the smoke check and the CPU-baseline leg
 * of the benchmark. The product library never calls it.
 *
 * It restates the reference's own algorithm (weak form evaluated with quadrature,
 * per-element inverse mass matrices, face integrals of the Rusanov-type flux, ghost
 * boundary fluxes, classic RK4 with the reference's stage structure) in plain C, with its
 * own Lagrange basis (Silvester's product formula on the equispaced lattice) and its own
 * quadrature (collapsed Gauss-Jacobi, exact to degree 2p+1), independent of the product's
 * nodal operators.
 */
#ifndef DG_ORACLE_H
#define DG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int dim;                     /* 2 or 3 */
    int order;                   /* p */
    int64_t num_elements;        /* K */
    int nodes_per_element;       /* Np */
    const double* node_coords;   /* [K][Np][3], Gmsh element node order (vertices first) */
    const int64_t* elem_verts;   /* [K][dim+1] global vertex ids */
    /* Boundary physical groups in ascending tag order: type 0 absorbing, 1 reflecting.
       group_nodes is the concatenation of sorted vertex-id sets, group_offsets[g]..[g+1]. */
    int num_groups;
    const int8_t* group_type;
    const int64_t* group_offsets;
    const int64_t* group_nodes;
} oracle_mesh;

typedef struct {
    double v0[3];
    double rho0;
    double c0;
} oracle_physics;

/* kind 0: monopole amp*sin(2 pi f t + phase) while t < duration (src/solver.cpp:458-467)
   kind 1: caller-evaluated value (formula / file sources, src/solver.cpp:468-482):
           value_fn returns nonzero to overwrite with *value at time t. */
typedef int (*oracle_value_fn)(void* user, int source, double t, double* value);

typedef struct {
    int kind;
    double x, y, z, size;
    double amp, freq, phase, duration;
} oracle_source;

typedef struct {
    double x, y, z, size;
} oracle_observer;

typedef struct {
    double t_start, t_end, dt, t_rate;
    int integrator;              /* 0 RK4, 1 forward Euler */
    int num_sources;
    const oracle_source* sources;
    oracle_value_fn value_fn;
    void* value_user;
    int num_observers;
    const oracle_observer* observers;
    int64_t max_steps;           /* capacity of the output arrays */
    int num_threads;             /* 0: OpenMP default */
} oracle_run_params;

typedef struct {
    int64_t steps;               /* steps executed */
    double* times;               /* [max_steps] label t of each step */
    double* residuals;           /* [max_steps][5]: p, rho, vx, vy, vz */
    double* observers;           /* [max_steps][num_observers][4]: p, vx, vy, vz */
    int64_t* save_steps;         /* [max_steps] steps at which a snapshot/VTU is taken */
    int64_t num_saves;
} oracle_run_output;

typedef struct oracle_ctx oracle_ctx;

/* Returns NULL on failure; oracle_error() then describes it. */
oracle_ctx* oracle_create(const oracle_mesh* mesh, const oracle_physics* phys);
void oracle_destroy(oracle_ctx* ctx);
const char* oracle_error(void);

/* One evaluation of k = dt * M^-1 (S - F) for every field (numStep with beta = 0 after
   updateFlux). u and k are [4][K*Np] in the caller's node order. */
int oracle_rhs(oracle_ctx* ctx, const double* u, double* k, double dt, int num_threads);

/* The time loop of rungeKutta / forwardEuler on state u ([4][K*Np], updated in place). */
int oracle_run(oracle_ctx* ctx, double* u, const oracle_run_params* prm, oracle_run_output* out);

/* Boundary flux at one face point (Mesh::updateFlux, src/Mesh.cpp:911-969): bc 1 reflecting
   (n . F of the trace with its normal velocity removed), bc 0 absorbing (RKR q with the
   matrix of src/Mesh.cpp:619-666 built from n, tangent t and bitangent s). */
void oracle_bc_flux(const oracle_physics* ph, int bc, const double n[3], const double t[3], const double s[3],
                    const double q[4], double out[4]);

/* Quadrature sizes used (for reports): volume and face points. */
int oracle_num_volume_points(const oracle_ctx* ctx);
int oracle_num_face_points(const oracle_ctx* ctx);
int64_t oracle_num_faces(const oracle_ctx* ctx);
/* Boundary type of local face lf of element e: 0 absorbing, 1 reflecting, -1 interior. */
int oracle_face_bc(const oracle_ctx* ctx, int64_t e, int lf);

#ifdef __cplusplus
}
#endif
#endif
