/*
 * dgfem_amd — C ABI of the MI355X (gfx950) DG acoustic solver core.
 *
 * Drop-in for the time-integration call of the reference DGFEM-Acoustic solver:
 *   void solver::rungeKutta(std::vector<std::vector<double>>& u, Mesh& mesh, Config config)
 *   (include/solver.h:25, src/solver.cpp:311-584)
 * and its sibling solver::forwardEuler (include/solver.h:16, src/solver.cpp:66-302).
 * Plain pointers and sizes only; status codes instead of exit(); one context per host
 * thread; the context owns its device memory, its graphs and two streams: the main one, and
 * one for the end-of-step kernel's graph branch inside captured RK4 steps, which every
 * captured sequence joins back into the main stream before it ends.
 *
 * State layout at this boundary is the reference's: u[field][element * Np + node] with
 * fields (p, vx, vy, vz) — plus rho for the 5-field variant — and the element's node order
 * as in the mesh file (Gmsh element node order; the first dim+1 nodes are the vertices).
 */
#ifndef DGFEM_AMD_H
#define DGFEM_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DG_API __attribute__((visibility("default")))

typedef int32_t dg_status;
enum {
    DG_OK = 0,
    DG_ERR_INVALID = 1,     /* bad argument */
    DG_ERR_MESH = 2,        /* malformed / unsupported mesh */
    DG_ERR_DEVICE = 3,      /* HIP runtime error or no device */
    DG_ERR_NONFINITE = 4,   /* non-finite state or residual detected during dg_run */
    DG_ERR_UNSUPPORTED = 5, /* order / dimension / option not built */
    DG_ERR_CALLBACK = 6     /* a caller callback returned nonzero */
};

enum { DG_FP64 = 0, DG_MIXED = 1 };                 /* precision */
enum { DG_FIELDS_REF4 = 0, DG_FIELDS_LEE5 = 1 };    /* (p,vx,vy,vz) or + rho */
/* src/Mesh.cpp:575-578; DG_BC_LEARNED (2D only) has no counterpart in the reference: the
   face takes its exterior state from the learned closure network (dg_set_closure). */
enum { DG_BC_ABSORBING = 0, DG_BC_REFLECTING = 1, DG_BC_LEARNED = 2 };
enum { DG_RK4 = 0, DG_EULER = 1 };                  /* config timeIntMethod */
/* Stage-kernel variants (dg_info.stage_kernel): the VALU kernel (default), the VALU kernel with
   its volume contraction on the FP32/FP64 matrix cores (vm), the block and the
   wave-autonomous FP32/FP64 matrix-core kernels, and the split-BF16 matrix-core kernels of the
   mixed precision (operands as 3 BF16 terms, six products: FP32 operand precision; or 2 terms,
   three products). Environment DGFEM_KERNEL = valu | vm | mm | wave | bf16x6 | bf16x3; an explicit
   request that is not built for the mesh or precision fails dg_create with DG_ERR_UNSUPPORTED.
   Without it: bf16x6 for the mixed precision on tetrahedra of order 5, vm on tetrahedra of
   order 3, valu otherwise. */
enum { DG_KERNEL_VALU = 0, DG_KERNEL_MM = 1, DG_KERNEL_WAVE = 2, DG_KERNEL_BF16X6 = 3, DG_KERNEL_BF16X3 = 4,
       DG_KERNEL_VM = 5 };

typedef struct dg_context dg_context;

typedef struct {
    int32_t dim;                       /* 2 or 3 */
    int32_t order;                     /* Lagrange order p (1..5) */
    int64_t num_elements;              /* K */
    int32_t nodes_per_element;         /* Np, must equal the order's node count */
    const double* node_coords;         /* [K][Np][3] */
    const int64_t* element_vertices;   /* [K][dim+1] vertex ids (face matching) */
    const int8_t* face_bc;             /* [K][dim+1] BC of boundary faces per Gmsh local face
                                          (tri edges {0,1},{1,2},{2,0}; tet faces {0,2,1},{0,1,3},
                                          {0,3,2},{3,1,2}); NULL: all absorbing (reference default) */
} dg_mesh_desc;

typedef struct {
    double v0[3];    /* mean flow */
    double rho0;     /* mean density */
    double c0;       /* speed of sound */
    int32_t fields;  /* DG_FIELDS_REF4 | DG_FIELDS_LEE5 */
} dg_physics;

typedef struct {
    int32_t precision;        /* DG_FP64 | DG_MIXED */
    int32_t use_graph;        /* 1: replay captured hipGraphs per batch of steps */
    int32_t steps_per_graph;  /* steps captured per graph (>= 1) */
    int32_t device;           /* HIP device ordinal */
} dg_options;

/* Mesh::Mesh + Mesh::precomputeMassMatrix (src/Mesh.cpp:20-724): FP64 precompute on the
   host, upload of the device-resident operators and geometry. */
DG_API dg_status dg_create(const dg_mesh_desc* mesh, const dg_physics* phys, const dg_options* opt,
                           dg_context** out);
DG_API void dg_destroy(dg_context* ctx);
/* Message of the last failure on ctx; with ctx == NULL, of the last failed dg_create. */
DG_API const char* dg_last_error(const dg_context* ctx);

/* u = [fields][K*Np] (src/dgalerkin.cpp:51). */
DG_API dg_status dg_set_state(dg_context* ctx, const double* u);
DG_API dg_status dg_get_state(dg_context* ctx, double* u);

/* Initial condition p += amp * exp(-|x - xc|^2 / size) at every DG node
   (src/dgalerkin.cpp:51-70); for the 5-field variant rho += that / c0^2. */
DG_API dg_status dg_add_gaussian(dg_context* ctx, double x, double y, double z, double size, double amp);

/* Hard sources (src/solver.cpp:340-358 node search, :455-483 overwrite). Node set:
   DG nodes with |x - xs|^2 < size^2. Returns the number of nodes through *num_nodes. */
DG_API dg_status dg_add_source_monopole(dg_context* ctx, double x, double y, double z, double size, double amp,
                                        double freq, double phase, double duration, int64_t* num_nodes);
/* Formula / file sources: value(user, t, &v) is evaluated on the host for every step;
   it returns nonzero when the pressure at the source nodes is to be overwritten with v. */
typedef int (*dg_value_fn)(void* user, double t, double* value);
DG_API dg_status dg_add_source_table(dg_context* ctx, double x, double y, double z, double size, dg_value_fn fn,
                                     void* user, int64_t* num_nodes);

/* Observer (src/solver.cpp:361-383, :552-573): DG nodes with |x - xo| < size, inverse
   distance weights 1/(d^2 + 1e-12). */
DG_API dg_status dg_add_observer(dg_context* ctx, double x, double y, double z, double size, int64_t* num_nodes);

typedef struct {
    double t_start, t_end, dt, t_rate;  /* the loop of src/solver.cpp:413-414 */
    int32_t integrator;                 /* DG_RK4 | DG_EULER */
    int32_t compute_residuals;          /* 1: per-step residuals vs the last snapshot */
    int64_t max_steps;                  /* stop after this many steps (<0: no limit) */
} dg_run_params;

typedef struct {
    int64_t step;
    double t;                /* label of the step (time before the update) */
    const double* residuals; /* [5]: p, rho, vx, vy, vz */
    int32_t num_observers;
    const double* observers; /* [num_observers][4]: p, vx, vy, vz */
    double step_seconds;     /* device time of the step's batch divided by its steps */
} dg_step_record;

typedef int (*dg_step_fn)(void* user, const dg_step_record* rec);
/* Save steps (src/solver.cpp:421-453): state [fields][K*Np] before the source overwrite. */
typedef int (*dg_save_fn)(void* user, int64_t step, double t, const double* state);

/* solver::rungeKutta / solver::forwardEuler time loop. Callbacks may be NULL. Steps run on the
   device in batches (at most steps_per_graph steps, never across a save step); the step records
   of a batch are delivered after it. When a step callback returns nonzero (DG_ERR_CALLBACK) or a
   step leaves a non-finite value in the state or in a residual (DG_ERR_NONFINITE; the state is
   checked whether or not residuals are computed), *steps_done counts the steps delivered up to
   and including that step, and the state held by the context is the one at the end of its
   batch. */
DG_API dg_status dg_run(dg_context* ctx, const dg_run_params* prm, dg_step_fn on_step, dg_save_fn on_save,
                        void* user, int64_t* steps_done);

/* Convenience: run and write times [max], residuals [max][5], observers [max][nobs][4]. */
DG_API dg_status dg_run_collect(dg_context* ctx, const dg_run_params* prm, double* times, double* residuals,
                                double* observers, int64_t capacity, int64_t* steps_done);

/* One right-hand side: k = dt * M^-1 (S - F) for state u, i.e. numStep after updateFlux
   with beta = 0 (src/solver.cpp:38-57). u and k are [fields][K*Np]. */
DG_API dg_status dg_rhs(dg_context* ctx, const double* u, double* k, double dt);

typedef struct {
    int32_t dim, order, nodes_per_element, nodes_per_face, fields, precision;
    int64_t num_elements, num_faces, num_boundary_faces;
    double min_inradius;
    int64_t device_bytes;
    int32_t stage_kernel;        /* DG_KERNEL_* chosen by dg_create */
    int32_t post_branch;         /* 1: captured RK4 steps run the end-of-step kernel on a graph branch
                                    of its own (stage grids of >= 4096 blocks, no observer node that
                                    is also a source node, graphs on) */
    int32_t elements_per_block;  /* elements per stage-kernel block */
    int64_t stage_blocks;        /* blocks per stage launch */
} dg_info;
DG_API dg_status dg_get_info(const dg_context* ctx, dg_info* info);

/* Per-kernel device time of `steps` eager RK4 steps measured with HIP events on the
   context's stream: stage_ms[4] (mean per launch of stages 1..4), source_ms, post_ms.
   The steps are real: they advance the context's state by `steps` RK4 steps, with the
   source clock of those steps restarting at t = 0 and residuals against the current
   snapshot. */
DG_API dg_status dg_profile_steps(dg_context* ctx, double dt, int32_t steps, double* stage_ms, double* source_ms,
                                  double* post_ms);

/* Wall time in ms (HIP events) of the device work of the last dg_run. */
DG_API double dg_last_run_device_ms(const dg_context* ctx);

/* ---- learned absorbing-boundary closure (no counterpart in the reference; its
   characteristic absorbing BC, src/Mesh.cpp:619-666, stays the fallback) ----
   Network: n_in = dim Nfp + 2 -> 64 -> 64 -> n_out = Nfp, tanh, BF16 weights, FP32
   accumulation; parameters in the order W1[64][n_in], b1[64], W2[64][64], b2[64],
   W3[n_out][64], b3[n_out] (row-major, out x in). On a face with outward normal n and
   tangents t_c (2D: t = (-n_y, n_x); 3D: t1 along the tangential mean flow
   v0 - (v0.n) n when its norm exceeds 1e-6 c0, else along e_a - (e_a.n) n for the axis a least
   aligned with n, t2 = n x t1), per face node P = p/(rho0 c0), N = v.n, T_c = v.t_c and the
   characteristics w_out = P + N, w_in = P - N: scale s = sqrt(mean(w_out^2 + sum T_c^2))
   + 1e-30, x = [w_out/s, T_1/s, (T_2/s,) v0.n/c0, vt0/c0] with vt0 = v0.t (2D) or
   |v0 - (v0.n) n| (3D); the network predicts the incoming characteristic w_hat = s y, and
   the exterior state is chosen so that the face flux is the upwind flux with incoming
   values w_hat (acoustic) and 0 (vorticity). The zero network is the exact characteristic
   condition. Shapes built: 2D Nfp = 2..6, 3D Nfp = 3, 6, 10. */

/* Network of the context's DG_BC_LEARNED faces (default: all zero, i.e. exterior state =
   interior trace). params: BF16 bit patterns, n_params = dg_closure_num_params(...). */
DG_API dg_status dg_set_closure(dg_context* ctx, const uint16_t* params, int64_t n_params);
/* The learned faces in ghost order: element ids and local faces (capacity entries). */
DG_API dg_status dg_learned_faces(const dg_context* ctx, int64_t* elem, int32_t* face, int64_t capacity,
                                  int64_t* num_faces);
/* Traces of the current state on faces (elem[i], face[i]): out [n][2][Nfp][fields] holds the
   interior trace and the neighbour's trace in the interior face-node order (zeros on the
   boundary); normals [n][3] the outward unit normals. */
DG_API dg_status dg_face_traces(dg_context* ctx, int64_t n, const int64_t* elem, const int32_t* face, double* out,
                                double* normals);
/* No counterpart in the reference: dg_face_traces for callers that cannot allocate device
   memory (code under graph capture, tight per-step loops such as the closure's training-data
   collection, harnesses that forbid hipMalloc). Same arguments, results and checks; the device
   staging (element ids, faces, traces, normals) lives in a fixed 8 MiB array per device, held
   until the process ends, and the call allocates nothing. It accepts n faces when
   align256(8n) + align256(4n) + align256(16 n Nfp fields) + align256(24n) <= 8 MiB
   (align256: rounded up to a multiple of 256; 12,408 faces of order-3 tetrahedra with 4
   fields), else returns DG_ERR_INVALID. Calls on one device share the array: calls that
   overlap in time on one device (several host threads) race. */
DG_API dg_status dg_face_traces_noalloc(dg_context* ctx, int64_t n, const int64_t* elem, const int32_t* face,
                                        double* out, double* normals);

DG_API int64_t dg_closure_num_params(int32_t n_in, int32_t n_out);
DG_API const char* dg_closure_last_error(void);
/* FP32 -> BF16 bit patterns, round to nearest even. */
DG_API void dg_closure_to_bf16(const float* x, int64_t n, uint16_t* out);
/* Features X [n][n_in], targets Y [n][n_out] (the neighbour trace's incoming characteristic
   divided by s; may be NULL) and scales S [n] (may be NULL) from traces [n][2][nfp][nf] and
   normals [n][3] of dim-dimensional faces as returned by dg_face_traces. */
DG_API dg_status dg_closure_features(int32_t dim, int32_t nfp, int32_t nf, int64_t n, const double* traces,
                                     const double* normals, const dg_physics* phys, float* X, float* Y, float* S);
/* Inference Y [n][n_out] of X [n][n_in] on `device`. */
DG_API dg_status dg_closure_forward(int32_t n_in, int32_t n_out, const uint16_t* params, const float* X, int64_t n,
                                    float* Y, int32_t device);
/* Mean-square loss sum|y - Y|^2 / (n n_out) over all samples and its gradient grad[n_params]
   (FP32, parameter order as above). The forward and backward products run on the BF16 matrix
   cores in blocks of 32 samples (FP32 accumulation; activations and errors split into BF16
   terms: three for the forward products, two for the backward ones); environment
   DGFEM_CLOSURE_TRAIN=valu selects the one-sample-per-wave FP32 VALU kernel instead. */
DG_API dg_status dg_closure_loss_grad(int32_t n_in, int32_t n_out, const uint16_t* params, const float* X,
                                      const float* Y, int64_t n, double* loss, float* grad, int32_t device);
typedef struct {
    int32_t epochs;
    int32_t batch;             /* samples per Adam step */
    double lr, beta1, beta2, eps;
    uint64_t seed;             /* shuffling */
} dg_train_params;
/* Adam on FP32 master parameters (in/out) with BF16 forward weights (out: params). The
   forward and backward passes run on the GPU; epoch_loss[epochs] (may be NULL) receives
   the mean loss of each epoch. */
DG_API dg_status dg_closure_train(int32_t n_in, int32_t n_out, float* master, uint16_t* params, const float* X,
                                  const float* Y, int64_t n, const dg_train_params* prm, double* epoch_loss,
                                  int32_t device);

/* ---- host-side mesh helpers (no device needed) ---- */

/* Seeded jittered box mesh on [lo,hi]^dim (n cells per side; triangles or Kuhn tets).
   Sizes first, then fill caller-allocated arrays:
     node_coords [K][Np][3], element_vertices [K][dim+1], face_side [K][dim+1]
   (face_side: -1 interior, else side 2*axis + (0: lo, 1: hi)). */
DG_API dg_status dg_box_mesh_size(int32_t dim, int32_t n, int32_t order, int64_t* num_elements,
                                  int32_t* nodes_per_element);
DG_API dg_status dg_box_mesh_fill(int32_t dim, int32_t n, int32_t order, double jitter, uint64_t seed, double lo,
                                  double hi, double* node_coords, int64_t* element_vertices, int8_t* face_side);

/* The reference's boundary-type rule (src/Mesh.cpp:579-605) over physical groups given as
   side masks (ascending tag order): face_bc[K][dim+1]. */
DG_API dg_status dg_resolve_bc_sides(int32_t dim, int64_t num_elements, const int64_t* element_vertices,
                                     const int8_t* face_side, int32_t num_groups, const uint32_t* group_side_mask,
                                     const int8_t* group_type, int8_t* face_bc);

/* ---- front-end helpers (no device needed) ---- */

/* Value of a `formula` source expression at time t with the reference's evaluator semantics
   (src/eqEdit.cpp:107-311); -1 for malformed expressions, as the reference. */
DG_API double dg_eval_formula(const char* expression, double t);

/* (time, value) table of a `file` source (.csv: src/utils.cpp:138-177; .wav: src/utils.cpp:
   240-258). Writes up to `capacity` rows into rows[capacity][2]; *num_rows = total rows. */
DG_API dg_status dg_read_source_table(const char* path, double* rows, int64_t capacity, int64_t* num_rows);

/* Parse a .conf/.json configuration and its mesh (src/configParser.cpp:33-450), resolve the
   boundary types, and write a JSON summary into out[capacity] (keys: mesh, dim, order,
   elements, nodes_per_element, time, integrator, mean_flow, bc_counts, sources, observers,
   initial_conditions, physical_bcs, boundary_faces: [element, Gmsh local face, 0 absorbing |
   1 reflecting] per boundary face, listed for meshes of at most 4096 boundary faces, else
   empty). */
DG_API dg_status dg_case_summary(const char* config_path, char* out, int64_t capacity);

/* Diagnostics of the host precompute (no device needed), for tests of the host logic.
   Reference operators of (dim, p) in the canonical lattice order:
     Dw [dim][Np][Np] weak derivatives, LIFT [Np][(dim+1)*Nfp], fmask [(dim+1)][Nfp],
     Nref [(dim+1)][3], lattice [Np][4] integer barycentrics. Any pointer may be NULL. */
DG_API dg_status dg_ref_operators(int32_t dim, int32_t order, double* Dw, double* LIFT, int32_t* fmask, double* Nref,
                                  int32_t* lattice);
/* Reference mass matrix M [Np][Np] of (dim, p) on Gmsh's reference simplex, canonical order. */
DG_API dg_status dg_ref_mass(int32_t dim, int32_t order, double* M);
/* Mesh precompute (src/Mesh.cpp:20-686 counterpart): Jinv [K][3][3], neighbour [K][dim+1]
   (-1-bc on the boundary), neighbour code [K][dim+1] (local face | perm << 2),
   face_perm [nperm][Nfp], to_canonical [K][Np]. Any output pointer may be NULL. */
DG_API dg_status dg_mesh_precompute(const dg_mesh_desc* mesh, double* Jinv, int32_t* neighbour, uint8_t* code,
                                    int32_t* face_perm, int32_t* to_canonical);

#ifdef __cplusplus
}
#endif
#endif /* DGFEM_AMD_H */
