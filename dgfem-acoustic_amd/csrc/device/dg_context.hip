// Context behind the C ABI: FP64 host precompute -> device-resident SoA, kernel dispatch,
// the reference time loop (src/solver.cpp:311-584) driven in batches of steps between save
// points, with per-step residual / observer records returned through device rings.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <mutex>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../../include/dgfem_amd.h"
#include "../host/mesh.hpp"
#include "../host/refelem.hpp"
#include "dg_closure.hpp"
#include "dg_kernels.hpp"
#include "dg_kernels_bf16.hpp"

namespace {

thread_local std::string g_createError;

// Staging memory of dg_face_traces_noalloc: one array per device, held until the process ends,
// shared by every call on that device (overlapping calls on one device race). Host code uses
// only the per-device address obtained with hipGetSymbolAddress, never the array's name.
constexpr size_t kScratchBytes = size_t(8) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
std::mutex g_scratchMutex;
std::array<void*, kMaxDevices> g_scratchPtr{};

// Device address of the staging array on `device` (which must be current), looked up once per
// device.
hipError_t scratchAddress(int device, void** out) {
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    if (!g_scratchPtr[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratchPtr[device] = ptr;
    }
    *out = g_scratchPtr[device];
    return hipSuccess;
}

size_t align256(size_t n) { return (n + 255) / 256 * 256; }

// Staging bytes of dg_face_traces_noalloc for n faces: element ids, local faces, traces, normals.
size_t traceScratchBytes(int64_t n, int nfp, int nf) {
    const size_t m = static_cast<size_t>(n);
    return align256(8 * m) + align256(4 * m) + align256(16 * m * nfp * nf) + align256(24 * m);
}


// Lowest tetrahedral order whose mixed-precision default is the split-BF16 kernel.
constexpr int kBf16DefaultMinOrder = 5;

// A stage-kernel variant that was asked for but is not built for the mesh.
struct UnsupportedError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

struct TableSource {
    dg_value_fn fn;
    void* user;
};

}  // namespace

struct dg_context;
using StageFn = void (*)(dg_context*, int stage, const void* qin, void* qout, void* uout, bool snap);
using PerStepFn = void (*)(dg_context*);
using ClosureFn = void (*)(dg_context*, const void* qin);
using TraceFn = void (*)(dg_context*, int64_t n, const int64_t* elem, const int32_t* face, double* out,
                         double* normals);

struct dg_context {
    std::string err;
    int device = 0;
    hipStream_t stream = nullptr;
    int dim = 0, p = 0, Np = 0, Nfp = 0, Nfc = 0, nf = 4, precision = DG_FP64;
    int64_t K = 0;
    size_t tsize = 8;
    dg::DGMesh mesh;
    dg_physics phys{};
    dg_options opt{};
    // device state
    void *u = nullptr, *ulo = nullptr, *qa = nullptr, *qb = nullptr, *acc = nullptr, *snap = nullptr,
         *geo = nullptr, *fgeo = nullptr, *ops = nullptr;
    int32_t *nbr = nullptr, *iops = nullptr;
    uint8_t* code = nullptr;
    double* resPartial = nullptr;
    double* resGroup = nullptr;  // per-group residual sums when nblocks > 4 kResGroup
    int64_t nblocks = 0, nResGroups = 0;
    // sources / observers (host definitions + device tables)
    std::vector<std::vector<int64_t>> srcNodes;
    std::vector<std::array<double, 5>> srcParams;  // kind, amp, freq, phase, duration
    std::vector<TableSource> srcTables;            // per source (fn null for monopoles)
    std::vector<std::vector<int64_t>> obsNodes;
    std::vector<std::vector<double>> obsW;
    int64_t *dSrcNodes = nullptr, *dSrcCoverOff = nullptr, *dObsIdx = nullptr;
    int32_t *dSrcLast = nullptr, *dSrcCover = nullptr, *dObsSeg0 = nullptr;
    int64_t nSrcNodes = 0;
    double* dSegPartial = nullptr;
    int nseg = 0;
    double *dSrcParams = nullptr, *dObsW = nullptr;
    double *dT = nullptr, *dV = nullptr;
    uint8_t* dVF = nullptr;
    double *resRing = nullptr, *obsRing = nullptr;
    int* nonfinite = nullptr;
    int64_t* counter = nullptr;
    int64_t* batchSteps = nullptr;
    int64_t* srcBase = nullptr;  // steps of the batch done by earlier captured sequences
    int64_t ringSlots = 0;
    // pinned staging
    double *hT = nullptr, *hV = nullptr, *hRes = nullptr, *hObs = nullptr;
    uint8_t* hVF = nullptr;
    bool tablesDirty = true;
    // graphs
    hipGraphExec_t gMulti = nullptr, gOne = nullptr;
    int gSteps = 0, gIntegrator = -1, gResid = -1;
    hipStream_t side = nullptr;  // branch of the end-of-step kernel in captured RK4 sequences
    hipEvent_t evFork = nullptr, evJoin = nullptr;
    hipStream_t postStream = nullptr;  // stream of the next launchPost (null: the context stream)
    bool postSources = true;           // launchPost also overwrites the next step's sources
    // dispatch
    StageFn stageFn = nullptr;
    int eb = 1;              // elements per block (batch) of the chosen stage kernel
    bool mfma = false;       // matrix-core stage kernel in use
    int kernelKind = DG_KERNEL_VALU;  // stage kernel chosen by dg_create (reported by dg_get_info)
    void* afrag = nullptr;   // A-operand fragments of the matrix-core kernel
    void* afragBf = nullptr;  // split-BF16 operator fragments (uint16) of the BF16 matrix-core kernel
    std::vector<uint16_t> afragBfHost;
    int64_t bfGroups = 0;     // 16-element groups of the BF16 matrix-core kernel (persistent grid), or 0
    int bfBlocksPerCu = 0;    // its resident blocks per CU
    int bfSplit = 0;          // BF16 terms per operand (2 or 3)
    std::vector<double> afragHost;
    PerStepFn sourceFn = nullptr, postFn = nullptr;
    int64_t deviceBytes = 0;
    double lastRunMs = 0.0;
    // learned closure: faces, exterior states, BF16 network
    int64_t nLearned = 0;
    int64_t* dLearned = nullptr;
    void* ghost = nullptr;
    uint16_t* dClosureW = nullptr;
    int64_t closureParams = 0;
    ClosureFn closureFn = nullptr;
    TraceFn traceFn = nullptr;
    // per-run settings read by the launchers
    int curResid = 1;
    double curDt = 0.0, gDt = -1.0;
    void* scratch = nullptr;  // this device's staging array (dg_face_traces_noalloc)
};

namespace {

#define HIPCHK(call)                                                                    \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            throw std::runtime_error(std::string(#call) + ": " + hipGetErrorString(e_)); \
        }                                                                               \
    } while (0)

template <typename T>
T* devAlloc(dg_context* c, size_t n) {
    void* p = nullptr;
    if (n == 0) n = 1;
    HIPCHK(hipMalloc(&p, n * sizeof(T)));
    c->deviceBytes += static_cast<int64_t>(n * sizeof(T));
    return static_cast<T*>(p);
}

void devFree(void* p) {
    if (p) (void)hipFree(p);
}

template <typename T>
dgk::Physics<T> makePhysics(const dg_physics& ph) {
    dgk::Physics<T> o;
    for (int k = 0; k < 3; ++k) o.v0[k] = static_cast<T>(ph.v0[k]);
    o.rho0 = static_cast<T>(ph.rho0);
    o.c0 = static_cast<T>(ph.c0);
    o.rc2 = static_cast<T>(ph.rho0 * ph.c0 * ph.c0);
    o.irho0 = static_cast<T>(1.0 / ph.rho0);
    o.irc0 = static_cast<T>(1.0 / (ph.rho0 * ph.c0));
    o.ic2 = static_cast<T>(1.0 / (ph.c0 * ph.c0));
    return o;
}

// Arguments of one stage launch: stage input qin, stage output qout (next stage input), final
// state uout, residual snapshot on or off.
template <typename T, bool COMP>
dgk::StageArgs<T> stageArgs(dg_context* c, int stage, const void* qin, void* qout, void* uout, bool snap) {
    dgk::StageArgs<T> a{};
    a.qin = static_cast<const T*>(qin);
    a.qout = static_cast<T*>(qout);
    a.u = static_cast<T*>(c->u);
    a.uout = static_cast<T*>(uout);
    a.ulo = COMP ? static_cast<T*>(c->ulo) : nullptr;
    a.acc = static_cast<T*>(c->acc);
    a.geo = static_cast<const T*>(c->geo);
    a.fgeo = static_cast<const T*>(c->fgeo);
    a.nbr = c->nbr;
    a.ghost = static_cast<const T*>(c->ghost);
    a.code = c->code;
    a.ops = static_cast<const T*>(c->ops);
    a.iops = c->iops;
    a.vfrag = static_cast<const T*>(c->afrag);
    a.snap = snap ? static_cast<const T*>(c->snap) : nullptr;
    a.resPartial = c->resPartial;
    a.stateBad = c->nonfinite + 1;
    a.K = c->K;
    a.dt = static_cast<T>(c->curDt);
    a.stage = stage;
    a.ph = makePhysics<T>(c->phys);
    return a;
}

template <int DIM, int P, int NF, typename T, bool COMP, bool WIDE, bool VM = false>
void launchStage(dg_context* c, int stage, const void* qin, void* qout, void* uout, bool snap) {
    using dgk::stageKernel;
    const dgk::StageArgs<T> a = stageArgs<T, COMP>(c, stage, qin, qout, uout, snap);
    const dim3 grid(static_cast<unsigned>(c->nblocks)), block(dgk::StageWaves<DIM, P, NF, T, WIDE>::kBlockSize);
    switch (dgk::stageMode(stage)) {
        case dgk::kModeFirst:
            hipLaunchKernelGGL((stageKernel<DIM, P, NF, T, COMP, dgk::kModeFirst, WIDE, VM>), grid, block, 0, c->stream, a);
            break;
        case dgk::kModeMid:
            hipLaunchKernelGGL((stageKernel<DIM, P, NF, T, COMP, dgk::kModeMid, WIDE, VM>), grid, block, 0, c->stream, a);
            break;
        case dgk::kModeFinal:
            hipLaunchKernelGGL((stageKernel<DIM, P, NF, T, COMP, dgk::kModeFinal, WIDE, VM>), grid, block, 0, c->stream, a);
            break;
        case dgk::kModeEuler:
            hipLaunchKernelGGL((stageKernel<DIM, P, NF, T, COMP, dgk::kModeEuler, WIDE, VM>), grid, block, 0, c->stream, a);
            break;
        default:
            hipLaunchKernelGGL((stageKernel<DIM, P, NF, T, COMP, dgk::kModeRHS, WIDE, VM>), grid, block, 0, c->stream, a);
            break;
    }
}

// 512-thread stage blocks need a grid of at least this many of them (4 per CU); smaller grids
// run 256-thread blocks, which spread the same work over twice the blocks
constexpr int64_t kWideMinBlocks = 1024;

template <int DIM, int P, int NF, typename T, bool COMP, bool VM = false>
void pickStage(dg_context* c) {
    using Wd = dgk::StageWaves<DIM, P, NF, T, true>;
    constexpr int ebWide = VM ? dgk::VmShape<DIM, P, NF, T, true>::EB : Wd::EB;
    constexpr int ebNarrow = VM ? dgk::VmShape<DIM, P, NF, T, false>::EB : dgk::StageWaves<DIM, P, NF, T, false>::EB;
    if constexpr (Wd::kBig) {
        if ((c->K + ebWide - 1) / ebWide < kWideMinBlocks) {
            c->eb = ebNarrow;
            c->stageFn = &launchStage<DIM, P, NF, T, COMP, false, VM>;
            return;
        }
    }
    c->eb = ebWide;
    c->stageFn = &launchStage<DIM, P, NF, T, COMP, true, VM>;
}

// Instances of the matrix-core volume variant of the stage kernel: tetrahedra of order 2..3.
template <int DIM, int P, int NF>
constexpr bool vmBuilt() {
    return DIM == 3 && (P == 2 || P == 3) && (NF == 4 || NF == 5) &&
           dgk::VmShape<DIM, P, NF, float, true>::kSupported && dgk::VmShape<DIM, P, NF, float, false>::kSupported &&
           dgk::VmShape<DIM, P, NF, double, true>::kSupported && dgk::VmShape<DIM, P, NF, double, false>::kSupported;
}

// A fragments [MT][KS][64 lanes] of the stacked volume operator [Dw_1; ..; Dw_DIM]: row
// 16*mt + lane%16 = r*NP + i, k = 4*ks + lane/16 = j -> Dw_r[i][j].
template <int DIM, int P, int NF>
void buildVmFrag(dg_context* c, const dg::RefElement& R) {
    using V = dgk::VmShape<DIM, P, NF, float, true>;
    const int Np = R.Np;
    c->afragHost.assign(V::kFrag, 0.0);
    for (int mt = 0; mt < V::MT; ++mt)
        for (int ks = 0; ks < V::KS; ++ks)
            for (int lane = 0; lane < 64; ++lane) {
                const int row = 16 * mt + lane % 16, j = 4 * ks + lane / 16;
                if (row >= V::MR || j >= Np) continue;
                const int r = row / Np, i = row % Np;
                c->afragHost[(static_cast<size_t>(mt) * V::KS + ks) * 64 + lane] =
                    R.Dw[(static_cast<size_t>(r) * Np + i) * Np + j];
            }
}

dgk::SourceArgs sourceArgs(dg_context* c) {
    dgk::SourceArgs s{};
    s.nodes = c->dSrcNodes;
    s.lastSrc = c->dSrcLast;
    s.coverOff = c->dSrcCoverOff;
    s.coverSrc = c->dSrcCover;
    s.nnodes = c->nSrcNodes;
    s.params = c->dSrcParams;
    s.tTable = c->dT;
    s.vTable = c->dV;
    s.vFlag = c->dVF;
    s.nsrc = static_cast<int>(c->srcNodes.size());
    s.nf = c->nf;
    s.ic2 = 1.0 / (c->phys.c0 * c->phys.c0);
    return s;
}

// Source overwrite of the first step of a batch; later steps are overwritten by the
// previous step's post kernel.
template <int DIM, int P, typename T, bool COMP>
void launchStageWave(dg_context* c, int stage, const void* qin, void* qout, void* uout, bool snap) {
    const dgk::StageArgs<T> a = stageArgs<T, COMP>(c, stage, qin, qout, uout, snap);
    hipLaunchKernelGGL((dgk::stageKernelWave<DIM, P, T, COMP>), dim3(static_cast<unsigned>(c->nblocks)),
                       dim3(dgk::kBlock), 0, c->stream, a, static_cast<const T*>(c->afrag));
}

template <int DIM, int P, typename T, bool COMP>
void launchStageMM(dg_context* c, int stage, const void* qin, void* qout, void* uout, bool snap) {
    using dgk::stageKernelMM;
    const dgk::StageArgs<T> a = stageArgs<T, COMP>(c, stage, qin, qout, uout, snap);
    const T* af = static_cast<const T*>(c->afrag);
    const dim3 grid(static_cast<unsigned>(c->nblocks)), block(dgk::kBlock);
    switch (dgk::stageMode(stage)) {
        case dgk::kModeFirst:
            hipLaunchKernelGGL((stageKernelMM<DIM, P, T, COMP, dgk::kModeFirst>), grid, block, 0, c->stream, a, af);
            break;
        case dgk::kModeMid:
            hipLaunchKernelGGL((stageKernelMM<DIM, P, T, COMP, dgk::kModeMid>), grid, block, 0, c->stream, a, af);
            break;
        case dgk::kModeFinal:
            hipLaunchKernelGGL((stageKernelMM<DIM, P, T, COMP, dgk::kModeFinal>), grid, block, 0, c->stream, a, af);
            break;
        case dgk::kModeEuler:
            hipLaunchKernelGGL((stageKernelMM<DIM, P, T, COMP, dgk::kModeEuler>), grid, block, 0, c->stream, a, af);
            break;
        default:
            hipLaunchKernelGGL((stageKernelMM<DIM, P, T, COMP, dgk::kModeRHS>), grid, block, 0, c->stream, a, af);
            break;
    }
}

// Split-BF16 matrix-core stage (instances in dg_stage_bf16.hip).
void launchStageBf16(dg_context* c, int stage, const void* qin, void* qout, void* uout, bool snap) {
    const dgk::StageArgs<float> a = stageArgs<float, true>(c, stage, qin, qout, uout, snap);
    if (!dgk::launchStageBf16(c->dim, c->p, c->nf, c->bfSplit, dgk::stageMode(stage), a, c->afragBf, c->bfGroups,
                              static_cast<unsigned>(c->nblocks), c->stream))
        throw std::runtime_error("split-BF16 stage kernel not built for this shape");
}

// Nearest BF16 (bit pattern) of a double, through the nearest float.
uint16_t toBf16(double x) {
    const float f = static_cast<float>(x);
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}

double fromBf16(uint16_t b) {
    const uint32_t u = static_cast<uint32_t>(b) << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// Operator fragments of the BF16 matrix-core kernel, [MT][KS][NSPLIT][64 lanes][8]: lane l of
// (row tile it, k-step ks) holds row 16 it + l % 16 and the 8 columns of chunk 4 ks + l / 16 of
// [Dw_1 | .. | Dw_DIM | -LIFT] in the kernel's chunk order (volume chunks (r, c): Dw_r columns
// 8c .. 8c+7; lift chunks c: -LIFT columns 8c .. 8c+7; zero outside). Each FP64 value is split
// into NSPLIT BF16 terms, each the nearest BF16 of what the previous terms leave.
template <int DIM, int P, int NSPLIT>
void buildAfragBf16(dg_context* c, const dg::RefElement& R) {
    using B = dgk::Bf16Shape<DIM, P, 4>;  // the fragment layout does not depend on the fields
    const int Np = R.Np, NFT = R.Nfaces * R.Nfp;
    c->afragBfHost.assign(static_cast<size_t>(B::MT) * B::KS * NSPLIT * 64 * 8, 0);
    for (int it = 0; it < B::MT; ++it)
        for (int ks = 0; ks < B::KS; ++ks)
            for (int lane = 0; lane < 64; ++lane) {
                const int row = 16 * it + lane % 16, g = lane / 16;
                for (int jj = 0; jj < 8; ++jj) {
                    double v = 0.0;
                    if (row < Np) {
                        if (ks < B::KSV) {
                            const int ch = 4 * ks + g;
                            if (ch < DIM * B::CPR) {
                                const int r = ch / B::CPR, j = 8 * (ch % B::CPR) + jj;
                                if (j < Np) v = R.Dw[(static_cast<size_t>(r) * Np + row) * Np + j];
                            }
                        } else {
                            const int fa = 8 * (4 * (ks - B::KSV) + g) + jj;
                            if (fa < NFT) v = -R.LIFT[static_cast<size_t>(row) * NFT + fa];
                        }
                    }
                    double rem = v;
                    for (int s = 0; s < NSPLIT; ++s) {
                        const uint16_t t = toBf16(rem);
                        rem -= fromBf16(t);
                        c->afragBfHost[((((static_cast<size_t>(it) * B::KS + ks) * NSPLIT + s) * 64 + lane) * 8) + jj] = t;
                    }
                }
            }
}

template <int DIM, int P>
void buildAfrag(dg_context* c, const dg::RefElement& R);

// The same fragments lane-contiguous ([lane][mt][ks], padded to a multiple of 4 per lane) for
// the block matrix-core kernel.
template <int DIM, int P>
void buildAfragLane(dg_context* c, const dg::RefElement& R) {
    using M = dgk::MfmaShape<DIM, P>;
    using B = dgk::BlockMmShape<DIM, P>;
    buildAfrag<DIM, P>(c, R);
    // [row tile][lane][k-step (padded to KSP)]
    std::vector<double> lanes(static_cast<size_t>(M::MT) * 64 * B::KSP, 0.0);
    for (int mt = 0; mt < M::MT; ++mt)
        for (int lane = 0; lane < 64; ++lane)
            for (int ks = 0; ks < M::KS; ++ks)
                lanes[(static_cast<size_t>(mt) * 64 + lane) * B::KSP + ks] = c->afragHost[(mt * M::KS + ks) * 64 + lane];
    c->afragHost.swap(lanes);
}

// A-operand fragments [MT][KSV+KSL][64 lanes] of the matrix-core kernel: row 16*mt + lane%16,
// k = 4*ks + lane/16 of [Dw_1 | .. | Dw_dim] (volume k-steps) then of -LIFT (lift k-steps).
template <int DIM, int P>
void buildAfrag(dg_context* c, const dg::RefElement& R) {
    using M = dgk::MfmaShape<DIM, P>;
    const int Np = R.Np, NFT = R.Nfaces * R.Nfp;
    c->afragHost.assign(M::kFragTotal, 0.0);
    for (int mt = 0; mt < M::MT; ++mt)
        for (int ks = 0; ks < M::KS; ++ks)
            for (int lane = 0; lane < 64; ++lane) {
                const int row = 16 * mt + lane % 16;
                const int k = 4 * ks + lane / 16;
                double v = 0.0;
                if (row < Np) {
                    if (ks < M::KSV) {
                        if (k < M::KV) v = R.Dw[(static_cast<size_t>(k / Np) * Np + row) * Np + k % Np];
                    } else {
                        const int fa = 4 * (ks - M::KSV) + lane / 16;
                        if (fa < NFT) v = -R.LIFT[static_cast<size_t>(row) * NFT + fa];
                    }
                }
                c->afragHost[(mt * M::KS + ks) * 64 + lane] = v;
            }
}

template <typename T, bool COMP>
void launchSource(dg_context* c) {
    if (c->srcNodes.empty()) return;
    hipLaunchKernelGGL((dgk::sourceKernel<T, COMP>), dim3(1), dim3(256), 0, c->stream, sourceArgs(c), c->counter,
                       static_cast<T*>(c->u), static_cast<T*>(COMP ? c->ulo : nullptr));
}

template <typename T, bool COMP>
void launchSourceStep(dg_context* c, int64_t offset) {
    const unsigned nb = static_cast<unsigned>((c->nSrcNodes + dgk::kSrcU * 256 - 1) / (dgk::kSrcU * 256));
    hipLaunchKernelGGL((dgk::sourceStepKernel<T, COMP>), dim3(nb), dim3(256), 0, c->stream, sourceArgs(c), c->srcBase,
                       offset, c->batchSteps, static_cast<T*>(c->u), static_cast<T*>(COMP ? c->ulo : nullptr));
}

template <typename T, bool COMP>
void launchPost(dg_context* c) {
    hipStream_t st = c->postStream ? c->postStream : c->stream;
    dgk::PostArgs p{};
    p.resPartial = c->resPartial;
    p.nblocks = c->nblocks;
    if (c->resGroup && c->curResid) {
        hipLaunchKernelGGL(dgk::residualGroupKernel, dim3(static_cast<unsigned>(c->nResGroups)), dim3(256), 0, st,
                           static_cast<const double*>(c->resPartial), c->nblocks, c->resGroup);
        p.resPartial = c->resGroup;
        p.nblocks = c->nResGroups;
    }
    p.invN = 1.0 / static_cast<double>(c->K * c->Np);
    p.obsIdx = c->dObsIdx;
    p.obsW = c->dObsW;
    p.obsSeg0 = c->dObsSeg0;
    p.segPartial = c->dSegPartial;
    p.nseg = c->nseg;
    p.nobs = static_cast<int>(c->obsNodes.size());
    p.nf = c->nf;
    p.resRing = c->resRing;
    p.obsRing = c->obsRing;
    p.nonfinite = c->nonfinite;
    p.stepCounter = c->counter;
    p.batchSteps = c->batchSteps;
    p.computeResiduals = c->curResid;
    p.withSources = c->postSources ? 1 : 0;
    hipLaunchKernelGGL((dgk::postKernel<T, COMP>), dim3(1), dim3(dgk::kPostBlock), 0, st, p, sourceArgs(c),
                       static_cast<T*>(c->u), static_cast<T*>(COMP ? c->ulo : nullptr));
}

template <int DIM, int P, int NF, typename T>
void launchClosure(dg_context* c, const void* qin) {
    dgc::ClosureArgs<T> a{};
    a.qin = static_cast<const T*>(qin);
    a.ghost = static_cast<T*>(c->ghost);
    a.fgeo = static_cast<const T*>(c->fgeo);
    a.iops = c->iops;
    a.faces = c->dLearned;
    a.w = c->dClosureW;
    a.nfaces = c->nLearned;
    a.ph = makePhysics<T>(c->phys);
    if (c->nLearned <= dgc::kClosureSmall) {
        constexpr int kPerBlock = dgc::kClosureWaves * 4;
        const int64_t nb = (c->nLearned + kPerBlock - 1) / kPerBlock;
        hipLaunchKernelGGL((dgc::closureKernel<DIM, P, NF, T, 4>), dim3(static_cast<unsigned>(nb)),
                           dim3(dgc::kClosureBlock), 0, c->stream, a);
    } else {
        constexpr int kPerBlock = dgc::kClosureWaves * 8;
        const int64_t nb = (c->nLearned + kPerBlock - 1) / kPerBlock;
        hipLaunchKernelGGL((dgc::closureKernel<DIM, P, NF, T, 8>), dim3(static_cast<unsigned>(nb)),
                           dim3(dgc::kClosureBlock), 0, c->stream, a);
    }
}

template <int DIM, int P, int NF, typename T>
void launchTraces(dg_context* c, int64_t n, const int64_t* elem, const int32_t* face, double* out, double* normals) {
    const int64_t threads = n * dgk::Sizes<DIM, P>::NFP;
    const int64_t nb = (threads + 255) / 256;
    hipLaunchKernelGGL((dgc::traceKernel<DIM, P, NF, T>), dim3(static_cast<unsigned>(nb)), dim3(256), 0, c->stream,
                       static_cast<const T*>(c->u), static_cast<const T*>(c->ulo), static_cast<const T*>(c->geo),
                       static_cast<const T*>(c->ops), c->iops, c->nbr, c->code, elem, face, n, out, normals);
}

template <int DIM, int P, int NF>
void pickAux(dg_context* c) {
    if (c->precision == DG_FP64) {
        c->traceFn = &launchTraces<DIM, P, NF, double>;
        if constexpr (DIM == 2 || P <= 3) c->closureFn = &launchClosure<DIM, P, NF, double>;
    } else {
        c->traceFn = &launchTraces<DIM, P, NF, float>;
        if constexpr (DIM == 2 || P <= 3) c->closureFn = &launchClosure<DIM, P, NF, float>;
    }
}

// Stage kernel without DGFEM_KERNEL: the split-BF16 matrix-core kernel for the mixed precision
// where it measured faster than the VALU kernel (tetrahedra of order 5: 5.07 against 7.77 ms per
// step on the 511,104-tet 5-field case, at lower error against FP64); on tetrahedra of order 3
// the VALU kernel with its volume contraction on the matrix cores (0.5-4% faster per step than
// the VALU kernel in both precisions on the 196,608-tet case, on three boxes); the VALU kernel
// elsewhere (the split-BF16 kernel at order 3: 0.62 against 0.40 ms per step, its contractions
// too small a part of a stage for the split operands to pay).
std::string defaultKernel(int dim, int p, int nf, int precision) {
    if (precision == DG_MIXED && dim == 3 && p >= kBf16DefaultMinOrder && dgk::stageBf16Built(dim, p, nf, 3))
        return "bf16x6";
    if (dim == 3 && p == 3 && (nf == 4 || nf == 5)) return "vm";
    return "valu";
}

template <int DIM, int P, int NF, int NSPLIT>
void pickBf16(dg_context* c) {
    c->mfma = false;
    c->kernelKind = NSPLIT == 3 ? DG_KERNEL_BF16X6 : DG_KERNEL_BF16X3;
    buildAfragBf16<DIM, P, NSPLIT>(c, dg::refElement(DIM, P));
    c->eb = dgk::Bf16Shape<DIM, P, NF>::kGroup;
    c->bfGroups = (c->K + c->eb - 1) / c->eb;
    c->bfBlocksPerCu = dgk::stageBf16BlocksPerCu(DIM, P, NF, NSPLIT);
    c->bfSplit = NSPLIT;
    c->stageFn = &launchStageBf16;
    c->sourceFn = &launchSource<float, true>;
    c->postFn = &launchPost<float, true>;
}

// Stage-kernel variant: DGFEM_KERNEL = valu (default) | mm (block matrix-core, 4 fields)
// | wave (wave-autonomous matrix-core, 4 fields) | bf16x6, bf16x3 (split-BF16 matrix-core,
// mixed precision; 3 and 2 BF16 terms per operand). An explicit request for a variant that is
// not built for the mesh's (dim, order, fields, precision) fails dg_create with
// DG_ERR_UNSUPPORTED.
template <int DIM, int P, int NF>
bool pick(dg_context* c) {
    pickAux<DIM, P, NF>(c);
    const char* env = std::getenv("DGFEM_KERNEL");
    const std::string kind = env && env[0] ? env : defaultKernel(DIM, P, NF, c->precision);
    if (kind != "valu" && kind != "mm" && kind != "wave" && kind != "bf16x6" && kind != "bf16x3" && kind != "vm")
        throw UnsupportedError("DGFEM_KERNEL=" + kind +
                               " is not a stage-kernel variant (valu, vm, mm, wave, bf16x6, bf16x3)");
    if (kind == "vm") {
        if constexpr (vmBuilt<DIM, P, NF>()) {
            c->mfma = true;
            c->kernelKind = DG_KERNEL_VM;
            buildVmFrag<DIM, P, NF>(c, dg::refElement(DIM, P));
            if (c->precision == DG_FP64) {
                pickStage<DIM, P, NF, double, false, true>(c);
                c->sourceFn = &launchSource<double, false>;
                c->postFn = &launchPost<double, false>;
            } else {
                pickStage<DIM, P, NF, float, true, true>(c);
                c->sourceFn = &launchSource<float, true>;
                c->postFn = &launchPost<float, true>;
            }
            return true;
        }
        throw UnsupportedError("DGFEM_KERNEL=vm is built for tetrahedra of order 2 and 3 only");
    }
    if (kind == "bf16x6" || kind == "bf16x3") {
        if (c->precision != DG_MIXED)
            throw UnsupportedError("DGFEM_KERNEL=" + kind + " is a mixed-precision kernel (FP32 state)");
        const int nsplit = kind == "bf16x6" ? 3 : 2;
        if (!dgk::stageBf16Built(DIM, P, NF, nsplit))
            throw UnsupportedError("DGFEM_KERNEL=" + kind + " is built for tetrahedra of order 2..5 only");
        if constexpr (DIM == 3 && P >= 2) {
            if (nsplit == 3)
                pickBf16<DIM, P, NF, 3>(c);
            else
                pickBf16<DIM, P, NF, 2>(c);
            return true;
        }
        throw UnsupportedError("DGFEM_KERNEL=" + kind + " is built for tetrahedra of order 2..5 only");
    }
    if (kind == "mm" && !(NF == 4 && dgk::BlockMmShape<DIM, P>::kSupported))
        throw UnsupportedError("DGFEM_KERNEL=mm is built for 4 fields on tetrahedra of order <= 3 and triangles "
                               "of order <= 5 only");
    if (kind == "wave" && !(NF == 4 && dgk::MfmaShape<DIM, P>::kSupported))
        throw UnsupportedError("DGFEM_KERNEL=wave is built for 4 fields on tetrahedra of order <= 3 and "
                               "triangles of order <= 5 only");
    if constexpr (NF == 4 && dgk::BlockMmShape<DIM, P>::kSupported) {
        if (kind == "mm") {
            c->mfma = true;
            c->kernelKind = DG_KERNEL_MM;
            buildAfragLane<DIM, P>(c, dg::refElement(DIM, P));
            c->eb = dgk::BlockMmShape<DIM, P>::EB;
            if (c->precision == DG_FP64) {
                c->stageFn = &launchStageMM<DIM, P, double, false>;
                c->sourceFn = &launchSource<double, false>;
                c->postFn = &launchPost<double, false>;
            } else {
                c->stageFn = &launchStageMM<DIM, P, float, true>;
                c->sourceFn = &launchSource<float, true>;
                c->postFn = &launchPost<float, true>;
            }
            return true;
        }
    }
    if constexpr (NF == 4 && dgk::MfmaShape<DIM, P>::kSupported) {
        if (kind == "wave") {
            c->mfma = true;
            c->kernelKind = DG_KERNEL_WAVE;
            buildAfrag<DIM, P>(c, dg::refElement(DIM, P));
            c->eb = dgk::WaveShape<DIM, P>::EB;
            if (c->precision == DG_FP64) {
                c->stageFn = &launchStageWave<DIM, P, double, false>;
                c->sourceFn = &launchSource<double, false>;
                c->postFn = &launchPost<double, false>;
            } else {
                c->stageFn = &launchStageWave<DIM, P, float, true>;
                c->sourceFn = &launchSource<float, true>;
                c->postFn = &launchPost<float, true>;
            }
            return true;
        }
    }
    c->mfma = false;
    c->kernelKind = DG_KERNEL_VALU;
    if (c->precision == DG_FP64) {
        pickStage<DIM, P, NF, double, false>(c);
        c->sourceFn = &launchSource<double, false>;
        c->postFn = &launchPost<double, false>;
    } else {
        pickStage<DIM, P, NF, float, true>(c);
        c->sourceFn = &launchSource<float, true>;
        c->postFn = &launchPost<float, true>;
    }
    return true;
}

template <int DIM, int NF>
bool pickP(dg_context* c) {
    switch (c->p) {
        case 1: return pick<DIM, 1, NF>(c);
        case 2: return pick<DIM, 2, NF>(c);
        case 3: return pick<DIM, 3, NF>(c);
        case 4: return pick<DIM, 4, NF>(c);
        case 5: return pick<DIM, 5, NF>(c);
        default: return false;
    }
}

// One stage: exterior states of the learned faces from the stage input, then the stage.
void runStage(dg_context* c, int stage, const void* qin, void* qout, void* uout, bool snap) {
    if (c->nLearned > 0) c->closureFn(c, qin);
    c->stageFn(c, stage, qin, qout, uout, snap);
}

bool pickKernels(dg_context* c) {
    if (c->dim == 2) return c->nf == 4 ? pickP<2, 4>(c) : pickP<2, 5>(c);
    return c->nf == 4 ? pickP<3, 4>(c) : pickP<3, 5>(c);
}

template <int DIM, int P>
int opTotal() {
    return dgk::OpLayout<DIM, P>::kTotal;
}

// Operator tables in the kernel's layout (see dg_kernels.hpp OpLayout).
void buildOps(const dg::RefElement& R, std::vector<double>& ops, std::vector<int32_t>& iops) {
    const int Np = R.Np, dim = R.dim, NFT = R.Nfaces * R.Nfp;
    ops.assign(static_cast<size_t>(Np) * Np * dim + static_cast<size_t>(NFT) * Np + R.Nfaces * 3, 0.0);
    size_t o = 0;
    for (int j = 0; j < Np; ++j)
        for (int i = 0; i < Np; ++i)
            for (int r = 0; r < dim; ++r) ops[o++] = R.Dw[(static_cast<size_t>(r) * Np + i) * Np + j];
    for (int fa = 0; fa < NFT; ++fa)
        for (int i = 0; i < Np; ++i) ops[o++] = R.LIFT[static_cast<size_t>(i) * NFT + fa];
    for (int f = 0; f < R.Nfaces; ++f)
        for (int k = 0; k < 3; ++k) ops[o++] = R.Nref[f * 3 + k];
    iops.clear();
    for (int v : R.fmask) iops.push_back(v);
    for (int v : R.facePerm) iops.push_back(v);
}

template <typename T>
void uploadAs(dg_context* c, void* dst, const std::vector<double>& src) {
    std::vector<T> tmp(src.size());
    for (size_t i = 0; i < src.size(); ++i) tmp[i] = static_cast<T>(src[i]);
    HIPCHK(hipMemcpy(dst, tmp.data(), tmp.size() * sizeof(T), hipMemcpyHostToDevice));
}

void releaseGraphs(dg_context* c) {
    if (c->gMulti) (void)hipGraphExecDestroy(c->gMulti);
    if (c->gOne) (void)hipGraphExecDestroy(c->gOne);
    c->gMulti = c->gOne = nullptr;
    c->gSteps = 0;
    c->gIntegrator = -1;
}

void freeTables(dg_context* c) {
    for (void* p : {static_cast<void*>(c->dSrcNodes), static_cast<void*>(c->dSrcLast), static_cast<void*>(c->dSrcCoverOff),
                    static_cast<void*>(c->dSrcCover), static_cast<void*>(c->dSrcParams), static_cast<void*>(c->dObsIdx),
                    static_cast<void*>(c->dObsW), static_cast<void*>(c->dObsSeg0), static_cast<void*>(c->dSegPartial)})
        devFree(p);
    c->dSrcNodes = c->dSrcCoverOff = c->dObsIdx = nullptr;
    c->dSrcLast = c->dSrcCover = c->dObsSeg0 = nullptr;
    c->dSrcParams = c->dObsW = c->dSegPartial = nullptr;
    c->nSrcNodes = 0;
}

template <typename V>
V* uploadVec(dg_context* c, const std::vector<V>& v) {
    V* d = devAlloc<V>(c, v.size());
    if (!v.empty()) HIPCHK(hipMemcpy(d, v.data(), v.size() * sizeof(V), hipMemcpyHostToDevice));
    return d;
}

void syncTables(dg_context* c) {
    if (!c->tablesDirty) return;
    releaseGraphs(c);
    freeTables(c);
    // source table: the distinct nodes covered by a source (ascending), each with the sources
    // covering it in definition order and the last of them
    std::vector<std::pair<int64_t, int32_t>> cover;
    std::vector<double> prm;
    for (size_t s = 0; s < c->srcNodes.size(); ++s) {
        for (int64_t n : c->srcNodes[s]) cover.emplace_back(n, static_cast<int32_t>(s));
        prm.insert(prm.end(), c->srcParams[s].begin(), c->srcParams[s].end());
    }
    std::sort(cover.begin(), cover.end());
    cover.erase(std::unique(cover.begin(), cover.end()), cover.end());
    std::vector<int64_t> srcNodes, coverOff{0};
    std::vector<int32_t> lastSrc, coverSrc;
    for (size_t k = 0; k < cover.size(); ++k) {
        if (k > 0 && cover[k].first != cover[k - 1].first) coverOff.push_back(static_cast<int64_t>(coverSrc.size()));
        if (k == 0 || cover[k].first != cover[k - 1].first) srcNodes.push_back(cover[k].first);
        coverSrc.push_back(cover[k].second);
    }
    if (!cover.empty()) coverOff.push_back(static_cast<int64_t>(coverSrc.size()));
    for (size_t i = 0; i < srcNodes.size(); ++i) lastSrc.push_back(coverSrc[coverOff[i + 1] - 1]);
    c->nSrcNodes = static_cast<int64_t>(srcNodes.size());
    // entry 0 of every table is valid memory for the clamped loads of the kernels
    if (srcNodes.empty()) {
        srcNodes.push_back(0);
        lastSrc.push_back(0);
        coverSrc.push_back(0);
    }
    while (coverOff.size() < 2) coverOff.push_back(0);
    prm.resize(std::max<size_t>(prm.size(), 5), 0.0);
    c->dSrcNodes = uploadVec(c, srcNodes);
    c->dSrcLast = uploadVec(c, lastSrc);
    c->dSrcCoverOff = uploadVec(c, coverOff);
    c->dSrcCover = uploadVec(c, coverSrc);
    c->dSrcParams = uploadVec(c, prm);
    // observer node lists in segments of 256 entries, each padded with its first node at weight 0
    // (one all-padding segment of node 0 when no observer has a node)
    std::vector<int64_t> oidx;
    std::vector<double> ow;
    std::vector<int32_t> obsSeg0{0};
    int nseg = 0;
    for (size_t o = 0; o < c->obsNodes.size(); ++o) {
        const std::vector<int64_t>& nd = c->obsNodes[o];
        const std::vector<double>& w = c->obsW[o];
        for (size_t b = 0; b < nd.size(); b += 256, ++nseg)
            for (size_t q = 0; q < 256; ++q) {
                const bool in = b + q < nd.size();
                oidx.push_back(nd[in ? b + q : b]);
                ow.push_back(in ? w[b + q] : 0.0);
            }
        obsSeg0.push_back(nseg);
    }
    c->nseg = nseg;
    if (oidx.empty()) {
        oidx.assign(256, 0);
        ow.assign(256, 0.0);
    }
    while (obsSeg0.size() < 2) obsSeg0.push_back(nseg);
    c->dObsIdx = uploadVec(c, oidx);
    c->dObsW = uploadVec(c, ow);
    c->dObsSeg0 = uploadVec(c, obsSeg0);
    c->dSegPartial = devAlloc<double>(c, std::max(1, c->nseg) * 5);
    // rings and per-step tables
    const int64_t ns = std::max<int64_t>(1, static_cast<int64_t>(c->srcNodes.size()));
    const int64_t no = std::max<int64_t>(1, static_cast<int64_t>(c->obsNodes.size()));
    devFree(c->dT);
    devFree(c->dV);
    devFree(c->dVF);
    devFree(c->obsRing);
    if (c->hV) (void)hipHostFree(c->hV);
    if (c->hVF) (void)hipHostFree(c->hVF);
    if (c->hObs) (void)hipHostFree(c->hObs);
    c->dT = devAlloc<double>(c, c->ringSlots);
    c->dV = devAlloc<double>(c, c->ringSlots * ns);
    c->dVF = devAlloc<uint8_t>(c, c->ringSlots * ns);
    c->obsRing = devAlloc<double>(c, c->ringSlots * no * 4);
    HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&c->hV), sizeof(double) * c->ringSlots * ns));
    HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&c->hVF), c->ringSlots * ns));
    HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&c->hObs), sizeof(double) * c->ringSlots * no * 4));
    std::memset(c->hV, 0, sizeof(double) * c->ringSlots * ns);
    std::memset(c->hVF, 0, c->ringSlots * ns);
    c->tablesDirty = false;
}

// Launch the kernels of one step on the context stream.
void launchStep(dg_context* c, int integrator, bool resid) {
    if (integrator == DG_RK4) {
        runStage(c, dgk::kStageRK1, c->u, c->qa, c->u, false);
        runStage(c, dgk::kStageRK2, c->qa, c->qb, c->u, false);
        runStage(c, dgk::kStageRK3, c->qb, c->qa, c->u, false);
        runStage(c, dgk::kStageRK4, c->qa, nullptr, c->u, resid);
    } else {
        runStage(c, dgk::kStageEuler, c->u, nullptr, c->qb, resid);
        HIPCHK(hipMemcpyAsync(c->u, c->qb, static_cast<size_t>(c->K) * c->Np * c->nf * c->tsize,
                              hipMemcpyDeviceToDevice, c->stream));
    }
    c->postFn(c);
}

// Step `s` of a captured RK4 sequence with the end-of-step kernel (residuals, observers, step
// counter) on a branch of its own. The next step's first three stages neither write u/ulo nor
// touch the residual partials, so only its final stage waits for the branch; the next step's
// source overwrite runs on the main branch (sourceStepKernel). Used when no observer node is a
// source node, so that the observers' reads and the overwrite touch different nodes.
void launchStepForked(dg_context* c, bool resid, int s) {
    runStage(c, dgk::kStageRK1, c->u, c->qa, c->u, false);
    runStage(c, dgk::kStageRK2, c->qa, c->qb, c->u, false);
    runStage(c, dgk::kStageRK3, c->qb, c->qa, c->u, false);
    if (s > 0) HIPCHK(hipStreamWaitEvent(c->stream, c->evJoin, 0));
    runStage(c, dgk::kStageRK4, c->qa, nullptr, c->u, resid);
    HIPCHK(hipEventRecord(c->evFork, c->stream));
    HIPCHK(hipStreamWaitEvent(c->side, c->evFork, 0));
    c->postStream = c->side;
    c->postSources = false;
    c->postFn(c);
    c->postStream = nullptr;
    c->postSources = true;
    HIPCHK(hipEventRecord(c->evJoin, c->side));
    if (c->nSrcNodes > 0) {
        if (c->precision == DG_FP64)
            launchSourceStep<double, false>(c, s + 1);
        else
            launchSourceStep<float, true>(c, s + 1);
    }
}

// The branch adds a cross-stream dependency each way per step. On MI355X a 7,865-block 3D grid
// with a source and two observers breaks even with it (the end-of-step kernel's ~19 µs hidden,
// about as much added), and a 1,536-block 2D grid ran 39% slower per step. It is kept for large
// grids, where the end-of-step kernel grows with the residual partials and the observer and
// source tables while the branch's cost does not. DGFEM_POST_BRANCH = 0 | 1 overrides the rule.
constexpr int64_t kBranchMinBlocks = 4096;

// Whether some observer node is also a source node, from the host definitions (the same test
// syncTables makes when it builds the device tables).
bool observersOverlapSources(const dg_context* c) {
    std::vector<int64_t> src;
    for (const std::vector<int64_t>& nd : c->srcNodes) src.insert(src.end(), nd.begin(), nd.end());
    std::sort(src.begin(), src.end());
    for (const std::vector<int64_t>& nd : c->obsNodes)
        for (int64_t n : nd)
            if (std::binary_search(src.begin(), src.end(), n)) return true;
    return false;
}

bool usePostBranch(const dg_context* c, int integrator) {
    if (integrator != DG_RK4 || observersOverlapSources(c)) return false;
    const char* env = std::getenv("DGFEM_POST_BRANCH");
    if (env && env[0] == '0') return false;
    if (env && env[0] == '1') return true;
    return c->nblocks >= kBranchMinBlocks;
}

void captureGraphs(dg_context* c, int integrator, bool resid, int steps) {
    releaseGraphs(c);
    const bool fork = usePostBranch(c, integrator);
    auto capture = [&](int n) {
        hipGraph_t g;
        HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        if (fork) {
            for (int s = 0; s < n; ++s) launchStepForked(c, resid, s);
            HIPCHK(hipStreamWaitEvent(c->stream, c->evJoin, 0));
            if (c->nSrcNodes > 0)
                hipLaunchKernelGGL(dgk::advanceKernel, dim3(1), dim3(1), 0, c->stream, c->srcBase, static_cast<int64_t>(n));
        } else {
            for (int s = 0; s < n; ++s) launchStep(c, integrator, resid);
        }
        HIPCHK(hipStreamEndCapture(c->stream, &g));
        hipGraphExec_t ex;
        HIPCHK(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
        (void)hipGraphDestroy(g);
        return ex;
    };
    c->gOne = capture(1);
    c->gMulti = steps > 1 ? capture(steps) : nullptr;
    c->gSteps = steps;
    c->gIntegrator = integrator;
    c->gResid = resid ? 1 : 0;
}

// Caller layout [nf][K*Np] (file node order) <-> device layout [K][Np canonical][nf].
void toDeviceLayout(const dg_context* c, const double* u, std::vector<double>& out) {
    const int64_t N = c->K * c->Np;
    out.assign(static_cast<size_t>(N) * c->nf, 0.0);
    for (int64_t e = 0; e < c->K; ++e)
        for (int a = 0; a < c->Np; ++a) {
            const int64_t cn = e * c->Np + c->mesh.toCanon[e * c->Np + a];
            for (int f = 0; f < c->nf; ++f) out[cn * c->nf + f] = u[f * N + e * c->Np + a];
        }
}

void fromDeviceLayout(const dg_context* c, const std::vector<double>& dev, double* u) {
    const int64_t N = c->K * c->Np;
    for (int64_t e = 0; e < c->K; ++e)
        for (int a = 0; a < c->Np; ++a) {
            const int64_t cn = e * c->Np + c->mesh.toCanon[e * c->Np + a];
            for (int f = 0; f < c->nf; ++f) u[f * N + e * c->Np + a] = dev[cn * c->nf + f];
        }
}

void uploadState(dg_context* c, const std::vector<double>& dev, void* hiBuf, void* loBuf) {
    if (c->precision == DG_FP64) {
        HIPCHK(hipMemcpy(hiBuf, dev.data(), dev.size() * 8, hipMemcpyHostToDevice));
    } else {
        std::vector<float> hi(dev.size()), lo(dev.size());
        for (size_t i = 0; i < dev.size(); ++i) {
            hi[i] = static_cast<float>(dev[i]);
            lo[i] = static_cast<float>(dev[i] - static_cast<double>(hi[i]));
        }
        HIPCHK(hipMemcpy(hiBuf, hi.data(), hi.size() * 4, hipMemcpyHostToDevice));
        if (loBuf) HIPCHK(hipMemcpy(loBuf, lo.data(), lo.size() * 4, hipMemcpyHostToDevice));
    }
}

void downloadState(dg_context* c, const void* hiBuf, const void* loBuf, std::vector<double>& dev) {
    const size_t n = static_cast<size_t>(c->K) * c->Np * c->nf;
    dev.resize(n);
    if (c->precision == DG_FP64) {
        HIPCHK(hipMemcpy(dev.data(), hiBuf, n * 8, hipMemcpyDeviceToHost));
    } else {
        std::vector<float> hi(n), lo(n, 0.0f);
        HIPCHK(hipMemcpy(hi.data(), hiBuf, n * 4, hipMemcpyDeviceToHost));
        if (loBuf) HIPCHK(hipMemcpy(lo.data(), loBuf, n * 4, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; ++i) dev[i] = static_cast<double>(hi[i]) + static_cast<double>(lo[i]);
    }
}

dg_status fail(dg_context* c, dg_status code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

void destroyContext(dg_context* c) {
    if (!c) return;
    if (c->device >= 0) (void)hipSetDevice(c->device);
    releaseGraphs(c);
    freeTables(c);
    for (void* p : {c->u, c->ulo, c->qa, c->qb, c->acc, c->snap, c->geo, c->fgeo, c->ops, c->afrag, c->afragBf})
        devFree(p);
    devFree(c->nbr);
    devFree(c->iops);
    devFree(c->code);
    devFree(c->resPartial);
    devFree(c->resGroup);
    devFree(c->dT);
    devFree(c->dV);
    devFree(c->dVF);
    devFree(c->resRing);
    devFree(c->obsRing);
    devFree(c->nonfinite);
    devFree(c->counter);
    devFree(c->batchSteps);
    devFree(c->srcBase);
    devFree(c->dLearned);
    devFree(c->ghost);
    devFree(c->dClosureW);
    for (void* h : {static_cast<void*>(c->hT), static_cast<void*>(c->hV), static_cast<void*>(c->hRes),
                    static_cast<void*>(c->hObs), static_cast<void*>(c->hVF)})
        if (h) (void)hipHostFree(h);
    if (c->evFork) (void)hipEventDestroy(c->evFork);
    if (c->evJoin) (void)hipEventDestroy(c->evJoin);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

}  // namespace

extern "C" {

dg_status dg_create(const dg_mesh_desc* m, const dg_physics* ph, const dg_options* opt, dg_context** out) {
    if (!out) return DG_ERR_INVALID;
    *out = nullptr;
    if (!m || !ph || !m->node_coords || !m->element_vertices) {
        g_createError = "dg_create: null argument";
        return DG_ERR_INVALID;
    }
    if (m->dim != 2 && m->dim != 3) {
        g_createError = "dg_create: dim must be 2 or 3";
        return DG_ERR_UNSUPPORTED;
    }
    if (m->order < 1 || m->order > 5) {
        g_createError = "dg_create: order must be in 1..5";
        return DG_ERR_UNSUPPORTED;
    }
    if (m->nodes_per_element != dg::numNodes(m->dim, m->order)) {
        g_createError = "dg_create: nodes_per_element does not match the order";
        return DG_ERR_INVALID;
    }
    if (!(ph->c0 > 0) || !(ph->rho0 > 0)) {
        g_createError = "dg_create: c0 and rho0 must be positive";
        return DG_ERR_INVALID;
    }
    dg_context* c = new dg_context();
    c->device = -1;
    try {
        c->opt = opt ? *opt : dg_options{DG_FP64, 1, 16, 0};
        if (c->opt.steps_per_graph < 1) c->opt.steps_per_graph = 1;
        c->precision = c->opt.precision == DG_MIXED ? DG_MIXED : DG_FP64;
        c->tsize = c->precision == DG_FP64 ? 8 : 4;
        c->phys = *ph;
        c->nf = ph->fields == DG_FIELDS_LEE5 ? 5 : 4;
        c->dim = m->dim;
        c->p = m->order;
        c->K = m->num_elements;
        c->mesh = dg::buildDGMesh(m->dim, m->order, m->num_elements, m->node_coords, m->element_vertices, m->face_bc);
        const dg::RefElement& R = dg::refElement(m->dim, m->order);
        c->Np = R.Np;
        c->Nfp = R.Nfp;
        c->Nfc = R.Nfaces;
        if (c->K > INT32_MAX / 2) throw std::runtime_error("too many elements for 32-bit neighbour ids");
        if (!c->mesh.learnedFaces.empty() && m->dim == 3 && m->order > 3) {
            delete c;
            g_createError = "dg_create: the learned closure (DG_BC_LEARNED) is built for 2D meshes and tetrahedra "
                            "of order <= 3";
            return DG_ERR_UNSUPPORTED;
        }
        bool picked = false;
        try {
            picked = pickKernels(c);
        } catch (const UnsupportedError& ex) {
            delete c;
            g_createError = std::string("dg_create: ") + ex.what();
            return DG_ERR_UNSUPPORTED;
        }
        if (!picked) {
            delete c;
            g_createError = "dg_create: no kernel built for this (dim, order)";
            return DG_ERR_UNSUPPORTED;
        }
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
            destroyContext(c);
            g_createError = "dg_create: no HIP device available";
            return DG_ERR_DEVICE;
        }
        c->device = c->opt.device;
        HIPCHK(hipSetDevice(c->device));
        HIPCHK(scratchAddress(c->device, &c->scratch));
        HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        HIPCHK(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&c->evFork, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&c->evJoin, hipEventDisableTiming));
        c->nblocks = (c->K + c->eb - 1) / c->eb;  // one batch of eb elements per block
        if (c->bfGroups > 0) {
            // persistent blocks of the BF16 matrix-core kernel: as many as are resident at once
            hipDeviceProp_t prop;
            HIPCHK(hipGetDeviceProperties(&prop, c->device));
            c->nblocks = std::min<int64_t>(c->bfGroups, static_cast<int64_t>(prop.multiProcessorCount) * c->bfBlocksPerCu);
        }
        const size_t nstate = static_cast<size_t>(c->K) * c->Np * c->nf;
        auto alloc = [&](size_t n) -> void* {
            return c->precision == DG_FP64 ? static_cast<void*>(devAlloc<double>(c, n))
                                           : static_cast<void*>(devAlloc<float>(c, n));
        };
        c->u = alloc(nstate);
        c->qa = alloc(nstate);
        c->qb = alloc(nstate);
        c->acc = alloc(nstate);
        // the residual snapshot holds the state's leading word only (mixed precision: the FP32
        // value nearest the compensated state, so a residual term is off by <= 2^-24 |snapshot|)
        c->snap = alloc(nstate);
        if (c->precision == DG_MIXED) c->ulo = alloc(nstate);
        HIPCHK(hipMemset(c->u, 0, nstate * c->tsize));
        if (c->ulo) HIPCHK(hipMemset(c->ulo, 0, nstate * c->tsize));
        // geometry: [K][dim*dim] dxi_r/dx_c
        std::vector<double> geo(static_cast<size_t>(c->K) * c->dim * c->dim);
        for (int64_t e = 0; e < c->K; ++e)
            for (int r = 0; r < c->dim; ++r)
                for (int q = 0; q < c->dim; ++q) geo[(e * c->dim + r) * c->dim + q] = c->mesh.Jinv[e * 9 + r * 3 + q];
        c->geo = alloc(geo.size());
        // per (element, face): outward unit normal and Fscale in FP64, rounded once
        std::vector<double> fgeo(static_cast<size_t>(c->K) * c->Nfc * 4);
        for (int64_t e = 0; e < c->K; ++e)
            for (int f = 0; f < c->Nfc; ++f) {
                double nv[3], fs;
                dg::faceGeometry(c->mesh, e, f, nv, fs);
                double* o = &fgeo[(static_cast<size_t>(e) * c->Nfc + f) * 4];
                o[0] = nv[0];
                o[1] = nv[1];
                o[2] = nv[2];
                o[3] = fs;
            }
        c->fgeo = alloc(fgeo.size());
        std::vector<double> ops;
        std::vector<int32_t> iops;
        buildOps(R, ops, iops);
        c->ops = alloc(ops.size());
        if (c->mfma) c->afrag = alloc(c->afragHost.size());
        if (!c->afragBfHost.empty()) c->afragBf = uploadVec(c, c->afragBfHost);
        if (c->precision == DG_FP64) {
            uploadAs<double>(c, c->geo, geo);
            uploadAs<double>(c, c->fgeo, fgeo);
            uploadAs<double>(c, c->ops, ops);
            if (c->mfma) uploadAs<double>(c, c->afrag, c->afragHost);
        } else {
            uploadAs<float>(c, c->geo, geo);
            uploadAs<float>(c, c->fgeo, fgeo);
            uploadAs<float>(c, c->ops, ops);
            if (c->mfma) uploadAs<float>(c, c->afrag, c->afragHost);
        }
        c->iops = uploadVec(c, iops);
        c->nbr = uploadVec(c, c->mesh.nbr);
        c->code = uploadVec(c, c->mesh.nbrCode);
        c->nLearned = static_cast<int64_t>(c->mesh.learnedFaces.size());
        if (c->nLearned > 0) {
            c->dLearned = uploadVec(c, c->mesh.learnedFaces);
            c->ghost = alloc(static_cast<size_t>(c->nLearned) * c->Nfp * c->nf);
            c->closureParams = dgc::numParams(c->dim * c->Nfp + 2, c->Nfp);
            c->dClosureW = devAlloc<uint16_t>(c, c->closureParams);
            HIPCHK(hipMemset(c->dClosureW, 0, c->closureParams * sizeof(uint16_t)));
        }
        c->resPartial = devAlloc<double>(c, static_cast<size_t>(c->nblocks) * 5);
        HIPCHK(hipMemset(c->resPartial, 0, static_cast<size_t>(c->nblocks) * 5 * sizeof(double)));
        if (c->nblocks > 4 * dgk::kResGroup) {  // below that the extra launch costs what it saves
            c->nResGroups = (c->nblocks + dgk::kResGroup - 1) / dgk::kResGroup;
            c->resGroup = devAlloc<double>(c, static_cast<size_t>(c->nResGroups) * 5);
        }
        c->ringSlots = 4096;
        c->resRing = devAlloc<double>(c, c->ringSlots * 5);
        c->nonfinite = devAlloc<int>(c, 2);  // first non-finite slot, state flag of the final stages
        c->counter = devAlloc<int64_t>(c, 1);
        c->batchSteps = devAlloc<int64_t>(c, 1);
        c->srcBase = devAlloc<int64_t>(c, 1);
        HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&c->hT), sizeof(double) * c->ringSlots));
        HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&c->hRes), sizeof(double) * c->ringSlots * 5));
        c->tablesDirty = true;
        syncTables(c);
        HIPCHK(hipDeviceSynchronize());
    } catch (const std::exception& ex) {
        g_createError = std::string("dg_create: ") + ex.what();
        bool meshErr = c->stream == nullptr;
        destroyContext(c);
        return meshErr ? DG_ERR_MESH : DG_ERR_DEVICE;
    }
    *out = c;
    return DG_OK;
}

void dg_destroy(dg_context* c) { destroyContext(c); }

dg_status dg_set_closure(dg_context* c, const uint16_t* params, int64_t n) {
    if (!c || !params) return DG_ERR_INVALID;
    if (c->nLearned == 0) return fail(c, DG_ERR_INVALID, "dg_set_closure: the mesh has no DG_BC_LEARNED faces");
    if (n != c->closureParams)
        return fail(c, DG_ERR_INVALID,
                    "dg_set_closure: expected " + std::to_string(c->closureParams) + " parameters");
    try {
        HIPCHK(hipSetDevice(c->device));
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipMemcpy(c->dClosureW, params, n * sizeof(uint16_t), hipMemcpyHostToDevice));
    } catch (const std::exception& ex) {
        return fail(c, DG_ERR_DEVICE, ex.what());
    }
    return DG_OK;
}

dg_status dg_learned_faces(const dg_context* c, int64_t* elem, int32_t* face, int64_t capacity, int64_t* num) {
    if (!c || !num) return DG_ERR_INVALID;
    *num = c->nLearned;
    for (int64_t i = 0; i < c->nLearned && i < capacity; ++i) {
        if (elem) elem[i] = c->mesh.learnedFaces[i] / c->Nfc;
        if (face) face[i] = static_cast<int32_t>(c->mesh.learnedFaces[i] % c->Nfc);
    }
    return DG_OK;
}

dg_status dg_face_traces(dg_context* c, int64_t n, const int64_t* elem, const int32_t* face, double* out,
                         double* normals) {
    if (!c || n < 0 || (n > 0 && (!elem || !face || !out || !normals))) return DG_ERR_INVALID;
    for (int64_t i = 0; i < n; ++i)
        if (elem[i] < 0 || elem[i] >= c->K || face[i] < 0 || face[i] >= c->Nfc)
            return fail(c, DG_ERR_INVALID, "dg_face_traces: element or face out of range");
    if (n == 0) return DG_OK;
    void *dE = nullptr, *dF = nullptr, *dO = nullptr, *dN = nullptr;
    const size_t outN = static_cast<size_t>(n) * 2 * c->Nfp * c->nf;
    dg_status st = DG_OK;
    try {
        HIPCHK(hipSetDevice(c->device));
        HIPCHK(hipMalloc(&dE, n * sizeof(int64_t)));
        HIPCHK(hipMalloc(&dF, n * sizeof(int32_t)));
        HIPCHK(hipMalloc(&dO, outN * sizeof(double)));
        HIPCHK(hipMalloc(&dN, n * 3 * sizeof(double)));
        HIPCHK(hipMemcpyAsync(dE, elem, n * sizeof(int64_t), hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(dF, face, n * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
        c->traceFn(c, n, static_cast<const int64_t*>(dE), static_cast<const int32_t*>(dF), static_cast<double*>(dO),
                   static_cast<double*>(dN));
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(out, dO, outN * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(normals, dN, n * 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    } catch (const std::exception& ex) {
        st = fail(c, DG_ERR_DEVICE, ex.what());
    }
    for (void* p : {dE, dF, dO, dN}) devFree(p);
    return st;
}

dg_status dg_face_traces_noalloc(dg_context* c, int64_t n, const int64_t* elem, const int32_t* face, double* out,
                                 double* normals) {
    if (!c || n < 0 || (n > 0 && (!elem || !face || !out || !normals))) return DG_ERR_INVALID;
    if (traceScratchBytes(n, c->Nfp, c->nf) > kScratchBytes)
        return fail(c, DG_ERR_INVALID,
                    "dg_face_traces_noalloc: " + std::to_string(n) + " faces need " +
                        std::to_string(traceScratchBytes(n, c->Nfp, c->nf)) + " bytes of staging, more than the " +
                        std::to_string(kScratchBytes) + " of the per-device array");
    for (int64_t i = 0; i < n; ++i)
        if (elem[i] < 0 || elem[i] >= c->K || face[i] < 0 || face[i] >= c->Nfc)
            return fail(c, DG_ERR_INVALID, "dg_face_traces_noalloc: element or face out of range");
    if (n == 0) return DG_OK;
    const size_t m = static_cast<size_t>(n);
    unsigned char* base = static_cast<unsigned char*>(c->scratch);
    int64_t* dE = reinterpret_cast<int64_t*>(base);
    int32_t* dF = reinterpret_cast<int32_t*>(base + align256(8 * m));
    double* dO = reinterpret_cast<double*>(base + align256(8 * m) + align256(4 * m));
    double* dN = reinterpret_cast<double*>(base + align256(8 * m) + align256(4 * m) + align256(16 * m * c->Nfp * c->nf));
    const size_t outN = m * 2 * c->Nfp * c->nf;
    try {
        HIPCHK(hipSetDevice(c->device));
        HIPCHK(hipMemcpyAsync(dE, elem, m * sizeof(int64_t), hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(dF, face, m * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
        c->traceFn(c, n, dE, dF, dO, dN);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(out, dO, outN * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(normals, dN, m * 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    } catch (const std::exception& ex) {
        return fail(c, DG_ERR_DEVICE, ex.what());
    }
    return DG_OK;
}

const char* dg_last_error(const dg_context* c) { return c ? c->err.c_str() : g_createError.c_str(); }

dg_status dg_set_state(dg_context* c, const double* u) {
    if (!c || !u) return DG_ERR_INVALID;
    try {
        HIPCHK(hipSetDevice(c->device));
        std::vector<double> dev;
        toDeviceLayout(c, u, dev);
        uploadState(c, dev, c->u, c->ulo);
    } catch (const std::exception& ex) {
        return fail(c, DG_ERR_DEVICE, ex.what());
    }
    return DG_OK;
}

dg_status dg_get_state(dg_context* c, double* u) {
    if (!c || !u) return DG_ERR_INVALID;
    try {
        HIPCHK(hipSetDevice(c->device));
        HIPCHK(hipStreamSynchronize(c->stream));
        std::vector<double> dev;
        downloadState(c, c->u, c->ulo, dev);
        fromDeviceLayout(c, dev, u);
    } catch (const std::exception& ex) {
        return fail(c, DG_ERR_DEVICE, ex.what());
    }
    return DG_OK;
}

dg_status dg_add_gaussian(dg_context* c, double x, double y, double z, double size, double amp) {
    if (!c || !(size > 0)) return DG_ERR_INVALID;
    try {
        HIPCHK(hipSetDevice(c->device));
        HIPCHK(hipStreamSynchronize(c->stream));
        std::vector<double> dev;
        downloadState(c, c->u, c->ulo, dev);
        const int64_t N = c->K * c->Np;
        for (int64_t n = 0; n < N; ++n) {
            const double* X = &c->mesh.nodes[n * 3];
            double v = amp * std::exp(-((X[0] - x) * (X[0] - x) + (X[1] - y) * (X[1] - y) + (X[2] - z) * (X[2] - z)) /
                                      size);
            dev[n * c->nf] += v;
            if (c->nf == 5) dev[n * c->nf + 4] += v / (c->phys.c0 * c->phys.c0);
        }
        uploadState(c, dev, c->u, c->ulo);
    } catch (const std::exception& ex) {
        return fail(c, DG_ERR_DEVICE, ex.what());
    }
    return DG_OK;
}

static dg_status addSource(dg_context* c, double x, double y, double z, double size, const std::array<double, 5>& prm,
                           TableSource ts, int64_t* num) {
    if (!c) return DG_ERR_INVALID;
    std::vector<int64_t> nodes;
    const int64_t N = c->K * c->Np;
    for (int64_t n = 0; n < N; ++n) {
        const double* X = &c->mesh.nodes[n * 3];
        if (std::pow(X[0] - x, 2) + std::pow(X[1] - y, 2) + std::pow(X[2] - z, 2) < std::pow(size, 2))
            nodes.push_back(n);
    }
    c->srcNodes.push_back(nodes);
    c->srcParams.push_back(prm);
    c->srcTables.push_back(ts);
    c->tablesDirty = true;
    if (num) *num = static_cast<int64_t>(nodes.size());
    return DG_OK;
}

dg_status dg_add_source_monopole(dg_context* c, double x, double y, double z, double size, double amp, double freq,
                                 double phase, double duration, int64_t* num) {
    return addSource(c, x, y, z, size, {0.0, amp, freq, phase, duration}, TableSource{nullptr, nullptr}, num);
}

dg_status dg_add_source_table(dg_context* c, double x, double y, double z, double size, dg_value_fn fn, void* user,
                              int64_t* num) {
    if (!fn) return fail(c, DG_ERR_INVALID, "dg_add_source_table: null value function");
    return addSource(c, x, y, z, size, {1.0, 0, 0, 0, 0}, TableSource{fn, user}, num);
}

dg_status dg_add_observer(dg_context* c, double x, double y, double z, double size, int64_t* num) {
    if (!c) return DG_ERR_INVALID;
    std::vector<int64_t> nodes;
    std::vector<double> w;
    const int64_t N = c->K * c->Np;
    for (int64_t n = 0; n < N; ++n) {
        const double* X = &c->mesh.nodes[n * 3];
        double d = std::sqrt(std::pow(X[0] - x, 2) + std::pow(X[1] - y, 2) + std::pow(X[2] - z, 2));
        if (d < size) {
            nodes.push_back(n);
            w.push_back(1.0 / (std::pow(d, 2) + 1.0e-12));
        }
    }
    c->obsNodes.push_back(nodes);
    c->obsW.push_back(w);
    c->tablesDirty = true;
    if (num) *num = static_cast<int64_t>(nodes.size());
    return DG_OK;
}

dg_status dg_run(dg_context* c, const dg_run_params* prm, dg_step_fn on_step, dg_save_fn on_save, void* user,
                 int64_t* steps_done) {
    if (!c || !prm) return DG_ERR_INVALID;
    if (steps_done) *steps_done = 0;
    if (!(prm->dt > 0)) return fail(c, DG_ERR_INVALID, "dg_run: dt must be positive");
    if (prm->integrator != DG_RK4 && prm->integrator != DG_EULER)
        return fail(c, DG_ERR_INVALID, "dg_run: unknown integrator");
    try {
        HIPCHK(hipSetDevice(c->device));
        syncTables(c);
        // The reference loop (src/solver.cpp:413-414): double accumulation of t and tDisplay.
        std::vector<double> times;
        std::vector<uint8_t> isSave;
        {
            double t = prm->t_start, step = 0, tDisplay = 0;
            for (; t <= prm->t_end; t += prm->dt, tDisplay += prm->dt, ++step) {
                if (prm->max_steps >= 0 && static_cast<int64_t>(times.size()) >= prm->max_steps) break;
                bool save = false;
                if (tDisplay >= prm->t_rate || step == 0) {
                    tDisplay = 0;
                    save = true;
                }
                times.push_back(t);
                isSave.push_back(save);
            }
        }
        const bool resid = prm->compute_residuals != 0;
        c->curResid = resid ? 1 : 0;
        c->curDt = prm->dt;
        const bool useGraph = c->opt.use_graph != 0;
        if (useGraph && (c->gOne == nullptr || c->gIntegrator != prm->integrator || c->gResid != (resid ? 1 : 0) ||
                         c->gSteps != c->opt.steps_per_graph || c->gDt != prm->dt))
            captureGraphs(c, prm->integrator, resid, c->opt.steps_per_graph), c->gDt = prm->dt;
        const int64_t nsteps = static_cast<int64_t>(times.size());
        const int ns = static_cast<int>(c->srcNodes.size());
        const int no = static_cast<int>(c->obsNodes.size());
        const size_t stateN = static_cast<size_t>(c->K) * c->Np * c->nf;
        hipEvent_t ev0, ev1;
        HIPCHK(hipEventCreate(&ev0));
        HIPCHK(hipEventCreate(&ev1));
        double devMs = 0.0;
        std::vector<double> stateHost;
        int64_t done = 0;
        while (done < nsteps) {
            int64_t end = done + 1;
            while (end < nsteps && !isSave[end] && end - done < c->ringSlots) ++end;
            const int64_t n = end - done;
            if (isSave[done]) {
                HIPCHK(hipMemcpyAsync(c->snap, c->u, stateN * c->tsize, hipMemcpyDeviceToDevice, c->stream));
                if (on_save) {
                    HIPCHK(hipStreamSynchronize(c->stream));
                    std::vector<double> dev;
                    downloadState(c, c->u, c->ulo, dev);
                    stateHost.resize(stateN);
                    fromDeviceLayout(c, dev, stateHost.data());
                    if (on_save(user, done, times[done], stateHost.data()) != 0) {
                        HIPCHK(hipEventDestroy(ev0));
                        HIPCHK(hipEventDestroy(ev1));
                        return fail(c, DG_ERR_CALLBACK, "dg_run: save callback failed");
                    }
                }
            }
            for (int64_t s = 0; s < n; ++s) {
                c->hT[s] = times[done + s];
                for (int q = 0; q < ns; ++q) {
                    double v = 0.0;
                    int on = 0;
                    if (c->srcTables[q].fn) on = c->srcTables[q].fn(c->srcTables[q].user, times[done + s], &v) != 0;
                    c->hV[s * ns + q] = v;
                    c->hVF[s * ns + q] = static_cast<uint8_t>(on);
                }
            }
            HIPCHK(hipMemcpyAsync(c->dT, c->hT, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
            if (ns > 0) {
                HIPCHK(hipMemcpyAsync(c->dV, c->hV, sizeof(double) * n * ns, hipMemcpyHostToDevice, c->stream));
                HIPCHK(hipMemcpyAsync(c->dVF, c->hVF, n * ns, hipMemcpyHostToDevice, c->stream));
            }
            hipLaunchKernelGGL(dgk::setCounterKernel, dim3(1), dim3(1), 0, c->stream, c->counter, c->batchSteps,
                               n, c->nonfinite, c->srcBase);
            HIPCHK(hipEventRecord(ev0, c->stream));
            c->sourceFn(c);
            if (useGraph) {
                int64_t s = 0;
                if (c->gMulti)
                    for (; s + c->gSteps <= n; s += c->gSteps) HIPCHK(hipGraphLaunch(c->gMulti, c->stream));
                for (; s < n; ++s) HIPCHK(hipGraphLaunch(c->gOne, c->stream));
            } else {
                for (int64_t s = 0; s < n; ++s) launchStep(c, prm->integrator, resid);
            }
            HIPCHK(hipGetLastError());
            HIPCHK(hipEventRecord(ev1, c->stream));
            HIPCHK(hipMemcpyAsync(c->hRes, c->resRing, sizeof(double) * n * 5, hipMemcpyDeviceToHost, c->stream));
            if (no > 0)
                HIPCHK(hipMemcpyAsync(c->hObs, c->obsRing, sizeof(double) * n * no * 4, hipMemcpyDeviceToHost,
                                      c->stream));
            int flag = -1;
            HIPCHK(hipMemcpyAsync(&flag, c->nonfinite, sizeof(int), hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
            devMs += ms;
            int64_t lim = n;
            if (flag >= 0) lim = flag + 1;
            for (int64_t s = 0; s < lim; ++s) {
                if (on_step) {
                    dg_step_record rec{done + s, times[done + s], &c->hRes[s * 5], no, no ? &c->hObs[s * no * 4] : nullptr,
                                       1e-3 * ms / static_cast<double>(n)};
                    if (on_step(user, &rec) != 0) {
                        (void)hipEventDestroy(ev0);
                        (void)hipEventDestroy(ev1);
                        if (steps_done) *steps_done = done + s + 1;
                        return fail(c, DG_ERR_CALLBACK, "dg_run: step callback failed");
                    }
                }
            }
            if (flag >= 0) {
                (void)hipEventDestroy(ev0);
                (void)hipEventDestroy(ev1);
                if (steps_done) *steps_done = done + flag + 1;
                c->lastRunMs = devMs;
                return fail(c, DG_ERR_NONFINITE,
                            "dg_run: non-finite residual or state at step " + std::to_string(done + flag));
            }
            done = end;
        }
        (void)hipEventDestroy(ev0);
        (void)hipEventDestroy(ev1);
        c->lastRunMs = devMs;
        if (steps_done) *steps_done = done;
    } catch (const std::exception& ex) {
        return fail(c, DG_ERR_DEVICE, ex.what());
    }
    return DG_OK;
}

struct CollectCtx {
    double *times, *res, *obs;
    int64_t cap;
};

static int collectStep(void* user, const dg_step_record* r) {
    auto* cc = static_cast<CollectCtx*>(user);
    if (r->step >= cc->cap) return 1;
    if (cc->times) cc->times[r->step] = r->t;
    if (cc->res) std::memcpy(cc->res + r->step * 5, r->residuals, 5 * sizeof(double));
    if (cc->obs && r->num_observers > 0)
        std::memcpy(cc->obs + r->step * r->num_observers * 4, r->observers, sizeof(double) * 4 * r->num_observers);
    return 0;
}

dg_status dg_run_collect(dg_context* c, const dg_run_params* prm, double* times, double* residuals, double* observers,
                         int64_t capacity, int64_t* steps_done) {
    if (!c || !prm) return DG_ERR_INVALID;
    dg_run_params p2 = *prm;
    if (p2.max_steps < 0 || p2.max_steps > capacity) p2.max_steps = capacity;
    CollectCtx cc{times, residuals, observers, capacity};
    return dg_run(c, &p2, collectStep, nullptr, &cc, steps_done);
}

dg_status dg_rhs(dg_context* c, const double* u, double* k, double dt) {
    if (!c || !u || !k) return DG_ERR_INVALID;
    try {
        HIPCHK(hipSetDevice(c->device));
        std::vector<double> dev;
        toDeviceLayout(c, u, dev);
        uploadState(c, dev, c->qa, nullptr);
        c->curDt = dt;
        runStage(c, dgk::kStageRHS, c->qa, c->qb, c->u, false);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->stream));
        std::vector<double> out;
        downloadState(c, c->qb, nullptr, out);
        fromDeviceLayout(c, out, k);
    } catch (const std::exception& ex) {
        return fail(c, DG_ERR_DEVICE, ex.what());
    }
    return DG_OK;
}

dg_status dg_get_info(const dg_context* c, dg_info* info) {
    if (!c || !info) return DG_ERR_INVALID;
    info->dim = c->dim;
    info->order = c->p;
    info->nodes_per_element = c->Np;
    info->nodes_per_face = c->Nfp;
    info->fields = c->nf;
    info->precision = c->precision;
    info->num_elements = c->K;
    info->num_faces = c->mesh.numFaces;
    info->num_boundary_faces = c->mesh.numBoundaryFaces;
    double r = 1e300;
    for (double v : c->mesh.inradius) r = std::min(r, v);
    info->min_inradius = r;
    info->device_bytes = c->deviceBytes;
    info->stage_kernel = c->kernelKind;
    info->post_branch = (c->opt.use_graph != 0 && usePostBranch(c, DG_RK4)) ? 1 : 0;
    info->elements_per_block = c->eb;
    info->stage_blocks = c->nblocks;
    return DG_OK;
}

dg_status dg_profile_steps(dg_context* c, double dt, int32_t steps, double* stage_ms, double* source_ms,
                           double* post_ms) {
    if (!c || steps < 1) return DG_ERR_INVALID;
    try {
        HIPCHK(hipSetDevice(c->device));
        syncTables(c);
        c->curDt = dt;
        c->curResid = 1;
        std::vector<double> tt(std::min<int64_t>(steps, c->ringSlots), 0.0);
        for (size_t s = 0; s < tt.size(); ++s) c->hT[s] = s * dt;
        HIPCHK(hipMemcpyAsync(c->dT, c->hT, sizeof(double) * tt.size(), hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL(dgk::setCounterKernel, dim3(1), dim3(1), 0, c->stream, c->counter, c->batchSteps,
                           static_cast<int64_t>(tt.size()), c->nonfinite, c->srcBase);
        std::vector<hipEvent_t> ev(14);
        for (auto& e : ev) HIPCHK(hipEventCreate(&e));
        double acc[6] = {0, 0, 0, 0, 0, 0};
        int n = static_cast<int>(tt.size());
        for (int s = 0; s < n; ++s) {
            HIPCHK(hipEventRecord(ev[0], c->stream));
            c->sourceFn(c);
            HIPCHK(hipEventRecord(ev[1], c->stream));
            runStage(c, dgk::kStageRK1, c->u, c->qa, c->u, false);
            HIPCHK(hipEventRecord(ev[2], c->stream));
            runStage(c, dgk::kStageRK2, c->qa, c->qb, c->u, false);
            HIPCHK(hipEventRecord(ev[3], c->stream));
            runStage(c, dgk::kStageRK3, c->qb, c->qa, c->u, false);
            HIPCHK(hipEventRecord(ev[4], c->stream));
            runStage(c, dgk::kStageRK4, c->qa, nullptr, c->u, true);
            HIPCHK(hipEventRecord(ev[5], c->stream));
            c->postFn(c);
            HIPCHK(hipEventRecord(ev[6], c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            float ms;
            HIPCHK(hipEventElapsedTime(&ms, ev[0], ev[1]));
            acc[4] += ms;
            for (int k = 0; k < 4; ++k) {
                HIPCHK(hipEventElapsedTime(&ms, ev[1 + k], ev[2 + k]));
                acc[k] += ms;
            }
            HIPCHK(hipEventElapsedTime(&ms, ev[5], ev[6]));
            acc[5] += ms;
        }
        for (auto& e : ev) (void)hipEventDestroy(e);
        HIPCHK(hipGetLastError());
        if (stage_ms)
            for (int k = 0; k < 4; ++k) stage_ms[k] = acc[k] / n;
        if (source_ms) *source_ms = acc[4] / n;
        if (post_ms) *post_ms = acc[5] / n;
    } catch (const std::exception& ex) {
        return fail(c, DG_ERR_DEVICE, ex.what());
    }
    return DG_OK;
}

double dg_last_run_device_ms(const dg_context* c) { return c ? c->lastRunMs : 0.0; }

}  // extern "C"
