// HIP kernels (gfx950) for the DG semi-discrete right-hand side and the RK4 / forward-Euler
// update of the reference solver (src/solver.cpp:38-57, :311-584; src/Mesh.cpp:734-975),
// in nodal form:
//   rhs = sum_x A_x sum_r (dxi_r/dx) Dw_r q  -  sum_f Fscale_f LIFT_f pi_f
//   k   = dt * rhs                      (numStep with beta = 0, src/solver.cpp:50-54)
// One fused kernel per RK stage: state of a block of elements staged in LDS, numerical
// flux at every face node (neighbour traces gathered from HBM), weak-derivative volume
// contraction, lift, and the stage's register update.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace dgk {

constexpr int kBlock = 256;

template <int DIM, int P>
struct Sizes {
    static constexpr int NP = DIM == 2 ? (P + 1) * (P + 2) / 2 : (P + 1) * (P + 2) * (P + 3) / 6;
    static constexpr int NFP = DIM == 2 ? P + 1 : (P + 1) * (P + 2) / 2;
    static constexpr int NFC = DIM + 1;
    static constexpr int NFT = NFC * NFP;
    static constexpr int NPERM = DIM == 2 ? 2 : 6;
    static constexpr int EB = kBlock / NP;  // elements per block
};

// Operator table layout (element type T), per (DIM, P):
//   DwT  [NP(j)][NP(i)][DIM(r)]  weak derivative Dw_r[i][j]
//   LT   [NFT(fa)][NP(i)]        LIFT[i][fa]
//   Nref [NFC][3]                reference outward normals / face parameter measure
// Integer table: fmask [NFC][NFP], fperm [NPERM][NFP].
template <int DIM, int P>
struct OpLayout {
    using S = Sizes<DIM, P>;
    static constexpr int kDw = 0;
    static constexpr int kL = kDw + S::NP * S::NP * DIM;
    static constexpr int kN = kL + S::NFT * S::NP;
    static constexpr int kTotal = kN + S::NFC * 3;
    static constexpr int kMask = 0;
    static constexpr int kPerm = S::NFT;
    static constexpr int kITotal = kPerm + S::NPERM * S::NFP;
};

template <typename T>
struct Physics {
    T v0[3];
    T rho0, c0;
    T rc2;    // rho0 c0^2
    T irho0;  // 1 / rho0
    T irc0;   // 1 / (rho0 c0)
    T ic2;    // 1 / c0^2
};

// Stage modes.
enum : int { kStageRK1 = 0, kStageRK2 = 1, kStageRK3 = 2, kStageRK4 = 3, kStageEuler = 4, kStageRHS = 5 };

template <typename T>
struct StageArgs {
    const T* __restrict__ qin;   // stage input [K][NP][NF]
    T* __restrict__ qout;        // next stage input (RK1..3) or k (RHS mode)
    T* __restrict__ u;           // base state (hi part when compensated)
    T* __restrict__ uout;        // where the final stage writes the new state
    T* __restrict__ ulo;         // compensation (mixed precision), else null
    T* __restrict__ acc;         // RK accumulator
    const T* __restrict__ geo;   // [K][DIM*DIM] dxi_r/dx_c
    const T* __restrict__ fgeo;  // [K][NFC][4] outward unit normal, Fscale = |J^-T Nref| (= sJ/|J|)
    const int32_t* __restrict__ nbr;   // [K][NFC]: >= 0 neighbour, -1 absorbing, -2 reflecting,
                                       // -3-i learned-closure face i (exterior state in ghost)
    const T* __restrict__ ghost;       // [learned faces][NFP][NF] exterior states, or null
    const uint8_t* __restrict__ code;  // [K][NFC]
    const T* __restrict__ ops;
    const int32_t* __restrict__ iops;
    const T* __restrict__ vfrag;      // (matrix-core volume) stacked [Dw_r] A fragments [MT][KS][64]
    const T* __restrict__ snap;       // residual snapshot (leading word of the state), RK4/Euler final stage
    double* __restrict__ resPartial;  // [gridDim][5]
    int* __restrict__ stateBad;       // set to 1 by a final stage that produced a non-finite state value
    int64_t K;
    T dt;
    int stage;
    Physics<T> ph;
};

// Blocks are dispatched round-robin over the 8 XCDs; give each XCD a contiguous range of
// elements so that neighbour traces are likely to hit in that XCD's L2.
__device__ inline int64_t xcdRemap(int64_t b, int64_t nb) {
    int64_t nb8 = nb - nb % 8;
    if (b >= nb8) return b;
    return (b % 8) * (nb8 / 8) + b / 8;
}

// Normal flux F(q).n for the fields (p, vx, vy, vz[, rho]).
template <int NF, typename T>
__device__ inline void normalFlux(const Physics<T>& ph, const T n[3], T vn0, const T* q, T* out) {
    T nv = n[0] * q[1] + n[1] * q[2] + n[2] * q[3];
    out[0] = vn0 * q[0] + ph.rc2 * nv;
    out[1] = vn0 * q[1] + n[0] * q[0] * ph.irho0;
    out[2] = vn0 * q[2] + n[1] * q[0] * ph.irho0;
    out[3] = vn0 * q[3] + n[2] * q[0] * ph.irho0;
    if constexpr (NF == 5) out[4] = vn0 * q[4] + ph.rho0 * nv;
}

// Numerical flux at one face node in the element's outward normal direction.
//  interior:   n.{F} + (c0/2)(q- - q+)    (src/Mesh.cpp:785-797 with sigma = +1)
//  reflecting: n.F(p-, v- - (v-.n)n)      (src/Mesh.cpp:911-949)
//  absorbing:  RKR q-                     (src/Mesh.cpp:619-666, :954-969)
template <int DIM, int NF, typename T>
__device__ inline void numericalFlux(const Physics<T>& ph, const T n[3], int nb, const T* qm, const T* qp,
                                     T* pi) {
    const T vn0 = ph.v0[0] * n[0] + ph.v0[1] * n[1] + ph.v0[2] * n[2];
    if (nb >= 0) {
        // F is linear in q: n.(F(q-) + F(q+))/2 = n.F(q- + q+)/2
        T qs[NF], fs[NF];
#pragma unroll
        for (int k = 0; k < NF; ++k) qs[k] = qm[k] + qp[k];
        normalFlux<NF>(ph, n, vn0, qs, fs);
        const T hc = T(0.5) * ph.c0;
#pragma unroll
        for (int k = 0; k < NF; ++k) pi[k] = T(0.5) * fs[k] + hc * (qm[k] - qp[k]);
    } else if (nb == -2) {
        T q[NF];
        const T vn = n[0] * qm[1] + n[1] * qm[2] + n[2] * qm[3];
        q[0] = qm[0];
        q[1] = qm[1] - vn * n[0];
        q[2] = qm[2] - vn * n[1];
        q[3] = qm[3] - vn * n[2];
        if constexpr (NF == 5) q[4] = qm[4];
        normalFlux<NF>(ph, n, vn0, q, pi);
    } else {
        const T a = ph.c0 + vn0;
        const T lam = (vn0 < T(0)) ? T(0) : T(-1);
        const T nv = n[0] * qm[1] + n[1] * qm[2] + n[2] * qm[3];
        const T q4 = T(0.25);
        pi[0] = q4 * (a * qm[0] + ph.c0 * ph.rho0 * a * nv);
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            // tangential projector: I - nn^T (3D); T T^T with T = (-ny, nx, 0) in 2D
            T tang = (DIM == 3 || r < 2) ? (qm[1 + r] - n[r] * nv) : T(0);
            pi[1 + r] = q4 * (n[r] * a * qm[0] * ph.irc0 + a * n[r] * nv - vn0 * lam * tang);
        }
        if constexpr (NF == 5) pi[4] = q4 * (a * qm[4] + ph.rho0 / ph.c0 * a * nv);
    }
}

template <typename T>
struct Vec4;
template <>
struct Vec4<float> {
    using V = float4;
};
template <>
struct Vec4<double> {
    using V = double4;
};

// One node's field vector: a single 16-byte (FP32) / 2x16-byte (FP64) access for 4 fields.
template <int NF, typename T>
__device__ inline void loadNode(const T* __restrict__ p, T (&v)[NF]) {
    if constexpr (NF == 4) {
        const typename Vec4<T>::V x = *reinterpret_cast<const typename Vec4<T>::V*>(p);
        v[0] = x.x;
        v[1] = x.y;
        v[2] = x.z;
        v[3] = x.w;
    } else {
#pragma unroll
        for (int k = 0; k < NF; ++k) v[k] = p[k];
    }
}

// Streamed once per stage (base state, accumulator, compensation, snapshot): non-temporal, so
// that the caches keep the stage input that neighbour gathers re-read.
template <int NF, typename T>
__device__ inline void loadNodeNT(const T* __restrict__ p, T (&v)[NF]) {
    if constexpr (NF == 4) {
        typedef T V __attribute__((ext_vector_type(4)));
        const V x = __builtin_nontemporal_load(reinterpret_cast<const V*>(p));
        v[0] = x[0];
        v[1] = x[1];
        v[2] = x[2];
        v[3] = x[3];
    } else {
#pragma unroll
        for (int k = 0; k < NF; ++k) v[k] = __builtin_nontemporal_load(p + k);
    }
}

template <int NF, typename T>
__device__ inline void storeNodeNT(T* __restrict__ p, const T (&v)[NF]) {
    if constexpr (NF == 4) {
        typedef T V __attribute__((ext_vector_type(4)));
        const V x = {v[0], v[1], v[2], v[3]};
        __builtin_nontemporal_store(x, reinterpret_cast<V*>(p));
    } else {
#pragma unroll
        for (int k = 0; k < NF; ++k) __builtin_nontemporal_store(v[k], p + k);
    }
}

template <int NF, typename T>
__device__ inline void storeNode(T* __restrict__ p, const T (&v)[NF]) {
    if constexpr (NF == 4) {
        typename Vec4<T>::V x;
        x.x = v[0];
        x.y = v[1];
        x.z = v[2];
        x.w = v[3];
        *reinterpret_cast<typename Vec4<T>::V*>(p) = x;
    } else {
#pragma unroll
        for (int k = 0; k < NF; ++k) p[k] = v[k];
    }
}

// Whether every value of a node's new state is finite (a non-finite state is reported by the
// end-of-step kernel whether or not residuals are computed).
template <int NF, typename T>
__device__ inline bool allFinite(const T (&v)[NF]) {
    bool ok = true;
#pragma unroll
    for (int k = 0; k < NF; ++k) ok = ok && __builtin_isfinite(v[k]);
    return ok;
}

template <typename T>
__device__ inline void twoSum(T a, T b, T& s, T& e) {
    s = a + b;
    T bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

// A double moved between lanes of a row of 16 by a DPP control (VALU; no LDS traffic).
template <int CTRL>
__device__ inline double dppMove(double v) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = static_cast<uint32_t>(
        __builtin_amdgcn_mov_dpp(static_cast<int>(static_cast<uint32_t>(b)), CTRL, 0xF, 0xF, false));
    const uint32_t hi = static_cast<uint32_t>(
        __builtin_amdgcn_mov_dpp(static_cast<int>(static_cast<uint32_t>(b >> 32)), CTRL, 0xF, 0xF, false));
    return __builtin_bit_cast(double, (static_cast<uint64_t>(hi) << 32) | lo);
}

// Sum over each row of 16 lanes, left in every lane of the row: partner lanes are paired
// symmetrically at every level (neighbour, pair of pairs, mirrored halves, mirrored row), so
// all lanes of a row compute the same value in the same order. Every lane of the wave must be
// active.
__device__ inline double rowSum16(double x) {
    x += dppMove<0xB1>(x);   // quad_perm [1,0,3,2]
    x += dppMove<0x4E>(x);   // quad_perm [2,3,0,1]
    x += dppMove<0x141>(x);  // row_half_mirror
    x += dppMove<0x140>(x);  // row_mirror
    return x;
}

// Residual partial sums of a block in a fixed order: rows of 16 threads by rowSum16, then the
// rows in ascending order; thread k < 5 writes out[k]. sRed holds kBlock/16 rows. Called by
// every thread of the block.
template <int kBlock>
__device__ inline void blockResiduals(const double (&res)[5], double (*sRed)[5], double* __restrict__ out) {
    const int tid = threadIdx.x;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        const double v = rowSum16(res[k]);
        if ((tid & 15) == 0) sRed[tid >> 4][k] = v;
    }
    __syncthreads();
    if (tid < 5) {
        double s = 0.0;
        for (int r = 0; r < kBlock / 16; ++r) s += sRed[r][tid];
        out[tid] = s;
    }
}

// Matrix-core operand types: FP32 uses v_mfma_f32_16x16x4f32 (D row = 4*(lane/16)+r), FP64
// v_mfma_f64_16x16x4f64 (D row = lane/16 + 4*r); A lane = row + 16*k, B lane = col + 16*k
// for both (layouts checked on gfx950).
typedef float mf4 __attribute__((ext_vector_type(4)));
typedef double md4 __attribute__((ext_vector_type(4)));

template <typename T>
struct MfmaTraits;
template <>
struct MfmaTraits<float> {
    using V = mf4;
    __device__ static V mma(float a, float b, V c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    __device__ static int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};
template <>
struct MfmaTraits<double> {
    using V = md4;
    __device__ static V mma(double a, double b, V c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    __device__ static int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};

template <int DIM, int P, int NF, typename T, bool WIDE = true>
struct StageWaves {
    static constexpr int NP = Sizes<DIM, P>::NP;
    static constexpr bool kBig = WIDE && NF == 4 && NP >= 10 && NP <= 20;
    // 4 fields, 10..20 nodes: 512-thread blocks (one operator table per 2x the elements) fit
    // 4 blocks = 8 waves per SIMD in LDS in FP32, 2 blocks = 4 waves per SIMD in FP64; the
    // kernel is asked to stay within the registers that allows. WIDE = false keeps 256-thread
    // blocks, for grids too small to give every CU several 512-thread blocks
    static constexpr int kBlockSize = kBig ? 512 : kBlock;
    static constexpr int value = kBig ? (sizeof(T) == 4 ? 8 : 4) : 1;
    static constexpr int EB = kBlockSize / NP;  // elements per block
};

// Volume contraction on the matrix cores ("vm" variant of the stage kernel): the stacked
// operator [Dw_1; ..; Dw_DIM] (DIM*NP rows x NP) times the staged state (NP rows x 16 columns =
// ECT elements x NF fields) per 16x16 tile; each wave owns one row tile (its A fragments in
// registers) and CPW column tiles. Dq goes node-major through LDS (aliasing the flux buffer)
// for the flux-Jacobian fold; fluxes, lift and update stay on the VALU.
template <int DIM, int P, int NF, typename T, bool WIDE>
struct VmShape {
    using W = StageWaves<DIM, P, NF, T, WIDE>;
    static constexpr int NP = Sizes<DIM, P>::NP;
    static constexpr int kBlockSize = W::kBlockSize;
    static constexpr int NW = kBlockSize / 64;
    static constexpr int ECT = 16 / NF;                       // elements per column tile
    static constexpr int EB = (kBlockSize / NP) / ECT * ECT;  // elements per block
    static constexpr int NCT = EB / ECT;
    static constexpr int MR = DIM * NP;                       // stacked rows
    static constexpr int MT = (MR + 15) / 16;                 // row tiles
    static constexpr int KS = (NP + 3) / 4;                   // k-steps
    static constexpr int G = NW / MT;                         // waves per row tile
    static constexpr int CPW = G > 0 ? NCT / G : 0;           // column tiles per wave
    static constexpr int DS = (DIM * NF + 3) / 4 * 4;         // per-node stride of staged Dq
    // per-element stride: NP*DS + 4, so the 4 elements of a column tile start 20 (FP32) or
    // 8 (FP64, in 32-bit words) banks apart when a row of the tile is stored
    static constexpr int ES = NP * DS + 4;
    static constexpr bool kSupported = G > 0 && NW % MT == 0 && NCT % G == 0 && EB > 0 && NF >= 4;
    static constexpr int kFrag = MT * KS * 64;
};

// Padded LDS row stride (in elements) for per-element blocks of n values: == 4 (mod 16).
__host__ __device__ constexpr int ldsStride(int n) { return n + ((4 - n % 16) + 16) % 16; }

// Stage modes of the VALU stage kernel (a template parameter, so that each mode's loads are
// straight-line code the compiler can issue together).
enum : int { kModeFirst = 0, kModeMid = 1, kModeFinal = 2, kModeEuler = 3, kModeRHS = 4 };

__host__ __device__ constexpr int stageMode(int stage) {
    return stage == kStageRK1 ? kModeFirst
         : (stage == kStageRK2 || stage == kStageRK3) ? kModeMid
         : stage == kStageRK4 ? kModeFinal
         : stage == kStageEuler ? kModeEuler : kModeRHS;
}

// Load scheduling: a block's latency is the number of dependent memory round trips, so every
// global load of a phase is issued before the first wait. Loads are unconditional, from clamped
// (always valid) addresses: a load under a branch that merges with a default value makes the
// compiler wait for it at the merge. Phase 1 (operator tables of the first
// batch, the element's own stage input, inverse Jacobians, face rows, neighbour codes, and the
// RK operands consumed at the end) -> LDS -> barrier; phase 2 (the neighbour traces of all of the
// thread's face nodes) -> fluxes -> barrier; then the contractions and the RK update.
template <int DIM, int P, int NF, typename T, bool COMP, int MODE, bool WIDE = true, bool VM = false>
__global__ __launch_bounds__((StageWaves<DIM, P, NF, T, WIDE>::kBlockSize), (StageWaves<DIM, P, NF, T, WIDE>::value)) void stageKernel(StageArgs<T> a) {
    using S = Sizes<DIM, P>;
    using L = OpLayout<DIM, P>;
    using V = VmShape<DIM, P, NF, T, WIDE>;
    using MF = MfmaTraits<T>;
    static_assert(!VM || V::kSupported, "matrix-core volume shape");
    constexpr int kBlock = StageWaves<DIM, P, NF, T, WIDE>::kBlockSize;  // threads per block
    constexpr int NP = S::NP, NFP = S::NFP, NFC = S::NFC, NFT = S::NFT;
    constexpr int EB = VM ? V::EB : StageWaves<DIM, P, NF, T, WIDE>::EB;
    constexpr int JJ = DIM * DIM;
    constexpr int TS = static_cast<int>(sizeof(T));
    // operator values staged in LDS: all of them, or only LIFT when the volume runs on MFMA
    constexpr int kOpsOff = VM ? L::kL : 0;
    constexpr int kOpsN = VM ? NFT * NP : L::kTotal;
    constexpr bool kOpsInLds = (kOpsN * TS) <= 40 * 1024;
    constexpr bool kNeedU = MODE == kModeMid || MODE == kModeFinal;
    constexpr bool kFinal = MODE == kModeFinal || MODE == kModeEuler;
    // per-thread load counts
    constexpr int kOps16 = kOpsInLds ? (kOpsN * TS) / 16 : 0;         // 16-byte chunks
    constexpr int kOpsRem = kOpsInLds ? kOpsN - kOps16 * (16 / TS) : 0;  // trailing elements
    constexpr int kOpsIt = (kOps16 + kBlock - 1) / kBlock;
    constexpr int kIIt = (L::kITotal + kBlock - 1) / kBlock;
    constexpr int kJIt = (EB * JJ + kBlock - 1) / kBlock;
    constexpr int kFIt = (EB * NFT + kBlock - 1) / kBlock;
    static_assert(EB * NP <= kBlock && EB * NFC <= kBlock, "one node and one face row per thread");
    static_assert(kOpsRem < kBlock, "operator table tail");
    using V4 = typename Vec4<T>::V;
    typedef uint32_t U4 __attribute__((ext_vector_type(4)));  // 16 raw bytes (no unions: stays in VGPRs)

    // per-element strides of the staged state and fluxes padded to 4 (mod 16) elements: the
    // 3-4 elements a wave spans then start in different LDS banks
    constexpr int QS = ldsStride(NP * NF), FS = ldsStride(NFT * NF);
    constexpr int kFD = (VM && EB * V::ES > EB * FS) ? EB * V::ES : EB * FS;
    __shared__ __align__(32) T sQ[EB * QS];
    __shared__ __align__(32) T sF[kFD];  // fluxes; with VM first the node-major Dq
    __shared__ T sJ[EB * JJ];
    // per (element, face): unit normal n, Fscale fs, then (NSV = 8) h n and h (v0.n) with
    // h = fs / 2 (the interior flux with its face scale folded in) where those 4 values per face
    // leave the blocks per CU that the LDS allows unchanged
    constexpr auto ldsBytes = [](int nsv) {
        return TS * (EB * QS + kFD + EB * JJ + EB * NFC * nsv + (kOpsInLds ? kOpsN : 1)) + 5 * EB * NFC +
               4 * L::kITotal + 8 * 5 * (kBlock / 16);
    };
    constexpr int NSV = (160 * 1024) / ldsBytes(8) >= (160 * 1024) / ldsBytes(4) ? 8 : 4;
    __shared__ __align__(32) T sN[EB * NFC * NSV];
    __shared__ __align__(16) int32_t sNb[EB * NFC];  // per (element, face): neighbour / boundary code
    __shared__ __align__(16) uint8_t sCd8[EB * NFC]; // per (element, face): neighbour face | perm << 2
    __shared__ __align__(16) T sOps[kOpsInLds ? kOpsN : 1];
    __shared__ int32_t sI[L::kITotal];
    __shared__ double sRed[kBlock / 16][5];

    const int tid = threadIdx.x;
    const T* ops = kOpsInLds ? sOps : a.ops;                       // (not VM) the whole table
    const T* lift = kOpsInLds ? sOps + (L::kL - kOpsOff) : a.ops + L::kL;  // LIFT^T [fa][i]
    // FP64 tables in LDS are held as two 32-bit planes (low words, then high words): a wave's
    // reads of 20 consecutive doubles repeated per element hit the 32 banks twice in three of
    // the four 16-lane passes of a 64-bit read (4.7 conflict cycles per read measured), the
    // 32-bit plane reads none
    constexpr bool kPlanes = kOpsInLds && sizeof(T) == 8;
    uint32_t* const sOpsW = reinterpret_cast<uint32_t*>(sOps);
    auto opAt = [&](int k) -> T {
        if constexpr (kPlanes) {
            typedef uint32_t W2 __attribute__((ext_vector_type(2)));
            return __builtin_bit_cast(double, W2{sOpsW[k], sOpsW[kOpsN + k]});
        } else {
            return ops[k];
        }
    };
    double res[5] = {0, 0, 0, 0, 0};

    // one batch of EB elements per block (one batch per block measured faster than persistent
    // blocks in both precisions); each XCD takes a contiguous range of batches
    const int64_t bi = xcdRemap(blockIdx.x, gridDim.x);  // the grid has one block per batch
    const int64_t e0 = bi * EB;
    const int nE = static_cast<int>((a.K - e0) < EB ? (a.K - e0) : EB);
    const bool active = tid < nE * NP;
    const int64_t myBase = (e0 * NP + tid) * NF;
    const int64_t ldBase = (e0 * NP + (active ? tid : 0)) * NF;  // clamped load address

    // ---- phase 1: issue every load ----
    U4 rOps[kOpsIt > 0 ? kOpsIt : 1];
    T rOpsTail = T(0);
    int32_t rI[kIIt];
    {
#pragma unroll
        for (int c = 0; c < kOpsIt; ++c) {
            const int x = tid + c * kBlock;
            rOps[c] = reinterpret_cast<const U4*>(a.ops + kOpsOff)[x < kOps16 ? x : kOps16 - 1];
        }
        if (kOpsRem > 0) rOpsTail = a.ops[kOpsOff + kOps16 * (16 / TS) + (tid < kOpsRem ? tid : 0)];
#pragma unroll
        for (int c = 0; c < kIIt; ++c) {
            const int x = tid + c * kBlock;
            rI[c] = a.iops[x < L::kITotal ? x : 0];
        }
    }
    // (VM) this wave's row tile of the stacked operator, KS fragments
    T rA[VM ? V::KS : 1];
    if constexpr (VM) {
        const int rt = (tid >> 6) % V::MT;
#pragma unroll
        for (int ks = 0; ks < V::KS; ++ks) rA[ks] = a.vfrag[(rt * V::KS + ks) * 64 + (tid & 63)];
    }
    T rq[NF];
    loadNode<NF>(a.qin + ldBase, rq);
    T rJ[kJIt];
#pragma unroll
    for (int c = 0; c < kJIt; ++c) {
        const int x = tid + c * kBlock;
        rJ[c] = a.geo[e0 * JJ + (x < nE * JJ ? x : 0)];
    }
    const bool faceRow = tid < nE * NFC;
    const V4 rN = reinterpret_cast<const V4*>(a.fgeo + e0 * NFC * 4)[faceRow ? tid : 0];
    const int eRow = tid < nE ? tid : 0;
    const int fRow = faceRow ? tid : 0;
    U4 rNb4;
    uint32_t rCd4 = 0;  // the element's 4 face codes as one word
    int32_t rNb1 = 0;
    uint8_t rCd1 = 0;
    if constexpr (NFC == 4) {
        rNb4 = reinterpret_cast<const U4*>(a.nbr + e0 * NFC)[eRow];
        rCd4 = reinterpret_cast<const uint32_t*>(a.code + e0 * NFC)[eRow];
    } else {
        rNb1 = a.nbr[e0 * NFC + fRow];
        rCd1 = a.code[e0 * NFC + fRow];
    }
    // RK operands of this thread's node (consumed after the contractions). The compensation
    // word enters only the final accumulation: stage inputs are the FP32 u + c k (measured: same
    // error against FP64 after 2000 steps, 4 B/DOF less per stage).
    T pu[NF], pacc[NF], plo[NF], ps[NF];
#pragma unroll
    for (int k = 0; k < NF; ++k) pu[k] = pacc[k] = plo[k] = ps[k] = T(0);
    if constexpr (kNeedU) {
        loadNodeNT<NF>(a.u + ldBase, pu);
        loadNodeNT<NF>(a.acc + ldBase, pacc);
    }
    if constexpr (COMP && kFinal) loadNodeNT<NF>(a.ulo + ldBase, plo);

    // ---- LDS ----
    {
#pragma unroll
        for (int c = 0; c < kOpsIt; ++c) {
            const int x = tid + c * kBlock;
            if (x < kOps16) {
                if constexpr (kPlanes) {
                    sOpsW[2 * x] = rOps[c].x;
                    sOpsW[2 * x + 1] = rOps[c].z;
                    sOpsW[kOpsN + 2 * x] = rOps[c].y;
                    sOpsW[kOpsN + 2 * x + 1] = rOps[c].w;
                } else {
                    reinterpret_cast<U4*>(sOps)[x] = rOps[c];
                }
            }
        }
        if (kOpsRem > 0 && tid < kOpsRem) {
            const int k = kOps16 * (16 / TS) + tid;
            if constexpr (kPlanes) {
                const uint64_t b = __builtin_bit_cast(uint64_t, static_cast<double>(rOpsTail));
                sOpsW[k] = static_cast<uint32_t>(b);
                sOpsW[kOpsN + k] = static_cast<uint32_t>(b >> 32);
            } else {
                sOps[k] = rOpsTail;
            }
        }
#pragma unroll
        for (int c = 0; c < kIIt; ++c) {
            const int x = tid + c * kBlock;
            if (x < L::kITotal) sI[x] = rI[c];
        }
    }
    if (active) {
        const int el = tid / NP;
        storeNode<NF>(&sQ[el * QS + (tid - el * NP) * NF], rq);
    }
#pragma unroll
    for (int c = 0; c < kJIt; ++c) {
        const int x = tid + c * kBlock;
        if (x < nE * JJ) sJ[x] = rJ[c];
    }
    if (faceRow) {
        if constexpr (NSV == 8) {
            const T h = T(0.5) * rN.w;
            V4 hn;
            hn.x = h * rN.x;
            hn.y = h * rN.y;
            hn.z = h * rN.z;
            hn.w = a.ph.v0[0] * hn.x + a.ph.v0[1] * hn.y + a.ph.v0[2] * hn.z;
            reinterpret_cast<V4*>(sN)[2 * tid] = rN;
            reinterpret_cast<V4*>(sN)[2 * tid + 1] = hn;
        } else {
            reinterpret_cast<V4*>(sN)[tid] = rN;
        }
    }
    if constexpr (NFC == 4) {
        if (tid < nE) {
            reinterpret_cast<U4*>(sNb)[tid] = rNb4;
            reinterpret_cast<uint32_t*>(sCd8)[tid] = rCd4;
        }
    } else {
        if (faceRow) {
            sNb[tid] = rNb1;
            sCd8[tid] = rCd1;
        }
    }
    __syncthreads();

    // ---- phase 2: neighbour traces of all of this thread's face nodes, then the fluxes ----
    // (boundary faces and idle threads read the batch's first node: a valid address whose
    // value the flux does not use)
    T qp[kFIt][NF];
#pragma unroll
    for (int c = 0; c < kFIt; ++c) {
        const int it = tid + c * kBlock;
        const T* src = a.qin + e0 * NP * NF;
        if (it < nE * NFT) {
            const int el = it / NFT, fa = it - el * NFT;
            const int f = fa / NFP, an = fa - f * NFP;
            const int efl = el * NFC + f;
            const int nb = sNb[efl];
            const int cd = sCd8[efl];
            const int fp = cd & 3, pc = cd >> 2;
            const int nn = sI[L::kMask + fp * NFP + sI[L::kPerm + pc * NFP + an]];
            if (nb >= 0) src = a.qin + (static_cast<int64_t>(nb) * NP + nn) * NF;
            else if (nb <= -3) src = a.ghost + (static_cast<int64_t>(-3 - nb) * NFP + an) * NF;
        }
        loadNode<NF>(src, qp[c]);
    }
    // snapshot for the residuals, after the gathers (its wait then coincides with theirs)
    if constexpr (kFinal) {
        loadNodeNT<NF>((a.snap ? a.snap : a.qin) + ldBase, ps);
    }
    T Dq[DIM][NF];
#pragma unroll
    for (int r = 0; r < DIM; ++r)
#pragma unroll
        for (int k = 0; k < NF; ++k) Dq[r][k] = T(0);
    if constexpr (VM) {
        // Dq_r = Dw_r q on the matrix cores (while the gathers are in flight), node-major into
        // sF, then into the node threads' registers
        const int lane = tid & 63, wave = tid >> 6;
        const int rt = wave % V::MT, j = lane & 15, kq = lane >> 4;
        int dOff[4];  // offsets of this lane's 4 result rows (node i, direction rd) in an element
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = 16 * rt + MF::row(lane, r);
            const int rd = row / NP, i = row - rd * NP;
            dOff[r] = row < V::MR ? i * V::DS + rd * NF : -1;
        }
#pragma unroll
        for (int m = 0; m < V::CPW; ++m) {
            const int ct = wave / V::MT + m * V::G;
            const int el = ct * V::ECT + j / NF, f = j % NF;
            const bool colOk = j < V::ECT * NF && el < nE;
            const T* qc = &sQ[(colOk ? el : 0) * QS + f];
            typename MF::V acc = {T(0), T(0), T(0), T(0)};
#pragma unroll
            for (int ks = 0; ks < V::KS; ++ks) {
                const int n = 4 * ks + kq;
                const T b = (colOk && (4 * V::KS == NP || n < NP)) ? qc[n * NF] : T(0);
                acc = MF::mma(rA[ks], b, acc);
            }
#pragma unroll
            for (int r = 0; r < 4; ++r)
                if (colOk && dOff[r] >= 0) sF[el * V::ES + dOff[r] + f] = acc[r];
        }
        __syncthreads();
        if (active) {
            const int el = tid / NP, i = tid - el * NP;
            const T* d = &sF[el * V::ES + i * V::DS];
#pragma unroll
            for (int r = 0; r < DIM; ++r)
#pragma unroll
                for (int k = 0; k < NF; ++k) Dq[r][k] = d[r * NF + k];
        }
        __syncthreads();
    }
#pragma unroll
    for (int c = 0; c < kFIt; ++c) {
        const int it = tid + c * kBlock;
        if (it < nE * NFT) {
            const int el = it / NFT, fa = it - el * NFT;
            const int f = fa / NFP, an = fa - f * NFP;
            const int efl = el * NFC + f;
            const T* nf = &sN[efl * NSV];
            const int nb = sNb[efl];
            const int own = sI[L::kMask + f * NFP + an];
            T qm[NF], pi[NF];
            loadNode<NF>(&sQ[el * QS + own * NF], qm);
            if (NSV == 8 && (nb >= 0 || nb <= -3)) {
                // interior (and learned-closure) faces, scaled by fs:
                // (fs/2) n.F(q- + q+) + fs (c0/2)(q- - q+), F linear in q
                const T hn[3] = {nf[4], nf[5], nf[6]};
                const T hv = nf[7], cf = nf[3] * (T(0.5) * a.ph.c0);
                const T* qp0 = qp[c];
                T qs[NF], qd[NF];
#pragma unroll
                for (int k = 0; k < NF; ++k) {
                    qs[k] = qm[k] + qp0[k];
                    qd[k] = qm[k] - qp0[k];
                }
                const T nvs = hn[0] * qs[1] + hn[1] * qs[2] + hn[2] * qs[3];
                const T t = a.ph.irho0 * qs[0];
                pi[0] = hv * qs[0] + a.ph.rc2 * nvs + cf * qd[0];
#pragma unroll
                for (int x = 0; x < 3; ++x) pi[1 + x] = hv * qs[1 + x] + hn[x] * t + cf * qd[1 + x];
                if constexpr (NF == 5) pi[4] = hv * qs[4] + a.ph.rho0 * nvs + cf * qd[4];
            } else {
                const T nv[3] = {nf[0], nf[1], nf[2]};
                const T fs = nf[3];
                numericalFlux<DIM, NF>(a.ph, nv, nb <= -3 ? 0 : nb, qm, qp[c], pi);
#pragma unroll
                for (int k = 0; k < NF; ++k) pi[k] *= fs;
            }
#pragma unroll
            for (int k = 0; k < NF; ++k) sF[el * FS + fa * NF + k] = pi[k];
        }
    }
    __syncthreads();

    // ---- volume contraction, flux-Jacobian fold, lift, register update ----
    if (active) {
        const int el = tid / NP, i = tid - el * NP;
        const T* q0 = &sQ[el * QS];
#pragma unroll 4
        for (int j = 0; j < (VM ? 0 : NP); ++j) {
            T q[NF];
#pragma unroll
            for (int k = 0; k < NF; ++k) q[k] = q0[j * NF + k];
#pragma unroll
            for (int r = 0; r < DIM; ++r) {
                const T d = opAt(L::kDw + (j * NP + i) * DIM + r);
#pragma unroll
                for (int k = 0; k < NF; ++k) Dq[r][k] += d * q[k];
            }
        }
        // rhs = sum_x A_x W_x with W_x = sum_r (dxi_r/dx_x) Dq_r, A_x the flux Jacobians, in the
        // order sum_r B_r Dq_r: a_r = sum_x (dxi_r/dx_x) v0_x multiplies every field, the
        // pressure-velocity coupling enters through sdiv = sum_r sum_x (dxi_r/dx_x) Dq_r[v_x]
        // and sum_r (dxi_r/dx_k) Dq_r[p]
        const T* J = &sJ[el * JJ];
        T ar[DIM];
#pragma unroll
        for (int r = 0; r < DIM; ++r) {
            T s = T(0);
#pragma unroll
            for (int x = 0; x < DIM; ++x) s += J[r * DIM + x] * a.ph.v0[x];
            ar[r] = s;
        }
        T rhs[NF];
        {
            T sdiv = T(0);
#pragma unroll
            for (int r = 0; r < DIM; ++r)
#pragma unroll
                for (int x = 0; x < DIM; ++x) sdiv += J[r * DIM + x] * Dq[r][1 + x];
#pragma unroll
            for (int k = 0; k < NF; ++k) {
                T s = T(0);
#pragma unroll
                for (int r = 0; r < DIM; ++r) s += ar[r] * Dq[r][k];
                rhs[k] = s;
            }
            rhs[0] += a.ph.rc2 * sdiv;
            if constexpr (NF == 5) rhs[4] += a.ph.rho0 * sdiv;
#pragma unroll
            for (int kv = 0; kv < DIM; ++kv) {
                T g = T(0);
#pragma unroll
                for (int r = 0; r < DIM; ++r) g += J[r * DIM + kv] * Dq[r][0];
                rhs[1 + kv] += a.ph.irho0 * g;
            }
        }
        const T* f0 = &sF[el * FS];
#pragma unroll 4
        for (int fa = 0; fa < NFT; ++fa) {
            const T l = kPlanes ? opAt(L::kL - kOpsOff + fa * NP + i) : lift[fa * NP + i];
#pragma unroll
            for (int k = 0; k < NF; ++k) rhs[k] -= l * f0[fa * NF + k];
        }

        const int64_t base = myBase;
        const T dt = a.dt;
        if constexpr (MODE == kModeRHS) {
            T o[NF];
#pragma unroll
            for (int k = 0; k < NF; ++k) o[k] = dt * rhs[k];
            storeNode<NF>(a.qout + base, o);
        } else if constexpr (MODE == kModeFirst || MODE == kModeMid) {
            const T half = (MODE == kModeMid && a.stage == kStageRK3) ? T(1) : T(0.5);
            T accn[NF], qn[NF];
#pragma unroll
            for (int k = 0; k < NF; ++k) {
                const T kk = dt * rhs[k];
                T uv;
                if constexpr (MODE == kModeFirst) {
                    uv = q0[i * NF + k];
                    accn[k] = kk;
                } else {
                    uv = pu[k];
                    accn[k] = pacc[k] + T(2) * kk;
                }
                qn[k] = uv + half * kk;
            }
            storeNodeNT<NF>(a.acc + base, accn);
            storeNode<NF>(a.qout + base, qn);
        } else {
            // final update: u += (k1 + 2k2 + 2k3 + k4)/6  (RK4)  or  u += k  (Euler)
            T unew[NF], lnew[NF];
            double lo[NF];
#pragma unroll
            for (int k = 0; k < NF; ++k) {
                const T kk = dt * rhs[k];
                const T delta = (MODE == kModeFinal) ? (pacc[k] + kk) / T(6) : kk;
                const T uv = (MODE == kModeFinal) ? pu[k] : q0[i * NF + k];
                if constexpr (COMP) {
                    T s, e, t, e2;
                    twoSum(plo[k], delta, t, e2);
                    twoSum(uv, t, s, e);
                    const T nlo = e + e2;
                    T h, l2;
                    twoSum(s, nlo, h, l2);
                    unew[k] = h;
                    lnew[k] = l2;
                    lo[k] = static_cast<double>(l2);
                } else {
                    unew[k] = uv + delta;
                    lo[k] = 0.0;
                }
            }
            storeNodeNT<NF>(a.uout + base, unew);
            if constexpr (COMP) storeNodeNT<NF>(a.ulo + base, lnew);
            if (!allFinite<NF>(unew)) a.stateBad[0] = 1;
            if (a.snap) {
                double d[4];
                double sv[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    sv[k] = static_cast<double>(ps[k]);
                    d[k] = (static_cast<double>(unew[k]) + lo[k]) - sv[k];
                }
                const double ic2 = static_cast<double>(a.ph.ic2);
                const double u0 = static_cast<double>(unew[0]) + lo[0];
                const double sp = sv[0] * ic2 - u0 * ic2;
                res[0] += d[0] * d[0];
                res[1] += sp * sp;
                res[2] += d[1] * d[1];
                res[3] += d[2] * d[2];
                res[4] += d[3] * d[3];
            }
        }
    }
    if constexpr (kFinal) {
        if (a.snap) {
            blockResiduals<kBlock>(res, sRed, a.resPartial + static_cast<int64_t>(blockIdx.x) * 5);
        }
    }
}

// ------------------------------------------------------------ MFMA building blocks ----
//
// Matrix-core form of the two contractions of a stage (4 fields):
//   rhs = [Dw_1 | .. | Dw_DIM] (NP x DIM*NP) . G  +  (-LIFT) (NP x NFT) . (Fscale pi)
// with G_r = sum_x (dxi_r/dx_x) F_x(q). Columns of the B operand are (4 elements x 4
// fields). FP32 uses v_mfma_f32_16x16x4f32 (D row = 4*(lane/16)+r), FP64
// v_mfma_f64_16x16x4f64 (D row = lane/16 + 4*r); A lane = row + 16*k, B lane = col + 16*k
// for both (layouts checked on gfx950).

template <int DIM, int P>
struct MfmaShape {
    using S = Sizes<DIM, P>;
    static constexpr int MT = (S::NP + 15) / 16;       // row tiles
    static constexpr int KV = DIM * S::NP;              // volume contraction length
    static constexpr int KSV = (KV + 3) / 4;            // k-steps (volume)
    static constexpr int KSL = (S::NFT + 3) / 4;        // k-steps (lift)
    static constexpr int KS = KSV + KSL;
    static constexpr bool kSupported = MT * KS <= 64;   // A fragments fit in registers
    static constexpr int kFragTotal = MT * KS * 64;
};

// ------------------------------------------------- wave-autonomous MFMA stage kernel ----
//
// Each wave owns groups of 4 elements and runs the whole stage for them without block-wide
// barriers: state to wave-private LDS, face normals, numerical fluxes, then one matrix-core
// accumulation rhs = [Dw_1|..|Dw_DIM] . G + (-LIFT) . (Fscale pi) where the B operand G is
// folded on the fly from the staged state (G_r = sum_x (dxi_r/dx_x) F_x(q)), then an LDS
// transpose so that the RK register update is done with coalesced 16-byte accesses.
// The A fragments (constant operators) are staged once per block in LDS.

__device__ inline void waveSync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}


template <int DIM, int P>
struct WaveShape {
    using M = MfmaShape<DIM, P>;
    static constexpr int NG = 2;            // element groups (of 4) per wave
    static constexpr int EB = 4 * 4 * NG;   // elements per block
};

template <int DIM, int P, typename T, bool COMP>
__global__ __launch_bounds__(kBlock) void stageKernelWave(
    StageArgs<T> a, const T* __restrict__ afrag) {
    using S = Sizes<DIM, P>;
    using L = OpLayout<DIM, P>;
    using M = MfmaShape<DIM, P>;
    using W = WaveShape<DIM, P>;
    using MT_ = MfmaTraits<T>;
    using V4 = typename Vec4<T>::V;
    constexpr int NP = S::NP, NFP = S::NFP, NFC = S::NFC, NFT = S::NFT, NF = 4;
    constexpr int JJ = DIM * DIM;
    constexpr int NCH = (4 * NP + 63) / 64;  // float4 chunks (nodes) per lane per group

    __shared__ T sA[M::kFragTotal];
    __shared__ int32_t sI[L::kITotal];
    __shared__ __align__(16) T sQ[4][4 * NP * NF];    // per wave: staged state of its group
    __shared__ __align__(16) T sF[4][M::KSL * 4 * 16];  // per wave: lift B rows (fa x 16 cols)
    __shared__ __align__(16) T sR[4][4 * NP * NF];    // per wave: rhs transpose
    __shared__ T sJ[4][4 * JJ];
    __shared__ T sN[4][4 * NFC * 4];
    __shared__ double sRed[kBlock / 16][5];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t blk = xcdRemap(blockIdx.x, gridDim.x);
    const int stage = a.stage;
    const bool needU = stage == kStageRK2 || stage == kStageRK3 || stage == kStageRK4;
    const bool isFinal = stage == kStageRK4 || stage == kStageEuler;

    for (int x = tid; x < M::kFragTotal; x += kBlock) sA[x] = afrag[x];
    for (int x = tid; x < L::kITotal; x += kBlock) sI[x] = a.iops[x];
    __syncthreads();

    const int col = lane & 15, cel = col >> 2, fld = col & 3, brow = lane >> 4;
    T* Q = sQ[wave];
    T* F = sF[wave];
    T* R = sR[wave];
    T* Jw = sJ[wave];
    T* Nw = sN[wave];
    double res[5] = {0, 0, 0, 0, 0};

    for (int g = 0; g < W::NG; ++g) {
        const int64_t e0 = (blk * 4 + wave) * (4 * W::NG) + 4 * g;  // first element of the group
        if (e0 >= a.K) break;
        const int nE = static_cast<int>((a.K - e0) < 4 ? (a.K - e0) : 4);
        const int nNodes = nE * NP;
        // -- stage the group's state and geometry; prefetch this lane's RK operands
        const V4* qin4 = reinterpret_cast<const V4*>(a.qin + e0 * NP * NF);
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            const int node = lane + 64 * c;
            if (node < nNodes) reinterpret_cast<V4*>(Q)[node] = qin4[node];
        }
        for (int x = lane; x < nE * JJ; x += 64) Jw[x] = a.geo[e0 * JJ + x];
        for (int x = lane; x < M::KSL * 4 * 16; x += 64) F[x] = T(0);
        waveSync();
        // -- unit normals and Fscale of the group's faces
        for (int x = lane; x < nE * NFC * 4; x += 64) Nw[x] = a.fgeo[e0 * NFC * 4 + x];
        waveSync();
        // -- numerical fluxes at the group's face nodes -> lift B rows
        for (int it = lane; it < nE * NFT; it += 64) {
            const int e = it / NFT, fa = it - e * NFT;
            const int f = fa / NFP, an = fa - f * NFP;
            const int64_t eg = e0 + e;
            const T* n4 = &Nw[(e * NFC + f) * 4];
            const T nv[3] = {n4[0], n4[1], n4[2]};
            const T fs = n4[3];
            const int own = sI[L::kMask + f * NFP + an];
            T qm[NF], qp[NF], pi[NF];
            {
                const V4 v = reinterpret_cast<const V4*>(Q)[e * NP + own];
                qm[0] = v.x; qm[1] = v.y; qm[2] = v.z; qm[3] = v.w;
            }
            const int nb = a.nbr[eg * NFC + f];
            if (nb >= 0) {
                const int cd = a.code[eg * NFC + f];
                const int fp = cd & 3, pc = cd >> 2;
                const int nn = sI[L::kMask + fp * NFP + sI[L::kPerm + pc * NFP + an]];
                const V4 v = reinterpret_cast<const V4*>(a.qin)[static_cast<int64_t>(nb) * NP + nn];
                qp[0] = v.x; qp[1] = v.y; qp[2] = v.z; qp[3] = v.w;
            } else if (nb <= -3) {
                const V4 v = reinterpret_cast<const V4*>(a.ghost)[static_cast<int64_t>(-3 - nb) * NFP + an];
                qp[0] = v.x; qp[1] = v.y; qp[2] = v.z; qp[3] = v.w;
            } else {
                qp[0] = qp[1] = qp[2] = qp[3] = T(0);
            }
            numericalFlux<DIM, NF>(a.ph, nv, nb <= -3 ? 0 : nb, qm, qp, pi);
            V4 out;
            out.x = fs * pi[0]; out.y = fs * pi[1]; out.z = fs * pi[2]; out.w = fs * pi[3];
            reinterpret_cast<V4*>(F)[fa * 4 + e] = out;
        }
        // this lane's element: rows of dxi/dx and v0 . row
        T Jl[DIM][DIM], vr[DIM];
        const bool colOk = cel < nE;
#pragma unroll
        for (int r = 0; r < DIM; ++r) {
            T s0 = T(0);
#pragma unroll
            for (int xx = 0; xx < DIM; ++xx) {
                Jl[r][xx] = colOk ? Jw[cel * JJ + r * DIM + xx] : T(0);
                s0 += Jl[r][xx] * a.ph.v0[xx];
            }
            vr[r] = s0;
        }
        waveSync();
        // -- matrix-core contraction; B of the volume part folded on the fly
        typename MT_::V acc[M::MT];
#pragma unroll
        for (int mt = 0; mt < M::MT; ++mt) acc[mt] = typename MT_::V{T(0), T(0), T(0), T(0)};
#pragma unroll 3
        for (int ks = 0; ks < M::KSV; ++ks) {
            const int kk = 4 * ks + brow;
            T b = T(0);
            if (kk < M::KV && colOk) {
                const int r = kk / NP, j = kk - r * NP;
                const V4 q = reinterpret_cast<const V4*>(Q)[cel * NP + j];
                T jr[DIM];
                T jv = T(0), vrr = T(0);
#pragma unroll
                for (int rr = 0; rr < DIM; ++rr)
                    if (rr == r) {
#pragma unroll
                        for (int xx = 0; xx < DIM; ++xx) jr[xx] = Jl[rr][xx];
                        vrr = vr[rr];
                    }
                jv = jr[0] * q.y;
                if (DIM > 1) jv += jr[1] * q.z;
                if (DIM > 2) jv += jr[DIM > 2 ? 2 : 0] * q.w;
                const T qf = fld == 0 ? q.x : (fld == 1 ? q.y : (fld == 2 ? q.z : q.w));
                b = vrr * qf;
                if (fld == 0)
                    b += a.ph.rc2 * jv;
                else if (fld - 1 < DIM)
                    b += jr[fld - 1 < DIM ? fld - 1 : 0] * q.x * a.ph.irho0;
            }
#pragma unroll
            for (int mt = 0; mt < M::MT; ++mt) acc[mt] = MT_::mma(sA[(mt * M::KS + ks) * 64 + lane], b, acc[mt]);
        }
#pragma unroll 2
        for (int ks = 0; ks < M::KSL; ++ks) {
            const T b = F[(4 * ks + brow) * 16 + col];
#pragma unroll
            for (int mt = 0; mt < M::MT; ++mt)
                acc[mt] = MT_::mma(sA[(mt * M::KS + M::KSV + ks) * 64 + lane], b, acc[mt]);
        }
        // -- transpose the rhs to [element][node][field]
#pragma unroll
        for (int mt = 0; mt < M::MT; ++mt)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = 16 * mt + MT_::row(lane, r);
                if (i < NP) R[(cel * NP + i) * NF + fld] = acc[mt][r];
            }
        waveSync();
        // -- register update, one node (4 fields) per lane and chunk
        const T dt = a.dt;
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            const int node = lane + 64 * c;
            if (node >= nNodes) continue;
            const int64_t off = e0 * NP + node;
            const V4 rv = reinterpret_cast<const V4*>(R)[node];
            const T kk[4] = {dt * rv.x, dt * rv.y, dt * rv.z, dt * rv.w};
            if (stage == kStageRHS) {
                V4 o;
                o.x = kk[0]; o.y = kk[1]; o.z = kk[2]; o.w = kk[3];
                reinterpret_cast<V4*>(a.qout)[off] = o;
                continue;
            }
            const V4 q0 = reinterpret_cast<const V4*>(Q)[node];
            const T qv[4] = {q0.x, q0.y, q0.z, q0.w};
            V4 pu[1], pacc[1], plo[1], ps[1];
            const int cc = c;
            c = 0;
            if (needU) {
                pu[0] = reinterpret_cast<const V4*>(a.u)[off];
                pacc[0] = reinterpret_cast<const V4*>(a.acc)[off];
            }
            plo[0] = V4{T(0), T(0), T(0), T(0)};
            if constexpr (COMP)
                if (isFinal) plo[0] = reinterpret_cast<const V4*>(a.ulo)[off];
            if (isFinal && a.snap) ps[0] = reinterpret_cast<const V4*>(a.snap)[off];
            const T uv4[4] = {pu[c].x, pu[c].y, pu[c].z, pu[c].w};
            const T ac4[4] = {pacc[c].x, pacc[c].y, pacc[c].z, pacc[c].w};
            T lo4[4] = {T(0), T(0), T(0), T(0)};
            if constexpr (COMP) {
                lo4[0] = plo[c].x; lo4[1] = plo[c].y; lo4[2] = plo[c].z; lo4[3] = plo[c].w;
            }
            if (stage <= kStageRK3) {
                const T half = (stage == kStageRK3) ? T(1) : T(0.5);
                T na[4], nq[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const T uv = (stage == kStageRK1) ? qv[k] : uv4[k];
                    na[k] = (stage == kStageRK1) ? kk[k] : ac4[k] + T(2) * kk[k];
                    nq[k] = uv + half * kk[k];
                }
                V4 oa, oq;
                oa.x = na[0]; oa.y = na[1]; oa.z = na[2]; oa.w = na[3];
                oq.x = nq[0]; oq.y = nq[1]; oq.z = nq[2]; oq.w = nq[3];
                reinterpret_cast<V4*>(a.acc)[off] = oa;
                reinterpret_cast<V4*>(a.qout)[off] = oq;
            } else {
                T hi[4], lw[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const T delta = (stage == kStageRK4) ? (ac4[k] + kk[k]) / T(6) : kk[k];
                    const T uv = (stage == kStageRK4) ? uv4[k] : qv[k];
                    if constexpr (COMP) {
                        T sum, e1, t, e2;
                        twoSum(lo4[k], delta, t, e2);
                        twoSum(uv, t, sum, e1);
                        twoSum(sum, e1 + e2, hi[k], lw[k]);
                    } else {
                        hi[k] = uv + delta;
                        lw[k] = T(0);
                    }
                }
                V4 oh;
                oh.x = hi[0]; oh.y = hi[1]; oh.z = hi[2]; oh.w = hi[3];
                reinterpret_cast<V4*>(a.uout)[off] = oh;
                if (!allFinite<4>(hi)) a.stateBad[0] = 1;
                if constexpr (COMP) {
                    V4 ol;
                    ol.x = lw[0]; ol.y = lw[1]; ol.z = lw[2]; ol.w = lw[3];
                    reinterpret_cast<V4*>(a.ulo)[off] = ol;
                }
                if (a.snap) {
                    const T s4[4] = {ps[c].x, ps[c].y, ps[c].z, ps[c].w};
                    double d[4], sv0 = 0.0, un0 = 0.0;
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        const double sv = static_cast<double>(s4[k]);
                        const double un = static_cast<double>(hi[k]) + static_cast<double>(lw[k]);
                        d[k] = un - sv;
                        if (k == 0) { sv0 = sv; un0 = un; }
                    }
                    const double ic2 = static_cast<double>(a.ph.ic2);
                    const double dr = sv0 * ic2 - un0 * ic2;
                    res[0] += d[0] * d[0];
                    res[1] += dr * dr;
                    res[2] += d[1] * d[1];
                    res[3] += d[2] * d[2];
                    res[4] += d[3] * d[3];
                }
            }
            c = cc;
        }
        waveSync();
    }
    if (isFinal && a.snap) blockResiduals<kBlock>(res, sRed, a.resPartial + static_cast<int64_t>(blockIdx.x) * 5);
}

// --------------------------------------------------- block matrix-core stage kernel ----
//
// 16 elements per 256-thread block (4 waves x one 16-column tile of 4 elements x 4 fields):
//  1. state, geometry and per-face data to LDS;
//  2. VALU: the flux-Jacobian-folded volume operand G_r[j] = sum_x (dxi_r/dx_x) A_x q_j once
//     per node, and the scaled numerical fluxes Fscale * pi once per face node, written as
//     the B operand in matrix-core fragment order ([k-step][column tile][lane]; conflict-free);
//  3. matrix core: rhs = [Dw_1 | .. | Dw_DIM | -LIFT] . B with the A fragments resident in
//     registers (one row tile per wave where the row tiles divide the 4 waves), two
//     accumulation chains per tile;
//  4. the results go to LDS node-major and the RK update is done with 16-byte accesses, as in
//     stageKernel.
template <int DIM, int P>
struct BlockMmShape {
    using M = MfmaShape<DIM, P>;
    static constexpr int EB = 16;                        // elements per block
    static constexpr int KROWS = 4 * M::KS;              // B rows (volume, then lift)
    static constexpr int kBFloats = M::KS * 4 * 64;      // B fragments of the 4 column tiles
    static constexpr int KSP = (M::KS + 3) / 4 * 4;      // A values per lane and row tile (padded)
    // MT in {1, 2, 4}: each wave takes one row tile and MT column tiles (A fragments of one row
    // tile per lane); otherwise every row tile of one column tile
    static constexpr bool kSplit = (4 % M::MT) == 0;
    static constexpr int kMtW = kSplit ? 1 : M::MT;      // row tiles per wave
    static constexpr int kCtW = kSplit ? M::MT : 1;      // column tiles per wave
    static constexpr int kAPerLane = kMtW * KSP;
    static constexpr bool kSupported = M::kSupported && Sizes<DIM, P>::NP <= KROWS;
};

template <int DIM, int P, typename T, bool COMP, int MODE>
__global__ __launch_bounds__(kBlock) void stageKernelMM(StageArgs<T> a, const T* __restrict__ afrag) {
    using S = Sizes<DIM, P>;
    using L = OpLayout<DIM, P>;
    using M = MfmaShape<DIM, P>;
    using B = BlockMmShape<DIM, P>;
    using MT_ = MfmaTraits<T>;
    using V4 = typename Vec4<T>::V;
    constexpr int NP = S::NP, NFP = S::NFP, NFC = S::NFC, NFT = S::NFT, NF = 4, EB = B::EB;
    constexpr int JJ = DIM * DIM;
    constexpr int LIFT0 = 4 * M::KSV;  // first B row of the lift part
    constexpr bool kNeedU = MODE == kModeMid || MODE == kModeFinal;
    constexpr bool kFinal = MODE == kModeFinal || MODE == kModeEuler;
    // per-thread item counts (loads of a phase are all issued before the first wait; addresses of
    // idle items are clamped to valid ones so that no load sits under a branch)
    constexpr int kQIt = (EB * NP + kBlock - 1) / kBlock;    // staged nodes
    constexpr int kJIt = (EB * JJ + kBlock - 1) / kBlock;    // inverse-Jacobian entries
    constexpr int kRIt = (EB * NFC + kBlock - 1) / kBlock;   // face rows
    constexpr int kIIt = (L::kITotal + kBlock - 1) / kBlock; // index-table entries
    constexpr int kFIt = (EB * NFT + kBlock - 1) / kBlock;   // face nodes
    constexpr int NIT = (4 * NP + 63) / 64;                  // update items per lane (it = lane + 64 c)
    constexpr int padV = LIFT0 - M::KV, padL = B::KROWS - LIFT0 - NFT;
    constexpr int kPadIt = ((padV + padL) * 64 + kBlock - 1) / kBlock;

    constexpr int QS = ldsStride(NP * NF);  // padded per-element stride (LDS banks)
    __shared__ __align__(32) T sB[B::kBFloats];
    __shared__ __align__(32) T sQ[EB * QS];
    __shared__ __align__(32) T sR[EB * QS];  // right-hand sides, node-major
    __shared__ T sJ[EB * JJ];
    __shared__ __align__(32) T sN[EB * NFC * 4];
    __shared__ int32_t sNb[EB * NFC];
    __shared__ int32_t sCd[EB * NFC];
    __shared__ int32_t sI[L::kITotal];
    __shared__ double sRed[kBlock / 16][5];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t blk = xcdRemap(blockIdx.x, gridDim.x);
    const int64_t e0 = blk * EB;
    const int nE = static_cast<int>((a.K - e0) < EB ? (a.K - e0) : EB);
    const int nEw = nE - 4 * wave;  // elements of this wave's column tile present in the block

    // ---- phase 1: issue every load ----
    T rq[kQIt][NF];
#pragma unroll
    for (int c = 0; c < kQIt; ++c) {
        const int x = tid + c * kBlock;
        loadNode<NF>(a.qin + (e0 * NP + (x < nE * NP ? x : 0)) * NF, rq[c]);
    }
    T rJ[kJIt];
#pragma unroll
    for (int c = 0; c < kJIt; ++c) {
        const int x = tid + c * kBlock;
        rJ[c] = a.geo[e0 * JJ + (x < nE * JJ ? x : 0)];
    }
    V4 rN[kRIt];
    int32_t rNb[kRIt], rCd[kRIt];
#pragma unroll
    for (int c = 0; c < kRIt; ++c) {
        const int x = tid + c * kBlock;
        const int xr = x < nE * NFC ? x : 0;
        rN[c] = reinterpret_cast<const V4*>(a.fgeo + e0 * NFC * 4)[xr];
        rNb[c] = a.nbr[e0 * NFC + xr];
        rCd[c] = a.code[e0 * NFC + xr];
    }
    int32_t rI[kIIt];
#pragma unroll
    for (int c = 0; c < kIIt; ++c) {
        const int x = tid + c * kBlock;
        rI[c] = a.iops[x < L::kITotal ? x : 0];
    }
    // RK operands of this lane's update items (node-major over the wave's 4 elements)
    int64_t uNode[NIT];
    bool uOk[NIT];
    T pu[NIT][NF], pacc[NIT][NF], plo[NIT][NF], ps[NIT][NF];
#pragma unroll
    for (int c = 0; c < NIT; ++c) {
        const int it = lane + 64 * c;
        const int el = it / NP, i = it - el * NP;
        uOk[c] = it < 4 * NP && el < nEw;
        uNode[c] = uOk[c] ? (e0 + 4 * wave + el) * NP + i : e0 * NP;
#pragma unroll
        for (int k = 0; k < NF; ++k) pu[c][k] = pacc[c][k] = plo[c][k] = ps[c][k] = T(0);
        if constexpr (kNeedU) {
            loadNodeNT<NF>(a.u + uNode[c] * NF, pu[c]);
            loadNodeNT<NF>(a.acc + uNode[c] * NF, pacc[c]);
        }
        if constexpr (COMP && kFinal) loadNodeNT<NF>(a.ulo + uNode[c] * NF, plo[c]);
    }

    // ---- LDS; zero the B padding rows ----
#pragma unroll
    for (int c = 0; c < kQIt; ++c) {
        const int x = tid + c * kBlock;
        if (x < nE * NP) {
            const int e = x / NP;
            storeNode<NF>(&sQ[e * QS + (x - e * NP) * NF], rq[c]);
        }
    }
#pragma unroll
    for (int c = 0; c < kJIt; ++c) {
        const int x = tid + c * kBlock;
        if (x < nE * JJ) sJ[x] = rJ[c];
    }
#pragma unroll
    for (int c = 0; c < kRIt; ++c) {
        const int x = tid + c * kBlock;
        if (x < nE * NFC) {
            reinterpret_cast<V4*>(sN)[x] = rN[c];
            sNb[x] = rNb[c];
            sCd[x] = rCd[c];
        }
    }
#pragma unroll
    for (int c = 0; c < kIIt; ++c) {
        const int x = tid + c * kBlock;
        if (x < L::kITotal) sI[x] = rI[c];
    }
#pragma unroll
    for (int c = 0; c < kPadIt; ++c) {
        const int x = tid + c * kBlock;
        if (x < (padV + padL) * 64) {
            const int pr = x / 64, col = x - pr * 64;
            const int kk = pr < padV ? M::KV + pr : LIFT0 + NFT + (pr - padV);
            sB[((kk >> 2) * 4 + (col >> 4)) * 64 + (kk & 3) * 16 + (col & 15)] = T(0);
        }
    }
    __syncthreads();

    // B fragment index of (row kk, column 4*e + field 0); the 4 fields are contiguous
    auto bIndex = [](int kk, int e) { return ((kk >> 2) * 4 + (e >> 2)) * 64 + (kk & 3) * 16 + (e & 3) * 4; };

    // ---- phase 2: neighbour traces of all of this thread's face nodes ----
    // (boundary faces and idle items read the block's first node: a valid address whose value
    // the flux does not use)
    T qp[kFIt][NF];
#pragma unroll
    for (int c = 0; c < kFIt; ++c) {
        const int it = tid + c * kBlock;
        const T* src = a.qin + e0 * NP * NF;
        if (it < nE * NFT) {
            const int e = it / NFT, fa = it - e * NFT;
            const int f = fa / NFP, an = fa - f * NFP;
            const int ef = e * NFC + f;
            const int nb = sNb[ef];
            const int cd = sCd[ef];
            const int fp = cd & 3, pc = cd >> 2;
            const int nn = sI[L::kMask + fp * NFP + sI[L::kPerm + pc * NFP + an]];
            if (nb >= 0) src = a.qin + (static_cast<int64_t>(nb) * NP + nn) * NF;
            else if (nb <= -3) src = a.ghost + (static_cast<int64_t>(-3 - nb) * NFP + an) * NF;
        }
        loadNode<NF>(src, qp[c]);
    }
    if constexpr (kFinal) {
#pragma unroll
        for (int c = 0; c < NIT; ++c) {
            loadNodeNT<NF>((a.snap ? a.snap : a.qin) + uNode[c] * NF, ps[c]);
        }
    }
    // A fragments of this lane and the wave's row tiles (constant operators), stored
    // [row tile][lane][k-step] so that they load as 16-byte vectors; issued here so that they
    // arrive with the traces
    const int mt0 = B::kSplit ? wave % M::MT : 0;                 // first row tile of the wave
    const int ct0 = B::kSplit ? (wave / M::MT) * M::MT : wave;    // first column tile of the wave
    T af[B::kAPerLane];
#pragma unroll
    for (int m = 0; m < B::kMtW; ++m) {
        const V4* src = reinterpret_cast<const V4*>(afrag + (static_cast<int64_t>(mt0 + m) * 64 + lane) * B::KSP);
#pragma unroll
        for (int x = 0; x < B::KSP / 4; ++x) {
            const V4 v = src[x];
            af[m * B::KSP + 4 * x] = v.x;
            af[m * B::KSP + 4 * x + 1] = v.y;
            af[m * B::KSP + 4 * x + 2] = v.z;
            af[m * B::KSP + 4 * x + 3] = v.w;
        }
    }

    // 2a. volume operand G_r[j] = sum_x J_rx F_x(q_j), F_x = (v0_x p + rho0 c0^2 v_x, v0_x v + e_x p / rho0)
    //     (one item per element node; overlaps the traces in flight)
#pragma unroll
    for (int c = 0; c < kQIt; ++c) {
        const int it = tid + c * kBlock;
        if (it < nE * NP) {
            const int e = it / NP, j = it - e * NP;
            const T* q = rq[c];
            const T* J = &sJ[e * JJ];
#pragma unroll
            for (int r = 0; r < DIM; ++r) {
                T w = T(0);
                T jx[3] = {T(0), T(0), T(0)};
#pragma unroll
                for (int x = 0; x < DIM; ++x) {
                    jx[x] = J[r * DIM + x];
                    w += jx[x] * a.ph.v0[x];
                }
                T jv = jx[0] * q[1];
                if (DIM > 1) jv += jx[1] * q[2];
                if (DIM > 2) jv += jx[DIM > 2 ? 2 : 0] * q[3];
                const T pq = q[0] * a.ph.irho0;
                T g[NF];
                g[0] = w * q[0] + a.ph.rc2 * jv;
                g[1] = w * q[1] + jx[0] * pq;
                g[2] = w * q[2] + jx[1] * pq;
                g[3] = w * q[3] + jx[2] * pq;
                storeNode<NF>(&sB[bIndex(r * NP + j, e)], g);
            }
        }
    }
    // 2b. scaled numerical fluxes (one item per element face node)
#pragma unroll
    for (int c = 0; c < kFIt; ++c) {
        const int it = tid + c * kBlock;
        if (it < nE * NFT) {
            const int e = it / NFT, fa = it - e * NFT;
            const int f = fa / NFP, an = fa - f * NFP;
            const int ef = e * NFC + f;
            const T nv[3] = {sN[ef * 4 + 0], sN[ef * 4 + 1], sN[ef * 4 + 2]};
            const T fs = sN[ef * 4 + 3];
            const int own = sI[L::kMask + f * NFP + an];
            const int nb = sNb[ef];
            T qm[NF], pi[NF], out[NF];
            loadNode<NF>(&sQ[e * QS + own * NF], qm);
            numericalFlux<DIM, NF>(a.ph, nv, nb <= -3 ? 0 : nb, qm, qp[c], pi);
#pragma unroll
            for (int k = 0; k < NF; ++k) out[k] = fs * pi[k];
            storeNode<NF>(&sB[bIndex(LIFT0 + fa, e)], out);
        }
    }
    __syncthreads();

    // ---- 3. matrix core: the wave's row tiles x column tiles (4 elements x 4 fields each) ----
    typename MT_::V acc[B::kMtW][B::kCtW][2];
#pragma unroll
    for (int m = 0; m < B::kMtW; ++m)
#pragma unroll
        for (int ct = 0; ct < B::kCtW; ++ct) acc[m][ct][0] = acc[m][ct][1] = typename MT_::V{T(0), T(0), T(0), T(0)};
#pragma unroll
    for (int ks = 0; ks < M::KS; ++ks) {
#pragma unroll
        for (int ct = 0; ct < B::kCtW; ++ct) {
            const T b = sB[(ks * 4 + ct0 + ct) * 64 + lane];
#pragma unroll
            for (int m = 0; m < B::kMtW; ++m)
                acc[m][ct][ks & 1] = MT_::mma(af[m * B::KSP + ks], b, acc[m][ct][ks & 1]);
        }
    }
    // 4a. results node-major: row = node, column lane%16 = 4 * element + field within the tile
#pragma unroll
    for (int m = 0; m < B::kMtW; ++m)
#pragma unroll
        for (int ct = 0; ct < B::kCtW; ++ct)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = 16 * (mt0 + m) + MT_::row(lane, r);
                const int col = lane & 15;
                if (row < NP) sR[(4 * (ct0 + ct) + (col >> 2)) * QS + row * NF + (col & 3)] = acc[m][ct][0][r] + acc[m][ct][1][r];
            }
    __syncthreads();

    // ---- 4b. RK update, node-major over this wave's 4 elements ----
    double res[5] = {0, 0, 0, 0, 0};
#pragma unroll
    for (int c = 0; c < NIT; ++c) {
        if (!uOk[c]) continue;
        const int it = lane + 64 * c;
        const int el = it / NP, i = it - el * NP;
        const int eb = 4 * wave + el;
        T rhs[NF], q0[NF];
        loadNode<NF>(&sR[eb * QS + i * NF], rhs);
        loadNode<NF>(&sQ[eb * QS + i * NF], q0);
        const int64_t base = uNode[c] * NF;
        const T dt = a.dt;
        if constexpr (MODE == kModeRHS) {
            T o[NF];
#pragma unroll
            for (int k = 0; k < NF; ++k) o[k] = dt * rhs[k];
            storeNode<NF>(a.qout + base, o);
        } else if constexpr (MODE == kModeFirst || MODE == kModeMid) {
            const T half = (MODE == kModeMid && a.stage == kStageRK3) ? T(1) : T(0.5);
            T accn[NF], qn[NF];
#pragma unroll
            for (int k = 0; k < NF; ++k) {
                const T kk = dt * rhs[k];
                T uv;
                if constexpr (MODE == kModeFirst) {
                    uv = q0[k];
                    accn[k] = kk;
                } else {
                    uv = pu[c][k];
                    accn[k] = pacc[c][k] + T(2) * kk;
                }
                qn[k] = uv + half * kk;
            }
            storeNodeNT<NF>(a.acc + base, accn);
            storeNode<NF>(a.qout + base, qn);
        } else {
            // final update: u += (k1 + 2k2 + 2k3 + k4)/6  (RK4)  or  u += k  (Euler)
            T unew[NF], lw[NF];
#pragma unroll
            for (int k = 0; k < NF; ++k) {
                const T kk = dt * rhs[k];
                const T delta = (MODE == kModeFinal) ? (pacc[c][k] + kk) / T(6) : kk;
                const T uv = (MODE == kModeFinal) ? pu[c][k] : q0[k];
                if constexpr (COMP) {
                    T s_, e_, t_, e2;
                    twoSum(plo[c][k], delta, t_, e2);
                    twoSum(uv, t_, s_, e_);
                    const T nlo = e_ + e2;
                    T h, l2;
                    twoSum(s_, nlo, h, l2);
                    unew[k] = h;
                    lw[k] = l2;
                } else {
                    unew[k] = uv + delta;
                    lw[k] = T(0);
                }
            }
            storeNodeNT<NF>(a.uout + base, unew);
            if constexpr (COMP) storeNodeNT<NF>(a.ulo + base, lw);
            if (!allFinite<NF>(unew)) a.stateBad[0] = 1;
            if (a.snap) {
                double sv[4], d[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    sv[k] = static_cast<double>(ps[c][k]);
                    d[k] = (static_cast<double>(unew[k]) + static_cast<double>(lw[k])) - sv[k];
                }
                const double ic2 = static_cast<double>(a.ph.ic2);
                const double u0 = static_cast<double>(unew[0]) + static_cast<double>(lw[0]);
                const double sp = sv[0] * ic2 - u0 * ic2;
                res[0] += d[0] * d[0];
                res[1] += sp * sp;
                res[2] += d[1] * d[1];
                res[3] += d[2] * d[2];
                res[4] += d[3] * d[3];
            }
        }
    }
    if constexpr (kFinal) {
        if (a.snap) blockResiduals<kBlock>(res, sRed, a.resPartial + static_cast<int64_t>(blockIdx.x) * 5);
    }
}

// ---------------------------------------------------------------- per-step kernels ----

struct SourceArgs {
    const int64_t* nodes;     // [nnodes] distinct element-node indices covered by a source (ascending)
    const int32_t* lastSrc;   // [nnodes] the last source (definition order) covering the node
    const int64_t* coverOff;  // [nnodes+1] into coverSrc
    const int32_t* coverSrc;  // the sources covering each node, in definition order
    int64_t nnodes;
    const double* params;     // [nsrc][5]: kind, amp, freq, phase, duration
    const double* tTable;     // [slots] step times
    const double* vTable;     // [slots][nsrc] table values (kind 1)
    const uint8_t* vFlag;     // [slots][nsrc] 1 = overwrite (kind 1)
    int nsrc;
    int nf;
    double ic2;               // 1/c0^2 (5-field variant: rho follows p)
};

// Value of source `src` at step `slot` (time t) and whether it overwrites (src/solver.cpp:455-483):
// monopole amp*sin(2 pi f t + phase) while t < duration; table sources (formula / file,
// evaluated on the host) when flagged.
__device__ inline bool sourceValue(const SourceArgs& s, int src, int64_t slot, double t, double& value) {
    const double* pr = s.params + src * 5;
    if (pr[0] == 0.0) {
        value = pr[1] * sin(2.0 * M_PI * pr[2] * t + pr[3]);
        return t < pr[4];
    }
    value = s.vTable[slot * s.nsrc + src];
    return s.vFlag[slot * s.nsrc + src] != 0;
}

// Hard-source overwrite of the pressure (and, with 5 fields, of rho = p/c0^2) at one node.
template <typename T, bool COMP>
__device__ inline void writeSource(const SourceArgs& s, int64_t node, double value, T* __restrict__ u,
                                   T* __restrict__ ulo) {
    const int64_t base = node * s.nf;
    const T hi = static_cast<T>(value);
    u[base] = hi;
    if constexpr (COMP) ulo[base] = static_cast<T>(value - static_cast<double>(hi));
    if (s.nf == 5) {
        const double rv = value * s.ic2;
        const T rh = static_cast<T>(rv);
        u[base + 4] = rh;
        if constexpr (COMP) ulo[base + 4] = static_cast<T>(rv - static_cast<double>(rh));
    }
}

// A thread's share of one round of the source table: nodes x0 + q * blockDim.x, q < kSrcU,
// loaded from clamped addresses before any use (the caller guarantees nnodes > 0).
constexpr int kSrcU = 2;
struct SourceRound {
    int64_t node[kSrcU];
    int32_t src[kSrcU];
    int64_t b[kSrcU];
    int64_t n[kSrcU];  // number of covering sources; 0 for a slot past the end
};

__device__ inline void loadSourceRound(const SourceArgs& s, int64_t x0, SourceRound& r) {
#pragma unroll
    for (int q = 0; q < kSrcU; ++q) {
        const int64_t x = x0 + static_cast<int64_t>(q) * blockDim.x;
        const bool ok = x < s.nnodes;
        const int64_t xc = ok ? x : 0;
        r.node[q] = s.nodes[xc];
        r.src[q] = s.lastSrc[xc];
        const int64_t b = s.coverOff[xc], e = s.coverOff[xc + 1];
        r.b[q] = b;
        r.n[q] = ok ? e - b : 0;
    }
}

// Every node takes the value of the last source covering it that is active at this step: the
// reference applies the sources in definition order, so a later active source overwrites an
// earlier one; a node whose sources are all inactive keeps its state.
__device__ inline void evalSourceRound(const SourceArgs& s, int64_t slot, double t, const SourceRound& r,
                                       bool (&act)[kSrcU], double (&v)[kSrcU]) {
#pragma unroll
    for (int q = 0; q < kSrcU; ++q) {
        v[q] = 0.0;
        act[q] = r.n[q] > 0 && sourceValue(s, r.src[q], slot, t, v[q]);
        for (int64_t k = r.b[q] + r.n[q] - 2; !act[q] && k >= r.b[q]; --k)
            act[q] = sourceValue(s, s.coverSrc[k], slot, t, v[q]);
    }
}

template <typename T, bool COMP>
__device__ inline void writeSourceRound(const SourceArgs& s, const SourceRound& r, const bool (&act)[kSrcU],
                                        const double (&v)[kSrcU], T* __restrict__ u, T* __restrict__ ulo) {
#pragma unroll
    for (int q = 0; q < kSrcU; ++q)
        if (act[q]) writeSource<T, COMP>(s, r.node[q], v[q], u, ulo);
}

// Source overwrite of step `slot` by one block.
template <typename T, bool COMP>
__device__ inline void applySources(const SourceArgs& s, int64_t slot, T* __restrict__ u, T* __restrict__ ulo) {
    if (s.nnodes == 0) return;
    const double t = s.tTable[slot];
    for (int64_t x0 = threadIdx.x; x0 < s.nnodes; x0 += static_cast<int64_t>(kSrcU) * blockDim.x) {
        SourceRound r;
        bool act[kSrcU];
        double v[kSrcU];
        loadSourceRound(s, x0, r);
        evalSourceRound(s, slot, t, r, act, v);
        writeSourceRound<T, COMP>(s, r, act, v, u, ulo);
    }
}

// Source overwrite for the step held in the step counter (first step of a batch).
template <typename T, bool COMP>
__global__ __launch_bounds__(256) void sourceKernel(SourceArgs s, const int64_t* __restrict__ counter,
                                                    T* __restrict__ u, T* __restrict__ ulo) {
    applySources<T, COMP>(s, *counter, u, ulo);
}

// Residual partial sums of many stage blocks folded into groups of kResGroup blocks (one block
// of 256 threads per group; thread t takes blocks t, t + 256, ... of its group in order, then
// rows of 16 lanes and the rows in order), so that the single-block end-of-step kernel reads one
// partial per group. Used when the stage grid has more than 4 kResGroup blocks.
constexpr int kResGroup = 4096;

[[maybe_unused]] static __global__ __launch_bounds__(256) void residualGroupKernel(const double* __restrict__ part, int64_t n,
                                                           double* __restrict__ out) {
    __shared__ double sRed[256 / 16][5];
    const int64_t g0 = static_cast<int64_t>(blockIdx.x) * kResGroup;
    const int64_t g1 = (g0 + kResGroup < n) ? g0 + kResGroup : n;
    constexpr int U = kResGroup / 256;
    double v[U][5];
#pragma unroll
    for (int q = 0; q < U; ++q) {
        const int64_t b = g0 + threadIdx.x + static_cast<int64_t>(q) * 256;
        const int64_t bc = b < g1 ? b : g0;
#pragma unroll
        for (int k = 0; k < 5; ++k) v[q][k] = part[bc * 5 + k];
    }
    double acc[5] = {0, 0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < U; ++q)
        if (g0 + threadIdx.x + static_cast<int64_t>(q) * 256 < g1)
#pragma unroll
            for (int k = 0; k < 5; ++k) acc[k] += v[q][k];
    blockResiduals<256>(acc, sRed, out + static_cast<int64_t>(blockIdx.x) * 5);
}

struct PostArgs {
    const double* resPartial;   // [nblocks][5] from the final stage
    int64_t nblocks;
    double invN;                // 1/(K*Np)
    const int64_t* obsIdx;      // [max(nseg,1)][256] element-node indices; observer node lists in
    const double* obsW;         //   segments of 256 entries, padded with weight 0 (IDW weights)
    const int32_t* obsSeg0;     // [nobs+1] first segment of each observer
    double* segPartial;         // [nseg][5]
    int nseg;
    int nobs;
    int nf;
    double* resRing;            // [slots][5]
    double* obsRing;            // [slots][nobs][4]
    int* nonfinite;             // [0]: first slot with a non-finite residual or state, or -1;
                                // [1]: set by a final stage that wrote a non-finite state value
    int64_t* stepCounter;
    const int64_t* batchSteps;  // number of steps in the current batch
    int computeResiduals;
    int withSources;            // 0: the next step's source overwrite runs in sourceStepKernel
};

// End-of-step work in one block of kPostBlock threads: residuals (src/solver.cpp:518-547;
// thread t sums the stage blocks t, t + kPostBlock, ... in that order, then a fixed-order tree),
// inverse-distance-weighted observers (src/solver.cpp:552-573; one wave per segment of 256
// node-list entries), then the source overwrite of the next step of the batch and the step
// counter advance. The loads that depend on no other load (counter, the first observer
// segment, the source table, observer ranges, residual partials) are issued first, then the
// ones that depend on them (state at observer nodes, step time and source parameters): with up
// to 4096 stage blocks and one observer segment per wave a step costs three memory round trips.
constexpr int kPostBlock = 1024;

template <typename T, bool COMP>
__global__ __launch_bounds__(kPostBlock) void postKernel(PostArgs p, SourceArgs s, T* __restrict__ u,
                                                         T* __restrict__ ulo) {
    constexpr int NW = kPostBlock / 64;
    constexpr int U = 4;
    __shared__ double red[NW][5];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

    // ---- round trip 1 ----
    const int64_t slot = *p.stepCounter;
    const int64_t nsteps = *p.batchSteps;
    const int stateBad = p.nonfinite[1];
    int64_t oIdx[4];
    double oW[4];
    {
        const int64_t x0 = static_cast<int64_t>(wave < p.nseg ? wave : 0) * 256;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            oIdx[q] = p.obsIdx[x0 + lane + 64 * q];
            oW[q] = p.obsW[x0 + lane + 64 * q];
        }
    }
    const int oc = tid < p.nobs ? tid : 0;
    const int sg0 = p.obsSeg0[oc], sg1 = p.obsSeg0[oc + 1];
    SourceRound sr;
    const bool haveSrc = p.withSources && s.nnodes > 0;
    if (haveSrc) loadSourceRound(s, tid, sr);
    // residual partial sums of this thread: stage blocks tid, tid + kPostBlock, ... in order,
    // U per round trip
    double acc[5] = {0, 0, 0, 0, 0};
    for (int64_t b0 = tid; b0 < p.nblocks; b0 += static_cast<int64_t>(U) * kPostBlock) {
        double rv[U][5];
#pragma unroll
        for (int q = 0; q < U; ++q) {
            const int64_t b = b0 + static_cast<int64_t>(q) * kPostBlock;
            const int64_t bc = b < p.nblocks ? b : p.nblocks - 1;
#pragma unroll
            for (int k = 0; k < 5; ++k) rv[q][k] = p.resPartial[bc * 5 + k];
        }
#pragma unroll
        for (int q = 0; q < U; ++q)
            if (b0 + static_cast<int64_t>(q) * kPostBlock < p.nblocks)
#pragma unroll
                for (int k = 0; k < 5; ++k) acc[k] += rv[q][k];
    }
    if (!p.computeResiduals)
#pragma unroll
        for (int k = 0; k < 5; ++k) acc[k] = 0.0;

    // ---- round trip 2 ----
    const bool srcStep = slot + 1 < nsteps;
    const int64_t slotNext = srcStep ? slot + 1 : slot;
    const double tNext = s.tTable[slotNext];
    double vq[4][4];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            vq[q][k] = static_cast<double>(u[oIdx[q] * p.nf + k]);
            if constexpr (COMP) vq[q][k] += static_cast<double>(ulo[oIdx[q] * p.nf + k]);
        }
    bool srcAct[kSrcU] = {};
    double srcVal[kSrcU] = {};
    if (haveSrc) evalSourceRound(s, slotNext, tNext, sr, srcAct, srcVal);

    // observer segments: one wave per segment, 4 entries per lane
    for (int sg = wave; sg < p.nseg; sg += NW) {
        if (sg != wave) {
            const int64_t x0 = static_cast<int64_t>(sg) * 256;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                oIdx[q] = p.obsIdx[x0 + lane + 64 * q];
                oW[q] = p.obsW[x0 + lane + 64 * q];
            }
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    vq[q][k] = static_cast<double>(u[oIdx[q] * p.nf + k]);
                    if constexpr (COMP) vq[q][k] += static_cast<double>(ulo[oIdx[q] * p.nf + k]);
                }
        }
        double ob[5] = {0, 0, 0, 0, 0};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
#pragma unroll
            for (int k = 0; k < 4; ++k) ob[k] += vq[q][k] * oW[q];
            ob[4] += oW[q];
        }
#pragma unroll
        for (int k = 0; k < 5; ++k)
            for (int off = 32; off > 0; off >>= 1) ob[k] += __shfl_down(ob[k], off, 64);
        if (lane == 0)
#pragma unroll
            for (int k = 0; k < 5; ++k) p.segPartial[sg * 5 + k] = ob[k];
    }
    // residuals: wave trees, then the waves in order
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        double v = acc[k];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if (lane == 0) red[wave][k] = v;
    }
    __syncthreads();  // also orders the segment partials (global) and every read of u above
    if (tid < 5) {
        double v = 0.0;
        for (int w = 0; w < NW; ++w) v += red[w][tid];
        v *= p.invN;
        p.resRing[slot * 5 + tid] = v;
        if (!isfinite(v)) atomicCAS(p.nonfinite, -1, static_cast<int>(slot));
    }
    if (tid == 0 && stateBad) atomicCAS(p.nonfinite, -1, static_cast<int>(slot));
    // observers: sum of the observer's segment partials in segment order
    for (int o = tid; o < p.nobs; o += kPostBlock) {
        const int s0 = o == tid ? sg0 : p.obsSeg0[o], s1 = o == tid ? sg1 : p.obsSeg0[o + 1];
        double ob[5] = {0, 0, 0, 0, 0};
        for (int sg = s0; sg < s1; ++sg)
            for (int k = 0; k < 5; ++k) ob[k] += p.segPartial[sg * 5 + k];
        for (int k = 0; k < 4; ++k) p.obsRing[(slot * p.nobs + o) * 4 + k] = ob[k] / ob[4];
    }
    // source overwrite of the next step of the batch
    if (srcStep && haveSrc) {
        writeSourceRound<T, COMP>(s, sr, srcAct, srcVal, u, ulo);
        for (int64_t x0 = tid + static_cast<int64_t>(kSrcU) * kPostBlock; x0 < s.nnodes;
             x0 += static_cast<int64_t>(kSrcU) * kPostBlock) {
            SourceRound r;
            loadSourceRound(s, x0, r);
            evalSourceRound(s, slotNext, tNext, r, srcAct, srcVal);
            writeSourceRound<T, COMP>(s, r, srcAct, srcVal, u, ulo);
        }
    }
    __syncthreads();
    if (tid == 0) *p.stepCounter = slot + 1;
}

[[maybe_unused]] static __global__ void setCounterKernel(int64_t* counter, int64_t* batchSteps, int64_t nsteps, int* flag,
                                                         int64_t* srcBase) {
    *counter = 0;
    *batchSteps = nsteps;
    flag[0] = -1;
    flag[1] = 0;
    *srcBase = 0;
}

// Source overwrite of step `*base + offset` of the batch (when that step exists) inside a captured
// sequence of steps, where the end-of-step kernel runs on a branch of its own: one round of kSrcU
// table entries per thread, as many 256-thread blocks as the table needs. `*base` is the number of
// steps of the batch that earlier sequences did (advanceKernel adds to it after each sequence).
template <typename T, bool COMP>
__global__ __launch_bounds__(256) void sourceStepKernel(SourceArgs s, const int64_t* __restrict__ base, int64_t offset,
                                                        const int64_t* __restrict__ batchSteps, T* __restrict__ u,
                                                        T* __restrict__ ulo) {
    const int64_t slot = *base + offset;
    const int64_t nsteps = *batchSteps;
    SourceRound r;
    loadSourceRound(s, static_cast<int64_t>(blockIdx.x) * kSrcU * blockDim.x + threadIdx.x, r);
    if (slot >= nsteps) return;
    bool act[kSrcU];
    double v[kSrcU];
    evalSourceRound(s, slot, s.tTable[slot], r, act, v);
    writeSourceRound<T, COMP>(s, r, act, v, u, ulo);
}

[[maybe_unused]] static __global__ void advanceKernel(int64_t* base, int64_t n) { *base += n; }

}  // namespace dgk
