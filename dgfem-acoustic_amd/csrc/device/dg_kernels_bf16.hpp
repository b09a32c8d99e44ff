// Split-BF16 matrix-core stage kernel (mixed precision: FP32 state, fluxes and accumulation).
//
// The two contractions of a stage are one product per group of 16 elements,
//   rhs = [Dw_1 | .. | Dw_DIM | -LIFT] . [G_1; ..; G_DIM; Fscale pi]
//   G_r = sum_x (dxi_r/dx_x) F_x(q)      (the flux Jacobian folded into the volume operand)
// (the nodal form of getElStiffVector + getElFlux + linEq, src/Mesh.cpp:734-751, :822-839,
// src/utils.cpp:406-412), computed on v_mfma_f32_16x16x32_bf16 with FP32 accumulation. FP32
// operands are carried as sums of BF16 terms: the operator is split on the host from its FP64
// values, the data on the device from its FP32 values (x = h + m + l, each term the nearest BF16
// of what the previous ones leave: unbiased rounding; a truncated last term measured a 2.4x
// larger error against FP64 after 300 RK4 steps at p = 5). NSPLIT = 3 keeps every product term
// down to 2^-16 relative (hh, hm, mh, hl, lh, mm: six matrix products), i.e. FP32 operand
// precision; NSPLIT = 2 keeps hh, hl, lh (three products) with operands carried to about 2^-17.
//
// Work layout: a persistent block of 4 waves takes contiguous groups of 16 elements (matrix
// column = element). The K dimension is cut in chunks of 8: volume chunks (direction r, nodes
// 8c..8c+7), then lift chunks (face nodes 8c..8c+7), four chunks per k-step (lane group l/16
// takes chunk 4 ks + l/16), volume and lift k-steps kept apart (zero chunks pad each part). Per
// group: the group's stage input is staged in LDS field-major with coalesced loads; the scaled
// numerical fluxes are computed one thread per face node (neighbour traces gathered from HBM)
// into LDS; then per k-step the four waves build the split data fragments of the k-step
// together into a double-buffered LDS slot and each wave multiplies the tiles it owns, its
// operator fragments held in registers for the whole kernel; finally the accumulators are
// transposed through LDS and the RK register update runs with coalesced accesses, one thread
// per element node. Each wave owns one row tile and a subset of the fields (4 / MT waves per
// row tile); its operator fragments stay in registers when they fit in 64 of them (orders
// <= 3), else they are read one k-step ahead from L2 (the whole operator is at most 108 KB).
#pragma once
#include "dg_kernels.hpp"

namespace dgk {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));

// Row stride (floats) of per-element LDS blocks read by 16 lanes at once as 16-byte vectors:
// == 4 (mod 8), so that 8 consecutive elements start in distinct groups of 4 banks.
__host__ __device__ constexpr int bf16ElemStride(int n) { return n + ((4 - n % 8) + 8) % 8; }

template <int DIM, int P, int NF>
struct Bf16Shape {
    using S = Sizes<DIM, P>;
    static constexpr int NP = S::NP, NFT = S::NFT;
    static constexpr int MT = (NP + 15) / 16;                // row tiles (element nodes)
    static constexpr int CPR = (NP + 7) / 8;                 // volume chunks per direction
    static constexpr int KSV = (DIM * CPR + 3) / 4;          // volume k-steps
    static constexpr int NCL = (NFT + 7) / 8;                // lift chunks
    static constexpr int KSL = (NCL + 3) / 4;                // lift k-steps
    static constexpr int KS = KSV + KSL;
    static constexpr int kWaves = 4;
    static constexpr int kBlockSize = 64 * kWaves;
    static constexpr int kGroup = 16;                         // elements per group
    static constexpr int QS = 8 * CPR;                        // staged state: row (element, field)
    static constexpr int QE = bf16ElemStride(NF * QS);        //   element stride
    static constexpr int PR = 8 * NCL;                        // staged fluxes: row (element, field)
    static constexpr int PE = bf16ElemStride(NF * PR);        //   element stride
    // tiles (row tile, field) of the waves: wave w takes row tile w % MT and the fields f with
    // f % WPT == w / MT (WPT waves per row tile; with 3 row tiles the fourth wave takes none)
    static constexpr int WPT = (4 % MT) == 0 ? 4 / MT : 1;
    static constexpr int FPW = (NF + WPT - 1) / WPT;         // fields per wave (at most)
    // the data fragments of all k-steps fit in 64 KB of LDS: built at once per group (one
    // barrier between building and multiplying), else per k-step into two slots
    static constexpr bool kAllFrag = KS * NF * 3 * 1024 <= 64 * 1024;
    static constexpr int NSLOT = kAllFrag ? KS : 2;
};

// Operator fragment words of one (DIM, P, NSPLIT): [MT][KS][NSPLIT][64 lanes] of 16 bytes.
template <int DIM, int P, int NSPLIT>
constexpr int bf16FragWords() {
    return Bf16Shape<DIM, P, 4>::MT * Bf16Shape<DIM, P, 4>::KS * NSPLIT * 64;
}

// LDS bytes of the kernel's block.
template <int DIM, int P, int NF, int NSPLIT>
constexpr int bf16LdsBytes() {
    using B = Bf16Shape<DIM, P, NF>;
    using L = OpLayout<DIM, P>;
    constexpr int NFC = Sizes<DIM, P>::NFC;
    return 4 * B::kGroup * (B::QE + B::PE) + B::NSLOT * NF * NSPLIT * 64 * 16 + 4 * L::kITotal +
           B::kGroup * NFC * (4 + 4 + 16) + 4 * B::kGroup * DIM * DIM + 8 * 5 * (B::kBlockSize / 16);
}

// Blocks per CU: what the LDS allows, at most 2 (2 waves per SIMD leave 256 registers a wave).
template <int DIM, int P, int NF, int NSPLIT>
constexpr int bf16BlocksPerCu() {
    constexpr int b = (160 * 1024) / bf16LdsBytes<DIM, P, NF, NSPLIT>();
    return b > 2 ? 2 : b;
}

// Built for (DIM, P, NF, NSPLIT): row tiles dividing the four waves or a wave per row tile, and
// at least one block per CU.
template <int DIM, int P, int NF, int NSPLIT>
constexpr bool bf16Supported() {
    return Bf16Shape<DIM, P, NF>::MT <= 4 && bf16BlocksPerCu<DIM, P, NF, NSPLIT>() >= 1;
}

__device__ inline uint32_t f2u(float x) { return __builtin_bit_cast(uint32_t, x); }
__device__ inline float u2f(uint32_t x) { return __builtin_bit_cast(float, x); }

// Nearest BF16 pair (round to nearest even, v_cvt_pk_bf16_f32), x low / y high.
__device__ inline uint32_t bf16Pair(f32x2_t v) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2_t));
}

// The pair's values as floats.
__device__ inline f32x2_t unpackPair(uint32_t w) { return f32x2_t{u2f(w << 16), u2f(w & 0xffff0000u)}; }

// BF16 terms of 8 values: frag[s] = term s of v[0..7], each term the nearest BF16 of what the
// previous terms leave (the remainders x - h and x - h - m are exact in FP32, so NSPLIT terms
// carry x to about 2^-(8 NSPLIT + 1) with unbiased rounding).
template <int NSPLIT>
__device__ inline void splitFrag(const float (&v)[8], u32x4_t (&frag)[NSPLIT]) {
    f32x2_t v0 = {v[0], v[1]}, v1 = {v[2], v[3]}, v2 = {v[4], v[5]}, v3 = {v[6], v[7]};
#pragma unroll
    for (int s = 0; s < NSPLIT; ++s) {
        const uint32_t w0 = bf16Pair(v0), w1 = bf16Pair(v1), w2 = bf16Pair(v2), w3 = bf16Pair(v3);
        frag[s] = u32x4_t{w0, w1, w2, w3};
        if (s + 1 < NSPLIT) {
            v0 = v0 - unpackPair(w0);
            v1 = v1 - unpackPair(w1);
            v2 = v2 - unpackPair(w2);
            v3 = v3 - unpackPair(w3);
        }
    }
}

__device__ inline f32x4_t mfmaBf16(u32x4_t a, u32x4_t b, f32x4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a), __builtin_bit_cast(bf16x8_t, b), c,
                                                   0, 0, 0);
}

// acc += A . B over the split terms: (h,h) (h,m) (m,h) (h,l) (l,h) (m,m) for three terms,
// (h,h) (h,l) (l,h) for two. The matrix instruction's FP32 accumulation is unbiased but not
// exactly rounded (measured on gfx950: mean |error| 0.47 ulp of the result, 0.25 for one RNE
// rounding), so the small cross products are summed into a fresh accumulator and only the
// (h,h) product and one RNE add touch the running sum: two roundings at its magnitude per
// k-step instead of one per product (with one per product, bf16x6 drifted 2x more than the
// FP32 VALU kernel over 300 RK4 steps; with two, less).
template <int NSPLIT>
__device__ inline f32x4_t mfmaSplit(const u32x4_t (&A)[NSPLIT], const u32x4_t (&B)[NSPLIT], f32x4_t c) {
    const f32x4_t z = {0.0f, 0.0f, 0.0f, 0.0f};
    f32x4_t corr;
    if constexpr (NSPLIT == 2) {
        corr = mfmaBf16(A[0], B[1], z);
        corr = mfmaBf16(A[1], B[0], corr);
    } else {
        corr = mfmaBf16(A[2], B[0], z);
        corr = mfmaBf16(A[0], B[2], corr);
        corr = mfmaBf16(A[1], B[1], corr);
        corr = mfmaBf16(A[1], B[0], corr);
        corr = mfmaBf16(A[0], B[1], corr);
    }
    c = mfmaBf16(A[0], B[0], c);
    return c + corr;
}

// Fields of field group fg (0..3): fg, and for 5 fields group 0 also rho (it shares J_r . v
// with p). The data fragments of a field are built by the threads of its group and, for the
// orders whose waves own fields, multiplied by the wave of that group.
template <int NF, int FG>
struct FieldGroup {
    static constexpr int kCount = (NF == 5 && FG == 0) ? 2 : 1;
    __host__ __device__ static constexpr int field(int j) { return j == 0 ? FG : 4; }
};

// Data fragments of k-step ks for the fields of group FG, lane slot bls (element column
// bls % 16, chunk 4 ks + bls / 16), written to slot[field][term][bls]. Volume chunks: G_r at 8
// nodes from the staged state (field-major rows); lift chunks: 8 staged scaled fluxes.
template <int DIM, int P, int NF, int NSPLIT, int FG>
__device__ inline void buildFragments(int ks, int bls, const float* __restrict__ sQ, const float* __restrict__ sP,
                                      const float* __restrict__ sJ, const Physics<float>& ph,
                                      u32x4_t (*slot)[NSPLIT][64]) {
    using B = Bf16Shape<DIM, P, NF>;
    using FGt = FieldGroup<NF, FG>;
    constexpr int NP = B::NP, CPR = B::CPR, KSV = B::KSV, NCL = B::NCL, QS = B::QS, QE = B::QE, PR = B::PR,
                  PE = B::PE, JJ = DIM * DIM;
    const int bcol = bls & 15, bg = bls >> 4;
    auto row8 = [](const float* p, float (&o)[8]) {
        const float4 lo = *reinterpret_cast<const float4*>(p);
        const float4 hi = *reinterpret_cast<const float4*>(p + 4);
        o[0] = lo.x; o[1] = lo.y; o[2] = lo.z; o[3] = lo.w;
        o[4] = hi.x; o[5] = hi.y; o[6] = hi.z; o[7] = hi.w;
    };
    if (ks < KSV) {
        const int v = 4 * ks + bg;
        const bool ok = v < DIM * CPR;
        const int r = ok ? v / CPR : 0;
        const int c = ok ? v - r * CPR : 0;
        float jr[DIM], w = 0.0f;
#pragma unroll
        for (int x = 0; x < DIM; ++x) {
            jr[x] = sJ[bcol * JJ + r * DIM + x];
            w += jr[x] * ph.v0[x];
        }
        const float* qrow = &sQ[bcol * QE + 8 * c];
        // rows read: p; the velocities for p (and rho); v_FG for the velocity groups
        float q[NF][8];
        row8(qrow, q[0]);
        if constexpr (FG == 0) {
#pragma unroll
            for (int x = 0; x < DIM; ++x) row8(qrow + (1 + x) * QS, q[1 + x]);
            if constexpr (NF == 5) row8(qrow + 4 * QS, q[4]);
        } else {
            row8(qrow + FG * QS, q[FG]);
        }
#pragma unroll
        for (int j = 0; j < FGt::kCount; ++j) {
            const int f = FGt::field(j);
            float vals[8];
#pragma unroll
            for (int jj = 0; jj < 8; ++jj) {
                float gv;
                if (FG == 0) {
                    float jv = 0.0f;
#pragma unroll
                    for (int x = 0; x < DIM; ++x) jv += jr[x] * q[1 + x][jj];
                    gv = (j == 0) ? w * q[0][jj] + ph.rc2 * jv : w * q[NF - 1][jj] + ph.rho0 * jv;
                } else if (FG - 1 < DIM) {
                    gv = w * q[FG][jj] + jr[FG - 1 < DIM ? FG - 1 : 0] * ph.irho0 * q[0][jj];
                } else {
                    gv = w * q[FG][jj];
                }
                vals[jj] = (ok && 8 * c + jj < NP) ? gv : 0.0f;
            }
            u32x4_t fr[NSPLIT];
            splitFrag<NSPLIT>(vals, fr);
#pragma unroll
            for (int s = 0; s < NSPLIT; ++s) slot[f][s][bls] = fr[s];
        }
    } else {
        const int c = 4 * (ks - KSV) + bg;
        const bool ok = c < NCL;
        const float* prow = &sP[bcol * PE + 8 * (ok ? c : 0)];
#pragma unroll
        for (int j = 0; j < FGt::kCount; ++j) {
            const int f = FGt::field(j);
            float vals[8];
            row8(prow + f * PR, vals);
            if (!ok)
#pragma unroll
                for (int jj = 0; jj < 8; ++jj) vals[jj] = 0.0f;
            u32x4_t fr[NSPLIT];
            splitFrag<NSPLIT>(vals, fr);
#pragma unroll
            for (int s = 0; s < NSPLIT; ++s) slot[f][s][bls] = fr[s];
        }
    }
}

template <int DIM, int P, int NF, int NSPLIT, int MODE>
__global__ __launch_bounds__(256, (bf16BlocksPerCu<DIM, P, NF, NSPLIT>())) void stageKernelBf16(
    StageArgs<float> a, const u32x4_t* __restrict__ afrag, int64_t nGroups) {
    using S = Sizes<DIM, P>;
    using L = OpLayout<DIM, P>;
    using B = Bf16Shape<DIM, P, NF>;
    constexpr int NP = S::NP, NFP = S::NFP, NFC = S::NFC, NFT = S::NFT;
    constexpr int MT = B::MT, KSV = B::KSV, KS = B::KS, G = B::kGroup;
    constexpr int QS = B::QS, QE = B::QE, PR = B::PR, PE = B::PE, BS = B::kBlockSize;
    constexpr int JJ = DIM * DIM;
    constexpr int WPT = B::WPT, FPW = B::FPW;
    // operator fragments of the wave's row tile: held in registers for the whole kernel, or, when
    // they would take more than 64 registers, read per k-step (from L2; one k-step ahead)
    constexpr bool kAStream = KS * NSPLIT * 4 > 64;
    constexpr int AREG = kAStream ? 1 : KS;
    constexpr bool kFinal = MODE == kModeFinal || MODE == kModeEuler;
    constexpr int kFluxIt = (G * NFT + BS - 1) / BS;            // face nodes per thread
    constexpr int kNodeIt = (G * NP + BS - 1) / BS;             // element nodes per thread
    constexpr int kGroupWords = G * NP * NF;                    // floats of a group's state
    constexpr int kStIt = (kGroupWords / 4 + BS - 1) / BS;      // staged 16-byte words per thread
    constexpr int kTbIt = (G * NFC + BS - 1) / BS;              // face rows per thread
    constexpr int kJIt = (G * JJ + BS - 1) / BS;                // geometry words per thread
    // orders of at most two row tiles issue a group's neighbour traces, RK operands and the next
    // group's staging ahead of its volume k-steps (the registers the larger orders lack)
    constexpr bool kPipe = MT <= 2;
    constexpr bool kAllFrag = B::kAllFrag && !kAStream;

    __shared__ __align__(16) float sQ[G * QE];                  // stage input, later the rhs
    __shared__ __align__(16) float sP[G * PE];                  // Fscale * numerical flux
    __shared__ u32x4_t sB[B::NSLOT][NF][NSPLIT][64];            // data fragments (per k-step or all)
    __shared__ int32_t sI[L::kITotal];
    __shared__ int32_t sNb[G * NFC];
    __shared__ int32_t sCd[G * NFC];
    __shared__ __align__(16) float sFg[G * NFC][4];
    __shared__ float sJ[G * JJ];
    __shared__ double sRed[BS / 16][5];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const Physics<float> ph = a.ph;
    const float dt = a.dt;

    for (int x = tid; x < L::kITotal; x += BS) sI[x] = a.iops[x];

    // this wave's row tile and field subset
    const bool mfmaWave = wave < MT * WPT;
    const int wit = mfmaWave ? wave % MT : 0;
    const int wsub = wave / MT;
    const u32x4_t* aw = afrag + static_cast<int64_t>(wit) * KS * NSPLIT * 64 + lane;  // [ks][s] at stride 64
    u32x4_t afr[AREG][NSPLIT];
    if constexpr (!kAStream) {
#pragma unroll
        for (int ks = 0; ks < KS; ++ks)
#pragma unroll
            for (int s = 0; s < NSPLIT; ++s) afr[ks][s] = aw[(ks * NSPLIT + s) * 64];
    }

    // contiguous range of groups per block; blocks of one XCD take adjacent ranges
    const int64_t nb = gridDim.x;
    const int64_t bi = xcdRemap(blockIdx.x, nb);
    const int64_t g0 = bi * nGroups / nb, g1 = (bi + 1) * nGroups / nb;
    double res[5] = {0, 0, 0, 0, 0};
    const int bls = tid & 63, bfg = tid >> 6;  // builder role: lane slot, field group

    // ---- staging of a group: its stage input (field-major per element), face data, geometry;
    //      loads into registers first, so that they can be issued a group ahead ----
    struct Staged {
        float4 q[kStIt];
        int nb[kTbIt], cd[kTbIt];
        float4 fg[kTbIt];
        float j[kJIt];
    };
    auto stageLoad = [&](int64_t grp, Staged& st) __attribute__((always_inline)) {
        const int64_t e0 = grp * G;
        const int nE = static_cast<int>((a.K - e0) < G ? (a.K - e0) : G);
        const float* src = a.qin + e0 * NP * NF;
        const int nw = nE * NP * NF, n4 = nw / 4;
#pragma unroll
        for (int c = 0; c < kStIt; ++c) {
            const int x = tid + c * BS;
            float4 v = {0.0f, 0.0f, 0.0f, 0.0f};
            if (x < n4) {
                v = reinterpret_cast<const float4*>(src)[x];
            } else if (x < kGroupWords / 4) {
                if (4 * x + 0 < nw) v.x = src[4 * x + 0];
                if (4 * x + 1 < nw) v.y = src[4 * x + 1];
                if (4 * x + 2 < nw) v.z = src[4 * x + 2];
                if (4 * x + 3 < nw) v.w = src[4 * x + 3];
            }
            st.q[c] = v;
        }
#pragma unroll
        for (int c = 0; c < kTbIt; ++c) {
            const int x = tid + c * BS;
            const int xc = x < G * NFC ? x : 0;
            const int el = xc / NFC;
            const int64_t ef = el < nE ? e0 * NFC + xc : e0 * NFC + xc % NFC;  // past the end: element e0's
            st.nb[c] = el < nE ? a.nbr[ef] : -1;
            st.cd[c] = a.code[ef];
            st.fg[c] = reinterpret_cast<const float4*>(a.fgeo)[ef];
            if (el >= nE) st.fg[c].w = 0.0f;  // no flux on faces of elements past the end
        }
#pragma unroll
        for (int c = 0; c < kJIt; ++c) {
            const int x = tid + c * BS;
            const int xc = x < G * JJ ? x : 0;
            st.j[c] = a.geo[xc < nE * JJ ? e0 * JJ + xc : e0 * JJ + xc % JJ];
        }
    };
    auto stageStore = [&](const Staged& st) __attribute__((always_inline)) {
#pragma unroll
        for (int c = 0; c < kStIt; ++c) {
            const int x = tid + c * BS;
            if (x >= kGroupWords / 4) continue;
            const float v[4] = {st.q[c].x, st.q[c].y, st.q[c].z, st.q[c].w};
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int y = 4 * x + k;
                const int el = y / (NP * NF), rem = y - el * (NP * NF);
                const int node = rem / NF, f = rem - node * NF;
                sQ[el * QE + f * QS + node] = v[k];
            }
        }
#pragma unroll
        for (int c = 0; c < kTbIt; ++c) {
            const int x = tid + c * BS;
            if (x >= G * NFC) continue;
            sNb[x] = st.nb[c];
            sCd[x] = st.cd[c];
            sFg[x][0] = st.fg[c].x;
            sFg[x][1] = st.fg[c].y;
            sFg[x][2] = st.fg[c].z;
            sFg[x][3] = st.fg[c].w;
        }
#pragma unroll
        for (int c = 0; c < kJIt; ++c) {
            const int x = tid + c * BS;
            if (x < G * JJ) sJ[x] = st.j[c];
        }
    };

    // ---- neighbour traces of the group's face nodes (one thread per element face node) ----
    auto fluxGather = [&](int nE, const float* __restrict__ qgrp, float (&qp)[kFluxIt][NF]) __attribute__((always_inline)) {
#pragma unroll
        for (int c = 0; c < kFluxIt; ++c) {
            const int it = tid + c * BS;
            const float* src = qgrp;  // a valid address for idle items and boundary faces
            if (it < nE * NFT) {
                const int el = it / NFT, fa = it - el * NFT;
                const int f = fa / NFP, an = fa - f * NFP;
                const int nbv = sNb[el * NFC + f];
                const int cd = sCd[el * NFC + f];
                const int nn = sI[L::kMask + (cd & 3) * NFP + sI[L::kPerm + (cd >> 2) * NFP + an]];
                if (nbv >= 0) src = a.qin + (static_cast<int64_t>(nbv) * NP + nn) * NF;
                else if (nbv <= -3) src = a.ghost + (static_cast<int64_t>(-3 - nbv) * NFP + an) * NF;
            }
            loadNode<NF>(src, qp[c]);
        }
    };
    // scaled numerical fluxes -> sP (and zeros in the lift chunks' rows past the last face node)
    auto fluxCommit = [&](int nE, const float (&qp)[kFluxIt][NF]) __attribute__((always_inline)) {
#pragma unroll
        for (int c = 0; c < kFluxIt; ++c) {
            const int it = tid + c * BS;
            if (it >= G * NFT) continue;
            const int el = it / NFT, fa = it - el * NFT;
            float out[NF];
#pragma unroll
            for (int k = 0; k < NF; ++k) out[k] = 0.0f;
            if (el < nE) {
                const int f = fa / NFP, an = fa - f * NFP;
                const int ef = el * NFC + f;
                const int own = sI[L::kMask + f * NFP + an];
                float qm[NF], pi[NF];
#pragma unroll
                for (int k = 0; k < NF; ++k) qm[k] = sQ[el * QE + k * QS + own];
                const float nv[3] = {sFg[ef][0], sFg[ef][1], sFg[ef][2]};
                const int nbv = sNb[ef];
                numericalFlux<DIM, NF>(ph, nv, nbv <= -3 ? 0 : nbv, qm, qp[c], pi);
#pragma unroll
                for (int k = 0; k < NF; ++k) out[k] = sFg[ef][3] * pi[k];
            }
#pragma unroll
            for (int k = 0; k < NF; ++k) sP[el * PE + k * PR + fa] = out[k];
        }
        if constexpr (PR > NFT) {
            for (int x = tid; x < G * NF * (PR - NFT); x += BS) {
                const int row = x / (PR - NFT), col = x - row * (PR - NFT);
                const int el = row / NF, k = row - el * NF;
                sP[el * PE + k * PR + NFT + col] = 0.0f;
            }
        }
    };

    // ---- data fragments of k-step ks -> slot buf (every thread: one lane slot, its field group) ----
    auto build = [&](int ks, int buf) __attribute__((always_inline)) {
        switch (bfg) {
            case 0: buildFragments<DIM, P, NF, NSPLIT, 0>(ks, bls, sQ, sP, sJ, ph, sB[buf]); break;
            case 1: buildFragments<DIM, P, NF, NSPLIT, 1>(ks, bls, sQ, sP, sJ, ph, sB[buf]); break;
            case 2: buildFragments<DIM, P, NF, NSPLIT, 2>(ks, bls, sQ, sP, sJ, ph, sB[buf]); break;
            default: buildFragments<DIM, P, NF, NSPLIT, 3>(ks, bls, sQ, sP, sJ, ph, sB[buf]); break;
        }
    };
    // k-steps [k0, k1) (slot k0 & 1 built and visible): build the next inside the range while
    // multiplying the current one; operator fragments from registers or one k-step ahead from L2
    auto ksteps = [&](int k0, int k1, f32x4_t (&acc)[FPW]) __attribute__((always_inline)) {
        u32x4_t anext[NSPLIT];
        if constexpr (kAStream) {
#pragma unroll
            for (int s = 0; s < NSPLIT; ++s) anext[s] = aw[(k0 * NSPLIT + s) * 64];
        }
    constexpr int kUnrollK = kAStream ? 1 : KS;
#pragma unroll kUnrollK
        for (int ks = k0; ks < k1; ++ks) {
            u32x4_t acur[NSPLIT];
            if constexpr (kAStream) {
#pragma unroll
                for (int s = 0; s < NSPLIT; ++s) acur[s] = anext[s];
                if (ks + 1 < k1)
#pragma unroll
                    for (int s = 0; s < NSPLIT; ++s) anext[s] = aw[((ks + 1) * NSPLIT + s) * 64];
            } else {
#pragma unroll
                for (int s = 0; s < NSPLIT; ++s) acur[s] = afr[ks][s];
            }
            if (ks + 1 < k1) build(ks + 1, (ks + 1) & 1);
            if (mfmaWave) {
#pragma unroll
                for (int j = 0; j < FPW; ++j) {
                    const int f = wsub + j * WPT;
                    if (f >= NF) continue;
                    u32x4_t bfr[NSPLIT];
#pragma unroll
                    for (int s = 0; s < NSPLIT; ++s) bfr[s] = sB[ks & 1][f][s][lane];
                    acc[j] = mfmaSplit<NSPLIT>(acur, bfr, acc[j]);
                }
            }
            __syncthreads();
        }
    };

    // products of the k-steps [k0, k1) whose fragments are all built (slot = k-step; operator
    // fragments in registers)
    auto multiply = [&](int k0, int k1, f32x4_t (&acc)[FPW]) __attribute__((always_inline)) {
        if (!mfmaWave) return;
#pragma unroll
        for (int ks = k0; ks < k1; ++ks)
#pragma unroll
            for (int j = 0; j < FPW; ++j) {
                const int f = wsub + j * WPT;
                if (f >= NF) continue;
                u32x4_t bfr[NSPLIT];
#pragma unroll
                for (int s = 0; s < NSPLIT; ++s) bfr[s] = sB[kAllFrag ? ks : 0][f][s][lane];
                acc[j] = mfmaSplit<NSPLIT>(afr[kAStream ? 0 : ks], bfr, acc[j]);
            }
    };

    if (g0 < g1) {
        Staged st;
        stageLoad(g0, st);
        stageStore(st);
    }
    __syncthreads();

    for (int64_t grp = g0; grp < g1; ++grp) {
        const int64_t e0 = grp * G;
        const int nE = static_cast<int>((a.K - e0) < G ? (a.K - e0) : G);
        const float* qgrp = a.qin + e0 * NP * NF;

        // pipelined orders: the group's neighbour traces, its RK operands and the next group's
        // staging are issued now and arrive during the volume k-steps
        float qp[kFluxIt][NF];
        float pu[kPipe ? kNodeIt : 1][NF], pacc[kPipe ? kNodeIt : 1][NF];
        Staged nx;
        const bool more = grp + 1 < g1;
        if constexpr (kPipe) {
            fluxGather(nE, qgrp, qp);
            if constexpr (MODE == kModeMid || MODE == kModeFinal) {
#pragma unroll
                for (int c = 0; c < kNodeIt; ++c) {
                    const int x = tid + c * BS;
                    const int64_t base = (e0 * NP + (x < nE * NP ? x : 0)) * NF;
                    loadNodeNT<NF>(a.u + base, pu[c]);
                    loadNodeNT<NF>(a.acc + base, pacc[c]);
                }
            }
            if (more) stageLoad(grp + 1, nx);
        }

        f32x4_t acc[FPW];
#pragma unroll
        for (int j = 0; j < FPW; ++j) acc[j] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};

        if constexpr (kAllFrag) {
            // every volume fragment and the fluxes; then every lift fragment beside the volume
            // products; then the lift products
#pragma unroll
            for (int ks = 0; ks < KSV; ++ks) build(ks, ks);
            if constexpr (!kPipe) fluxGather(nE, qgrp, qp);
            fluxCommit(nE, qp);
            __syncthreads();
#pragma unroll
            for (int ks = KSV; ks < KS; ++ks) build(ks, ks);
            multiply(0, KSV, acc);
            __syncthreads();
            multiply(KSV, KS, acc);
        } else {
            // volume k-steps (the staged state)
            build(0, 0);
            __syncthreads();
            ksteps(0, KSV, acc);
            // numerical fluxes, then the lift k-steps
            if constexpr (!kPipe) fluxGather(nE, qgrp, qp);
            fluxCommit(nE, qp);
            __syncthreads();
            build(KSV, KSV & 1);
            __syncthreads();
            ksteps(KSV, KS, acc);
        }

        // ---- right-hand sides to LDS (field-major per element; the stage input is no longer read) ----
        if (mfmaWave) {
            const int col = lane & 15, g4 = lane >> 4;
            const int i0 = 16 * wit + 4 * g4;
#pragma unroll
            for (int j = 0; j < FPW; ++j) {
                const int f = wsub + j * WPT;
                if (f >= NF) continue;
                float* dst = &sQ[col * QE + f * QS + i0];
                if (i0 + 3 < NP) {
                    *reinterpret_cast<float4*>(dst) = float4{acc[j][0], acc[j][1], acc[j][2], acc[j][3]};
                } else {
#pragma unroll
                    for (int rr = 0; rr < 4; ++rr)
                        if (i0 + rr < NP) dst[rr] = acc[j][rr];
                }
            }
        }
        __syncthreads();

        // ---- register update, one thread per element node (coalesced) ----
#pragma unroll 2
        for (int c = 0; c < kNodeIt; ++c) {
            const int x = tid + c * BS;
            if (x >= nE * NP) continue;
            const int el = x / NP, i = x - el * NP;
            const int64_t base = (e0 * NP + x) * NF;
            float rhs[NF];
#pragma unroll
            for (int f = 0; f < NF; ++f) rhs[f] = sQ[el * QE + f * QS + i];
            if constexpr (MODE == kModeRHS) {
                float o[NF];
#pragma unroll
                for (int k = 0; k < NF; ++k) o[k] = dt * rhs[k];
                storeNode<NF>(a.qout + base, o);
            } else if constexpr (MODE == kModeFirst || MODE == kModeMid) {
                const float half = (MODE == kModeMid && a.stage == kStageRK3) ? 1.0f : 0.5f;
                float uv[NF], pa[NF], accn[NF], qn[NF];
                if constexpr (MODE == kModeFirst) {
                    loadNode<NF>(a.qin + base, uv);
                } else if constexpr (kPipe) {
#pragma unroll
                    for (int k = 0; k < NF; ++k) {
                        uv[k] = pu[kPipe ? c : 0][k];
                        pa[k] = pacc[kPipe ? c : 0][k];
                    }
                } else {
                    loadNodeNT<NF>(a.u + base, uv);
                    loadNodeNT<NF>(a.acc + base, pa);
                }
#pragma unroll
                for (int k = 0; k < NF; ++k) {
                    const float kk = dt * rhs[k];
                    accn[k] = (MODE == kModeFirst) ? kk : pa[k] + 2.0f * kk;
                    qn[k] = uv[k] + half * kk;
                }
                storeNodeNT<NF>(a.acc + base, accn);
                storeNode<NF>(a.qout + base, qn);
            } else {
                float uv[NF], pa[NF], plo[NF], ps[NF], unew[NF], lnew[NF];
                if constexpr (MODE == kModeFinal) {
                    if constexpr (kPipe) {
#pragma unroll
                        for (int k = 0; k < NF; ++k) {
                            uv[k] = pu[kPipe ? c : 0][k];
                            pa[k] = pacc[kPipe ? c : 0][k];
                        }
                    } else {
                        loadNodeNT<NF>(a.u + base, uv);
                        loadNodeNT<NF>(a.acc + base, pa);
                    }
                } else {
                    loadNode<NF>(a.qin + base, uv);
                }
                loadNodeNT<NF>(a.ulo + base, plo);
                if (a.snap) loadNodeNT<NF>(a.snap + base, ps);
#pragma unroll
                for (int k = 0; k < NF; ++k) {
                    const float kk = dt * rhs[k];
                    const float delta = (MODE == kModeFinal) ? (pa[k] + kk) / 6.0f : kk;
                    float s_, e_, t_, e2;
                    twoSum(plo[k], delta, t_, e2);
                    twoSum(uv[k], t_, s_, e_);
                    twoSum(s_, e_ + e2, unew[k], lnew[k]);
                }
                storeNodeNT<NF>(a.uout + base, unew);
                storeNodeNT<NF>(a.ulo + base, lnew);
                if (!allFinite<NF>(unew)) a.stateBad[0] = 1;
                if (a.snap) {
                    double d[4];
#pragma unroll
                    for (int k = 0; k < 4; ++k)
                        d[k] = (static_cast<double>(unew[k]) + static_cast<double>(lnew[k])) - static_cast<double>(ps[k]);
                    const double ic2 = static_cast<double>(ph.ic2);
                    const double u0 = static_cast<double>(unew[0]) + static_cast<double>(lnew[0]);
                    const double sp = static_cast<double>(ps[0]) * ic2 - u0 * ic2;
                    res[0] += d[0] * d[0];
                    res[1] += sp * sp;
                    res[2] += d[1] * d[1];
                    res[3] += d[2] * d[2];
                    res[4] += d[3] * d[3];
                }
            }
        }
        __syncthreads();  // the next group's staging overwrites the LDS tables

        // the next group's staging: from the registers (pipelined orders) or loaded now
        if (more) {
            if constexpr (!kPipe) stageLoad(grp + 1, nx);
            stageStore(nx);
        }
        __syncthreads();
    }
    if constexpr (kFinal) {
        if (a.snap) blockResiduals<BS>(res, sRed, a.resPartial + static_cast<int64_t>(blockIdx.x) * 5);
    }
}

}  // namespace dgk

namespace dgk {

// Launch of the split-BF16 stage kernel (defined in dg_stage_bf16.hip, where its instances are
// built): false when no kernel is built for (dim, p, nf, nsplit).
bool launchStageBf16(int dim, int p, int nf, int nsplit, int mode, const StageArgs<float>& a, const void* afrag,
                     int64_t nGroups, unsigned grid, hipStream_t stream);
// Whether a kernel is built for (dim, p, nf, nsplit).
bool stageBf16Built(int dim, int p, int nf, int nsplit);
// Blocks per CU of the kernel for (dim, p, nf, nsplit) (0 when not built).
int stageBf16BlocksPerCu(int dim, int p, int nf, int nsplit);

}  // namespace dgk
