// Split-BF16 matrix-core stage kernel (mixed precision: FP32 state, fluxes and accumulation).
//
// The two contractions of a stage are one product per element group,
//   rhs = [Dw_1 | .. | Dw_DIM | -LIFT] . [G_1; ..; G_DIM; Fscale pi]
//   G_r = sum_x (dxi_r/dx_x) F_x(q)      (the flux Jacobian folded into the volume operand)
// (the nodal form of getElStiffVector + getElFlux + linEq, src/Mesh.cpp:734-751, :822-839,
// src/utils.cpp:406-412), computed on v_mfma_f32_16x16x32_bf16 with FP32 accumulation. FP32
// operands are carried as sums of BF16 terms: the operator is split on the host from its FP64
// values, the data on the device from its FP32 values (x = h + m + l, each term the nearest
// BF16 of what the previous ones leave: unbiased rounding; a truncated last term measured a
// 2.4x larger error against FP64 after 300 RK4 steps at p = 5). NSPLIT = 3 keeps every product
// term down to 2^-16 relative (hh, hm, mh, hl, lh, mm: six matrix products), i.e. FP32 operand
// precision; NSPLIT = 2 keeps hh, hl, lh (three products) with operands carried to about 2^-17.
//
// Work layout: one wave per group of 16 elements (matrix column = element), one 16-row tile
// per 16 element nodes, one accumulator tile per (row tile, field). The K dimension is cut in
// chunks of 8: volume chunks (direction r, nodes 8c..8c+7) then lift chunks (face nodes
// 8c..8c+7), four chunks per k-step (lane group l/16 takes chunk 4 ks + l/16), volume and lift
// k-steps kept apart (zero chunks pad each part). The operator fragments of all k-steps are
// staged once per persistent block in LDS; each lane builds the data fragments of its chunk
// and element from the stage input (volume: G_r at 8 nodes; lift: the numerical flux at 8
// face nodes, neighbour traces gathered from HBM) and splits them. The RK register update
// works directly on the accumulators: lane l holds nodes 16 it + 4 (l/16) + 0..3 of element
// l % 16, all fields.
#pragma once
#include "dg_kernels.hpp"

namespace dgk {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

template <int DIM, int P>
struct Bf16Shape {
    using S = Sizes<DIM, P>;
    static constexpr int NP = S::NP, NFT = S::NFT;
    static constexpr int MT = (NP + 15) / 16;               // row tiles (element nodes)
    static constexpr int CPR = (NP + 7) / 8;                // volume chunks per direction
    static constexpr int KSV = (DIM * CPR + 3) / 4;         // volume k-steps
    static constexpr int KSL = ((NFT + 7) / 8 + 3) / 4;     // lift k-steps
    static constexpr int KS = KSV + KSL;
    static constexpr int kWaves = 8;                         // waves per block
    static constexpr int kBlockSize = 64 * kWaves;
    static constexpr int kGroup = 16;                        // elements per wave group
};

// Operator fragment bytes of one (DIM, P, NSPLIT): [MT][KS][NSPLIT][64 lanes][8 bf16].
template <int DIM, int P, int NSPLIT>
constexpr int bf16FragBytes() {
    return Bf16Shape<DIM, P>::MT * Bf16Shape<DIM, P>::KS * NSPLIT * 64 * 16;
}

// Whether the kernel is built for (DIM, P, NSPLIT): the operator fragments fit in LDS beside
// the per-wave tables.
template <int DIM, int P, int NSPLIT>
constexpr bool bf16Supported() {
    return bf16FragBytes<DIM, P, NSPLIT>() <= 112 * 1024;
}

// Orders a wave's LDS accesses across lanes (the wave's own tables; no other wave reads them).
__device__ inline void waveLdsSync() {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ inline uint32_t f2u(float x) { return __builtin_bit_cast(uint32_t, x); }
__device__ inline float u2f(uint32_t x) { return __builtin_bit_cast(float, x); }

typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));

// Nearest BF16 pair (round to nearest even, v_cvt_pk_bf16_f32), x low / y high.
__device__ inline uint32_t bf16Pair(f32x2_t v) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2_t));
}

// The pair's values as floats.
__device__ inline f32x2_t unpackPair(uint32_t w) { return f32x2_t{u2f(w << 16), u2f(w & 0xffff0000u)}; }

// Data fragments of 8 values per field: frag[f][s] = BF16 term s of vals[f][0..7], each term the
// nearest BF16 of what the previous terms leave (the remainders x - h and x - h - m are exact
// in FP32, so NSPLIT terms carry x to about 2^-(8 NSPLIT + 1) with unbiased rounding).
template <int NF, int NSPLIT>
__device__ inline void buildFrag(const float (&vals)[NF][8], u32x4_t (&frag)[NF][NSPLIT]) {
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        f32x2_t v0 = {vals[f][0], vals[f][1]}, v1 = {vals[f][2], vals[f][3]};
        f32x2_t v2 = {vals[f][4], vals[f][5]}, v3 = {vals[f][6], vals[f][7]};
#pragma unroll
        for (int s = 0; s < NSPLIT; ++s) {
            const uint32_t w0 = bf16Pair(v0), w1 = bf16Pair(v1), w2 = bf16Pair(v2), w3 = bf16Pair(v3);
            frag[f][s] = u32x4_t{w0, w1, w2, w3};
            if (s + 1 < NSPLIT) {
                v0 = v0 - unpackPair(w0);
                v1 = v1 - unpackPair(w1);
                v2 = v2 - unpackPair(w2);
                v3 = v3 - unpackPair(w3);
            }
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
// k-step instead of one per product.
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

template <int NF>
__device__ inline void loadNodeAny(const float* __restrict__ p, float (&v)[NF]) {
    loadNode<NF, float>(p, v);
}

template <int DIM, int P, int NF, int NSPLIT, int MODE>
__global__ __launch_bounds__((Bf16Shape<DIM, P>::kBlockSize), 1) void stageKernelBf16(StageArgs<float> a,
                                                                                      const u32x4_t* __restrict__ afrag,
                                                                                      int64_t nGroups) {
    using S = Sizes<DIM, P>;
    using L = OpLayout<DIM, P>;
    using B = Bf16Shape<DIM, P>;
    constexpr int NP = S::NP, NFP = S::NFP, NFC = S::NFC, NFT = S::NFT;
    constexpr int MT = B::MT, CPR = B::CPR, KSV = B::KSV, KS = B::KS, W = B::kWaves, G = B::kGroup;
    constexpr int JJ = DIM * DIM;
    constexpr int kFrag = MT * KS * NSPLIT * 64;  // 16-byte fragment words
    constexpr bool kFinal = MODE == kModeFinal || MODE == kModeEuler;

    __shared__ u32x4_t sA[kFrag];
    __shared__ int32_t sI[L::kITotal];
    __shared__ int32_t sNb[W][G * NFC];
    __shared__ int32_t sCd[W][G * NFC];
    __shared__ __align__(16) float sFg[W][G * NFC][4];
    __shared__ double sRed[B::kBlockSize / 16][5];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int col = lane & 15, grp4 = lane >> 4;

    // operator fragments and index tables, once per block
    for (int x = tid; x < kFrag; x += B::kBlockSize) sA[x] = afrag[x];
    for (int x = tid; x < L::kITotal; x += B::kBlockSize) sI[x] = a.iops[x];
    __syncthreads();

    // contiguous range of groups per block; blocks of one XCD take adjacent ranges
    const int64_t nb = gridDim.x;
    const int64_t bi = xcdRemap(blockIdx.x, nb);
    const int64_t g0 = bi * nGroups / nb, g1 = (bi + 1) * nGroups / nb;
    double res[5] = {0, 0, 0, 0, 0};
    const float dt = a.dt;
    const Physics<float> ph = a.ph;

    for (int64_t grp = g0 + wave; grp < g1; grp += W) {
        const int64_t e0 = grp * G;
        const bool colOk = e0 + col < a.K;
        const int64_t e = colOk ? e0 + col : a.K - 1;  // clamped element of this lane's column

        // per-element face data of the group -> wave-private LDS (elements past the end: the
        // last element's)
        for (int x = lane; x < G * NFC; x += 64) {
            const int64_t el = e0 + x / NFC;
            const int64_t ef = (el < a.K ? el : a.K - 1) * NFC + x % NFC;
            sNb[wave][x] = a.nbr[ef];
            sCd[wave][x] = a.code[ef];
            const float4 v = reinterpret_cast<const float4*>(a.fgeo)[ef];
            sFg[wave][x][0] = v.x;
            sFg[wave][x][1] = v.y;
            sFg[wave][x][2] = v.z;
            sFg[wave][x][3] = v.w;
        }
        // this lane's element: dxi_r/dx_x, w_r = (dxi_r/dx) . v0
        float Jl[DIM][DIM], wr[DIM];
#pragma unroll
        for (int r = 0; r < DIM; ++r) {
            wr[r] = 0.0f;
#pragma unroll
            for (int x = 0; x < DIM; ++x) {
                Jl[r][x] = a.geo[e * JJ + r * DIM + x];
                wr[r] += Jl[r][x] * ph.v0[x];
            }
        }
        waveLdsSync();

        f32x4_t acc[MT][NF];
#pragma unroll
        for (int it = 0; it < MT; ++it)
#pragma unroll
            for (int f = 0; f < NF; ++f) acc[it][f] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};

        // volume k-steps: chunk (r, c) = G_r at nodes 8c .. 8c+7 of element e
#pragma unroll 1
        for (int ks = 0; ks < KSV; ++ks) {
            float vals[NF][8];
            const int v = 4 * ks + grp4;
            const bool ok = v < DIM * CPR;
            const int r = ok ? v / CPR : 0;
            const int c = ok ? v - r * CPR : 0;
            float jr[DIM], w = 0.0f;
#pragma unroll
            for (int x = 0; x < DIM; ++x) jr[x] = 0.0f;
#pragma unroll
            for (int rr = 0; rr < DIM; ++rr)
                if (rr == r) {
#pragma unroll
                    for (int x = 0; x < DIM; ++x) jr[x] = Jl[rr][x];
                    w = wr[rr];
                }
            float q[8][NF];
#pragma unroll
            for (int jj = 0; jj < 8; ++jj) {
                const int j = 8 * c + jj;
                loadNodeAny<NF>(a.qin + (e * NP + (j < NP ? j : NP - 1)) * NF, q[jj]);
            }
#pragma unroll
            for (int jj = 0; jj < 8; ++jj) {
                const bool in = ok && 8 * c + jj < NP;
                float jv = 0.0f;
#pragma unroll
                for (int x = 0; x < DIM; ++x) jv += jr[x] * q[jj][1 + x];
                const float pq = q[jj][0] * ph.irho0;
                float gv[NF];
                gv[0] = w * q[jj][0] + ph.rc2 * jv;
#pragma unroll
                for (int x = 0; x < 3; ++x)
                    gv[1 + x] = w * q[jj][1 + x] + (x < DIM ? jr[x < DIM ? x : 0] * pq : 0.0f);
                if constexpr (NF == 5) gv[4] = w * q[jj][4] + ph.rho0 * jv;
#pragma unroll
                for (int f = 0; f < NF; ++f) vals[f][jj] = in ? gv[f] : 0.0f;
            }
            u32x4_t bfr[NF][NSPLIT];
            buildFrag<NF, NSPLIT>(vals, bfr);
#pragma unroll
            for (int it = 0; it < MT; ++it) {
                u32x4_t afr[NSPLIT];
#pragma unroll
                for (int s = 0; s < NSPLIT; ++s) afr[s] = sA[((it * KS + ks) * NSPLIT + s) * 64 + lane];
#pragma unroll
                for (int f = 0; f < NF; ++f) acc[it][f] = mfmaSplit<NSPLIT>(afr, bfr[f], acc[it][f]);
            }
        }
        // lift k-steps: chunk c = scaled numerical flux at face nodes 8c .. 8c+7 of element e
#pragma unroll 1
        for (int ks = KSV; ks < KS; ++ks) {
            float vals[NF][8];
            const int c = 4 * (ks - KSV) + grp4;
            float qm[8][NF], qp[8][NF];
            int nbs[8], efs[8];
#pragma unroll
            for (int jj = 0; jj < 8; ++jj) {
                const int fa = 8 * c + jj;
                const int fac = fa < NFT ? fa : 0;
                const int f = fac / NFP, an = fac - f * NFP;
                const int ef = col * NFC + f;
                const int nbv = sNb[wave][ef];
                const int cd = sCd[wave][ef];
                const int own = sI[L::kMask + f * NFP + an];
                const int nn = sI[L::kMask + (cd & 3) * NFP + sI[L::kPerm + (cd >> 2) * NFP + an]];
                const float* src = a.qin + (e * NP + own) * NF;
                if (nbv >= 0) src = a.qin + (static_cast<int64_t>(nbv) * NP + nn) * NF;
                else if (nbv <= -3) src = a.ghost + (static_cast<int64_t>(-3 - nbv) * NFP + an) * NF;
                loadNodeAny<NF>(a.qin + (e * NP + own) * NF, qm[jj]);
                loadNodeAny<NF>(src, qp[jj]);
                nbs[jj] = nbv;
                efs[jj] = ef;
            }
#pragma unroll
            for (int jj = 0; jj < 8; ++jj) {
                const bool in = 8 * c + jj < NFT;
                const int ef = efs[jj];
                const float nv[3] = {sFg[wave][ef][0], sFg[wave][ef][1], sFg[wave][ef][2]};
                const float fs = sFg[wave][ef][3];
                float pi[NF];
                numericalFlux<DIM, NF>(ph, nv, nbs[jj] <= -3 ? 0 : nbs[jj], qm[jj], qp[jj], pi);
#pragma unroll
                for (int k = 0; k < NF; ++k) vals[k][jj] = in ? fs * pi[k] : 0.0f;
            }
            u32x4_t bfr[NF][NSPLIT];
            buildFrag<NF, NSPLIT>(vals, bfr);
#pragma unroll
            for (int it = 0; it < MT; ++it) {
                u32x4_t afr[NSPLIT];
#pragma unroll
                for (int s = 0; s < NSPLIT; ++s) afr[s] = sA[((it * KS + ks) * NSPLIT + s) * 64 + lane];
#pragma unroll
                for (int f = 0; f < NF; ++f) acc[it][f] = mfmaSplit<NSPLIT>(afr, bfr[f], acc[it][f]);
            }
        }

        // ---- register update of this lane's nodes (rows 16 it + 4 grp4 + rr of element e) ----
        if (colOk) {
#pragma unroll
            for (int it = 0; it < MT; ++it)
#pragma unroll
                for (int rr = 0; rr < 4; ++rr) {
                    const int i = 16 * it + 4 * grp4 + rr;
                    if (i >= NP) continue;
                    const int64_t base = (e * NP + i) * NF;
                    float rhs[NF];
#pragma unroll
                    for (int f = 0; f < NF; ++f) rhs[f] = acc[it][f][rr];
                    if constexpr (MODE == kModeRHS) {
                        float o[NF];
#pragma unroll
                        for (int k = 0; k < NF; ++k) o[k] = dt * rhs[k];
                        storeNode<NF>(a.qout + base, o);
                    } else if constexpr (MODE == kModeFirst || MODE == kModeMid) {
                        const float half = (MODE == kModeMid && a.stage == kStageRK3) ? 1.0f : 0.5f;
                        float uv[NF], pacc[NF], accn[NF], qn[NF];
                        if constexpr (MODE == kModeFirst) {
                            loadNode<NF>(a.qin + base, uv);
                        } else {
                            loadNodeNT<NF>(a.u + base, uv);
                            loadNodeNT<NF>(a.acc + base, pacc);
                        }
#pragma unroll
                        for (int k = 0; k < NF; ++k) {
                            const float kk = dt * rhs[k];
                            accn[k] = (MODE == kModeFirst) ? kk : pacc[k] + 2.0f * kk;
                            qn[k] = uv[k] + half * kk;
                        }
                        storeNodeNT<NF>(a.acc + base, accn);
                        storeNode<NF>(a.qout + base, qn);
                    } else {
                        float uv[NF], pacc[NF], plo[NF], ps[NF], unew[NF], lnew[NF];
                        if constexpr (MODE == kModeFinal) {
                            loadNodeNT<NF>(a.u + base, uv);
                            loadNodeNT<NF>(a.acc + base, pacc);
                        } else {
                            loadNode<NF>(a.qin + base, uv);
                        }
                        loadNodeNT<NF>(a.ulo + base, plo);
                        if (a.snap) loadNodeNT<NF>(a.snap + base, ps);
#pragma unroll
                        for (int k = 0; k < NF; ++k) {
                            const float kk = dt * rhs[k];
                            const float delta = (MODE == kModeFinal) ? (pacc[k] + kk) / 6.0f : kk;
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
                                d[k] = (static_cast<double>(unew[k]) + static_cast<double>(lnew[k])) -
                                       static_cast<double>(ps[k]);
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
        }
        // the next group reuses the wave's face tables
        waveLdsSync();
    }
    if constexpr (kFinal) {
        if (a.snap) blockResiduals<B::kBlockSize>(res, sRed, a.resPartial + static_cast<int64_t>(blockIdx.x) * 5);
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

}  // namespace dgk
