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
    const int32_t* __restrict__ nbr;   // [K][NFC]
    const uint8_t* __restrict__ code;  // [K][NFC]
    const T* __restrict__ ops;
    const int32_t* __restrict__ iops;
    const T* __restrict__ snap;       // snapshot (hi) for residuals, RK4/Euler final stage
    const T* __restrict__ snaplo;     // snapshot compensation or null
    double* __restrict__ resPartial;  // [gridDim][5]
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
        T fm[NF], fp[NF];
        normalFlux<NF>(ph, n, vn0, qm, fm);
        normalFlux<NF>(ph, n, vn0, qp, fp);
        const T hc = T(0.5) * ph.c0;
#pragma unroll
        for (int k = 0; k < NF; ++k) pi[k] = T(0.5) * (fm[k] + fp[k]) + hc * (qm[k] - qp[k]);
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
__device__ inline void twoSum(T a, T b, T& s, T& e) {
    s = a + b;
    T bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

template <int DIM, int P, int NF, typename T, bool COMP>
__global__ __launch_bounds__(kBlock) void stageKernel(StageArgs<T> a) {
    using S = Sizes<DIM, P>;
    using L = OpLayout<DIM, P>;
    constexpr int NP = S::NP, NFP = S::NFP, NFC = S::NFC, NFT = S::NFT, EB = S::EB;
    constexpr int JJ = DIM * DIM;
    constexpr bool kOpsInLds = (L::kTotal * sizeof(T)) <= 40 * 1024;
    constexpr int kOpsLds = kOpsInLds ? L::kTotal : 1;

    __shared__ T sQ[EB * NP * NF];
    __shared__ T sF[EB * NFT * NF];
    __shared__ T sJ[EB * JJ];
    __shared__ T sOps[kOpsLds];
    __shared__ int32_t sI[L::kITotal];
    __shared__ double sRed[kBlock / 64][5];

    const int tid = threadIdx.x;
    const int64_t blk = xcdRemap(blockIdx.x, gridDim.x);
    const int64_t e0 = blk * EB;
    const int nE = static_cast<int>((a.K - e0) < EB ? (a.K - e0) : EB);

    if constexpr (kOpsInLds)
        for (int x = tid; x < L::kTotal; x += kBlock) sOps[x] = a.ops[x];
    for (int x = tid; x < L::kITotal; x += kBlock) sI[x] = a.iops[x];
    const T* ops = kOpsInLds ? sOps : a.ops;
    {
        const T* src = a.qin + e0 * NP * NF;
        for (int x = tid; x < nE * NP * NF; x += kBlock) sQ[x] = src[x];
        const T* g = a.geo + e0 * JJ;
        for (int x = tid; x < nE * JJ; x += kBlock) sJ[x] = g[x];
    }
    __syncthreads();

    // ---- numerical fluxes at face nodes, scaled by Fscale = sJ/|J| ----
    for (int it = tid; it < nE * NFT; it += kBlock) {
        const int el = it / NFT, fa = it - el * NFT;
        const int f = fa / NFP, an = fa - f * NFP;
        const int64_t eg = e0 + el;
        const T* J = &sJ[el * JJ];
        T nv[3] = {T(0), T(0), T(0)};
#pragma unroll
        for (int c = 0; c < DIM; ++c) {
            T s = T(0);
#pragma unroll
            for (int r = 0; r < DIM; ++r) s += J[r * DIM + c] * ops[L::kN + f * 3 + r];
            nv[c] = s;
        }
        const T fs = sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
        const T ifs = T(1) / fs;
        nv[0] *= ifs;
        nv[1] *= ifs;
        nv[2] *= ifs;
        const int own = sI[L::kMask + f * NFP + an];
        T qm[NF], qp[NF], pi[NF];
#pragma unroll
        for (int k = 0; k < NF; ++k) qm[k] = sQ[(el * NP + own) * NF + k];
        const int nb = a.nbr[eg * NFC + f];
        if (nb >= 0) {
            const int cd = a.code[eg * NFC + f];
            const int fp = cd & 3, pc = cd >> 2;
            const int nn = sI[L::kMask + fp * NFP + sI[L::kPerm + pc * NFP + an]];
            const T* src = a.qin + (static_cast<int64_t>(nb) * NP + nn) * NF;
#pragma unroll
            for (int k = 0; k < NF; ++k) qp[k] = src[k];
        } else {
#pragma unroll
            for (int k = 0; k < NF; ++k) qp[k] = T(0);
        }
        numericalFlux<DIM, NF>(a.ph, nv, nb, qm, qp, pi);
#pragma unroll
        for (int k = 0; k < NF; ++k) sF[(el * NFT + fa) * NF + k] = fs * pi[k];
    }
    __syncthreads();

    // ---- volume contraction, flux-Jacobian fold, lift, register update ----
    double res[5] = {0, 0, 0, 0, 0};
    if (tid < nE * NP) {
        const int el = tid / NP, i = tid - el * NP;
        const int64_t node = (e0 + el) * NP + i;
        T Dq[DIM][NF];
#pragma unroll
        for (int r = 0; r < DIM; ++r)
#pragma unroll
            for (int k = 0; k < NF; ++k) Dq[r][k] = T(0);
        const T* q0 = &sQ[el * NP * NF];
#pragma unroll 4
        for (int j = 0; j < NP; ++j) {
            T q[NF];
#pragma unroll
            for (int k = 0; k < NF; ++k) q[k] = q0[j * NF + k];
#pragma unroll
            for (int r = 0; r < DIM; ++r) {
                const T d = ops[L::kDw + (j * NP + i) * DIM + r];
#pragma unroll
                for (int k = 0; k < NF; ++k) Dq[r][k] += d * q[k];
            }
        }
        // W_x = sum_r (dxi_r/dx_x) Dq_r ; rhs = sum_x A_x W_x
        const T* J = &sJ[el * JJ];
        T W[3][NF];
#pragma unroll
        for (int x = 0; x < 3; ++x)
#pragma unroll
            for (int k = 0; k < NF; ++k) {
                T s = T(0);
                if (x < DIM) {
#pragma unroll
                    for (int r = 0; r < DIM; ++r) s += J[r * DIM + x] * Dq[r][k];
                }
                W[x][k] = s;
            }
        T rhs[NF];
        {
            T rp = T(0), rr = T(0);
#pragma unroll
            for (int x = 0; x < DIM; ++x) {
                rp += a.ph.v0[x] * W[x][0] + a.ph.rc2 * W[x][1 + x];
                if constexpr (NF == 5) rr += a.ph.v0[x] * W[x][4] + a.ph.rho0 * W[x][1 + x];
            }
            rhs[0] = rp;
#pragma unroll
            for (int kv = 0; kv < 3; ++kv) {
                T s = T(0);
#pragma unroll
                for (int x = 0; x < DIM; ++x) s += a.ph.v0[x] * W[x][1 + kv];
                if (kv < DIM) s += W[kv][0] * a.ph.irho0;
                rhs[1 + kv] = s;
            }
            if constexpr (NF == 5) rhs[4] = rr;
        }
        const T* f0 = &sF[el * NFT * NF];
#pragma unroll 4
        for (int fa = 0; fa < NFT; ++fa) {
            const T l = ops[L::kL + fa * NP + i];
#pragma unroll
            for (int k = 0; k < NF; ++k) rhs[k] -= l * f0[fa * NF + k];
        }

        const int64_t base = node * NF;
        const T dt = a.dt;
        if (a.stage == kStageRHS) {
#pragma unroll
            for (int k = 0; k < NF; ++k) a.qout[base + k] = dt * rhs[k];
        } else if (a.stage <= kStageRK3) {
            const T half = (a.stage == kStageRK3) ? T(1) : T(0.5);
#pragma unroll
            for (int k = 0; k < NF; ++k) {
                const T kk = dt * rhs[k];
                T uv, ul = T(0);
                if (a.stage == kStageRK1) {
                    uv = q0[i * NF + k];
                    a.acc[base + k] = kk;
                } else {
                    uv = a.u[base + k];
                    a.acc[base + k] = a.acc[base + k] + T(2) * kk;
                }
                if constexpr (COMP) ul = a.ulo[base + k];
                if constexpr (COMP)
                    a.qout[base + k] = uv + (ul + half * kk);
                else
                    a.qout[base + k] = uv + half * kk;
            }
        } else {
            // final update: u += (k1 + 2k2 + 2k3 + k4)/6  (RK4)  or  u += k  (Euler)
            T unew[NF];
            double lo[NF];
#pragma unroll
            for (int k = 0; k < NF; ++k) {
                const T kk = dt * rhs[k];
                const T delta = (a.stage == kStageRK4) ? (a.acc[base + k] + kk) / T(6) : kk;
                const T uv = (a.stage == kStageRK4) ? a.u[base + k] : q0[i * NF + k];
                if constexpr (COMP) {
                    T s, e, t, e2;
                    const T ul = a.ulo[base + k];
                    twoSum(ul, delta, t, e2);
                    twoSum(uv, t, s, e);
                    const T nlo = e + e2;
                    T h, l2;
                    twoSum(s, nlo, h, l2);
                    a.uout[base + k] = h;
                    a.ulo[base + k] = l2;
                    unew[k] = h;
                    lo[k] = static_cast<double>(l2);
                } else {
                    const T un = uv + delta;
                    a.uout[base + k] = un;
                    unew[k] = un;
                    lo[k] = 0.0;
                }
            }
            if (a.snap) {
                double d[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    double sv = static_cast<double>(a.snap[base + k]);
                    if constexpr (COMP) sv += static_cast<double>(a.snaplo[base + k]);
                    d[k] = (static_cast<double>(unew[k]) + lo[k]) - sv;
                }
                const double ic2 = static_cast<double>(a.ph.ic2);
                double sp = 0.0;
                {
                    double s0 = static_cast<double>(a.snap[base]);
                    if constexpr (COMP) s0 += static_cast<double>(a.snaplo[base]);
                    double u0 = static_cast<double>(unew[0]) + lo[0];
                    sp = s0 * ic2 - u0 * ic2;
                }
                res[0] = d[0] * d[0];
                res[1] = sp * sp;
                res[2] = d[1] * d[1];
                res[3] = d[2] * d[2];
                res[4] = d[3] * d[3];
            }
        }
    }
    if ((a.stage == kStageRK4 || a.stage == kStageEuler) && a.snap) {
        // block reduction of the residual partial sums (deterministic order)
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            double v = res[k];
            for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
            if ((tid & 63) == 0) sRed[tid >> 6][k] = v;
        }
        __syncthreads();
        if (tid < 5) {
            double s = 0.0;
            for (int w = 0; w < kBlock / 64; ++w) s += sRed[w][tid];
            a.resPartial[static_cast<int64_t>(blockIdx.x) * 5 + tid] = s;
        }
    }
}

// ---------------------------------------------------------------- per-step kernels ----

struct SourceArgs {
    const int64_t* nodeOffsets;  // [nsrc+1] into nodeIdx
    const int64_t* nodeIdx;      // flat element-node indices (canonical)
    const double* params;        // [nsrc][5]: kind, amp, freq, phase, duration
    const double* tTable;        // [slots] step times
    const double* vTable;        // [slots][nsrc] table values (kind 1)
    const uint8_t* vFlag;        // [slots][nsrc] 1 = overwrite (kind 1)
    const int64_t* stepCounter;  // slot of the current step
    int nsrc;
    int nf;
};

// Hard-source overwrite of the pressure at the source nodes (src/solver.cpp:455-483).
template <typename T, bool COMP>
__global__ void sourceKernel(SourceArgs s, T* __restrict__ u, T* __restrict__ ulo, double ic2) {
    const int64_t slot = *s.stepCounter;
    const double t = s.tTable[slot];
    for (int src = blockIdx.y; src < s.nsrc; src += gridDim.y) {
        const double* pr = s.params + src * 5;
        double value;
        bool active;
        if (pr[0] == 0.0) {
            active = t < pr[4];
            value = pr[1] * sin(2.0 * M_PI * pr[2] * t + pr[3]);
        } else {
            active = s.vFlag[slot * s.nsrc + src] != 0;
            value = s.vTable[slot * s.nsrc + src];
        }
        if (!active) continue;
        const int64_t b = s.nodeOffsets[src], e = s.nodeOffsets[src + 1];
        for (int64_t x = b + blockIdx.x * blockDim.x + threadIdx.x; x < e; x += gridDim.x * blockDim.x) {
            const int64_t base = s.nodeIdx[x] * s.nf;
            const T hi = static_cast<T>(value);
            u[base] = hi;
            if constexpr (COMP) ulo[base] = static_cast<T>(value - static_cast<double>(hi));
            if (s.nf == 5) {
                const double rv = value * ic2;
                const T rh = static_cast<T>(rv);
                u[base + 4] = rh;
                if constexpr (COMP) ulo[base + 4] = static_cast<T>(rv - static_cast<double>(rh));
            }
        }
    }
}

struct PostArgs {
    const double* resPartial;   // [nblocks][5]
    int64_t nblocks;
    double invN;                // 1/(K*Np)
    const int64_t* obsOffsets;  // [nobs+1]
    const int64_t* obsIdx;      // element-node indices
    const double* obsW;         // IDW weights
    int nobs;
    int nf;
    double* resRing;            // [slots][5]
    double* obsRing;            // [slots][nobs][4]
    int* nonfinite;             // first slot with a non-finite residual, or -1
    int64_t* stepCounter;
    int computeResiduals;
};

// Residual reduction (src/solver.cpp:518-547) and inverse-distance-weighted observer
// samples (src/solver.cpp:552-573) for the step just finished; advances the step slot.
template <typename T, bool COMP>
__global__ __launch_bounds__(256) void postKernel(PostArgs p, const T* __restrict__ u, const T* __restrict__ ulo) {
    __shared__ double red[256][5];
    const int tid = threadIdx.x;
    const int64_t slot = *p.stepCounter;
    double s[5] = {0, 0, 0, 0, 0};
    if (p.computeResiduals)
        for (int64_t b = tid; b < p.nblocks; b += 256)
            for (int k = 0; k < 5; ++k) s[k] += p.resPartial[b * 5 + k];
    for (int k = 0; k < 5; ++k) red[tid][k] = s[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w)
            for (int k = 0; k < 5; ++k) red[tid][k] += red[tid + w][k];
        __syncthreads();
    }
    if (tid < 5) {
        double v = red[0][tid] * p.invN;
        p.resRing[slot * 5 + tid] = v;
        if (!isfinite(v)) atomicCAS(p.nonfinite, -1, static_cast<int>(slot));
    }
    // observers: one wave per observer
    const int wave = tid >> 6, lane = tid & 63;
    for (int o = wave; o < p.nobs; o += 4) {
        double acc[5] = {0, 0, 0, 0, 0};
        for (int64_t x = p.obsOffsets[o] + lane; x < p.obsOffsets[o + 1]; x += 64) {
            const double w = p.obsW[x];
            const int64_t base = p.obsIdx[x] * p.nf;
            for (int k = 0; k < 4; ++k) {
                double v = static_cast<double>(u[base + k]);
                if constexpr (COMP) v += static_cast<double>(ulo[base + k]);
                acc[k] += v * w;
            }
            acc[4] += w;
        }
        for (int k = 0; k < 5; ++k)
            for (int off = 32; off > 0; off >>= 1) acc[k] += __shfl_down(acc[k], off, 64);
        if (lane == 0)
            for (int k = 0; k < 4; ++k) p.obsRing[(slot * p.nobs + o) * 4 + k] = acc[k] / acc[4];
    }
    __syncthreads();
    if (tid == 0) *p.stepCounter = slot + 1;
}

__global__ inline void setCounterKernel(int64_t* counter, int64_t value, int* flag) {
    *counter = value;
    *flag = -1;
}

}  // namespace dgk
