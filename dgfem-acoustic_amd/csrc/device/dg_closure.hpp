// Learned absorbing-boundary closure (no counterpart in the reference solver; its
// characteristic absorbing flux, src/Mesh.cpp:619-666 and :954-969, stays the fallback).
//
// On a truncation face the exterior state of the interior numerical flux
// n.{F} + (c0/2)(q- - q+) is built so that the flux carries a prescribed incoming
// characteristic predicted by a small MLP from the outgoing part of the interior trace.
// Face frame: n (outward unit normal) and DIM-1 tangents; 2D t = (-n_y, n_x); 3D t1 along
// the tangential mean flow v0 - vn0 n when |v0 - vn0 n| > 1e-6 c0 (else along e_a - (e_a.n) n
// for the axis a least aligned with n) and t2 = n x t1. Per face node j:
//   P = p / (rho0 c0), N = v.n, T_c = v.t_c; characteristics of the normal flux Jacobian:
//   w_out = P + N (speed vn0 + c0), w_in = P - N (speed vn0 - c0), T_c (speed vn0)
//   scale s = sqrt(mean_j (w_out^2 + sum_c T_c^2)) + 1e-30
//   x = [w_out/s, T_1/s, (T_2/s,) vn0/c0, vt0/c0]  (n_in = DIM Nfp + 2; 2D vt0 = v0.t,
//                                                  3D vt0 = |v0 - vn0 n|)
//   y = W3 tanh(W2 tanh(W1 x + b1) + b2) + b3     (n_out = Nfp), w_hat = s y
// Exterior state (upwind flux with incoming values w_hat and 0):
//   w_out+ = w_out-
//   w_in+  = (2 lam w_hat - (lam + c0) w_in-) / (lam - c0),  lam = vn0 - c0
//   T_c+   = T_c- (vn0 >= 0);  -(vn0 + c0) T_c- / (vn0 - c0) (vn0 < 0: no incoming vorticity)
//   p+ = rho0 c0 (w_out+ + w_in+)/2,  v+ = (w_out+ - w_in+)/2 n + sum_c T_c+ t_c (2D: vz+ = vz-),
//   rho+ = rho- + (p+ - p-)/c0^2
// The network never sees w_in-, so it cannot feed the interior's incoming content back in;
// the zero network is the exact characteristic (first-order non-reflecting) condition. The
// map is homogeneous of degree one in the trace and equivariant under rotations of the
// face (in 3D when the tangential mean flow is nonzero; else the tangents follow the axes). Weights are BF16, activations and accumulation FP32. Parameter order (row-major,
// out x in): W1[64][n_in], b1[64], W2[64][64], b2[64], W3[n_out][64], b3[n_out].
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "dg_kernels.hpp"
#include "dg_kernels_bf16.hpp"

namespace dgc {

constexpr int kH = 64;  // hidden width

__host__ __device__ constexpr int numParams(int nin, int nout) {
    return kH * nin + kH + kH * kH + kH + nout * kH + nout;
}

__host__ __device__ inline float bf16ToFloat(uint16_t h) {
    union {
        uint32_t u;
        float f;
    } v;
    v.u = static_cast<uint32_t>(h) << 16;
    return v.f;
}

// Round to nearest even (NaN kept quiet).
__host__ __device__ inline uint16_t floatToBf16(float f) {
    union {
        float f;
        uint32_t u;
    } v;
    v.f = f;
    if ((v.u & 0x7f800000u) == 0x7f800000u && (v.u & 0x7fffffu)) return static_cast<uint16_t>((v.u >> 16) | 0x40);
    const uint32_t lsb = (v.u >> 16) & 1u;
    return static_cast<uint16_t>((v.u + 0x7fffu + lsb) >> 16);
}

// LDS image of the network in FP32 with odd row strides (conflict-free row and column
// walks across a wave).
template <int NIN, int NOUT>
struct NetLds {
    static constexpr int S1 = NIN | 1;  // W1 row stride
    static constexpr int S2 = kH + 1;   // W2 / W3 row stride
    static constexpr int kW1 = 0;
    static constexpr int kB1 = kW1 + kH * S1;
    static constexpr int kW2 = kB1 + kH;
    static constexpr int kB2 = kW2 + kH * S2;
    static constexpr int kW3 = kB2 + kH;
    static constexpr int kB3 = kW3 + NOUT * S2;
    static constexpr int kTotal = kB3 + NOUT;
};

// LDS offset of parameter x (canonical order) in the image.
template <int NIN, int NOUT>
__device__ inline int netOffset(int x) {
    using N = NetLds<NIN, NOUT>;
    if (x < kH * NIN) return N::kW1 + (x / NIN) * N::S1 + x % NIN;
    if (x < kH * NIN + kH) return N::kB1 + (x - kH * NIN);
    if (x < kH * NIN + kH + kH * kH) {
        const int y = x - (kH * NIN + kH);
        return N::kW2 + (y / kH) * N::S2 + y % kH;
    }
    if (x < kH * NIN + 2 * kH + kH * kH) return N::kB2 + (x - (kH * NIN + kH + kH * kH));
    if (x < kH * NIN + 2 * kH + kH * kH + NOUT * kH) {
        const int y = x - (kH * NIN + 2 * kH + kH * kH);
        return N::kW3 + (y / kH) * N::S2 + y % kH;
    }
    return N::kB3 + (x - (kH * NIN + 2 * kH + kH * kH + NOUT * kH));
}

// Block-cooperative load of BF16 parameters (canonical order) into the LDS image by NT
// threads: every 16-byte chunk of the thread is loaded (clamped addresses) before the first
// LDS write, then the trailing P mod 8 values. The caller publishes the image with a barrier.
template <int NIN, int NOUT, int NT>
__device__ inline void loadNet(const uint16_t* __restrict__ w, float* s, int tid) {
    constexpr int P = numParams(NIN, NOUT);
    constexpr int NC = P / 8, NTAIL = P - 8 * NC;
    constexpr int IT = (NC + NT - 1) / NT;
    typedef uint32_t U4 __attribute__((ext_vector_type(4)));
    U4 r[IT];
#pragma unroll
    for (int c = 0; c < IT; ++c) {
        const int x = tid + c * NT;
        r[c] = reinterpret_cast<const U4*>(w)[x < NC ? x : NC - 1];
    }
    uint16_t rt = 0;
    if constexpr (NTAIL > 0) rt = w[8 * NC + (tid < NTAIL ? tid : 0)];
#pragma unroll
    for (int c = 0; c < IT; ++c) {
        const int x = tid + c * NT;
        if (x < NC) {
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const uint32_t word = r[c][k >> 1];
                const uint16_t h = static_cast<uint16_t>((k & 1) ? (word >> 16) : (word & 0xffffu));
                s[netOffset<NIN, NOUT>(8 * x + k)] = bf16ToFloat(h);
            }
        }
    }
    if constexpr (NTAIL > 0)
        if (tid < NTAIL) s[netOffset<NIN, NOUT>(8 * NC + tid)] = bf16ToFloat(rt);
}

__device__ inline void waveFence() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Forward pass of one sample by one wave. sx[NIN] holds the input; on return sa1/sa2 hold
// the hidden activations and sy[NOUT] the output.
template <int NIN, int NOUT>
__device__ inline void mlpForwardWave(const float* net, const float* sx, float* sa1, float* sa2, float* sy, int lane) {
    using N = NetLds<NIN, NOUT>;
    {
        float z = net[N::kB1 + lane];
        const float* w = net + N::kW1 + lane * N::S1;
#pragma unroll
        for (int i = 0; i < NIN; ++i) z = fmaf(w[i], sx[i], z);
        sa1[lane] = tanhf(z);
    }
    waveFence();
    {
        float z = net[N::kB2 + lane];
        const float* w = net + N::kW2 + lane * N::S2;
#pragma unroll 16
        for (int k = 0; k < kH; ++k) z = fmaf(w[k], sa1[k], z);
        sa2[lane] = tanhf(z);
    }
    waveFence();
    if (lane < NOUT) {
        float z = net[N::kB3 + lane];
        const float* w = net + N::kW3 + lane * N::S2;
#pragma unroll 16
        for (int k = 0; k < kH; ++k) z = fmaf(w[k], sa2[k], z);
        sy[lane] = z;
    }
    waveFence();
}

template <typename T>
struct ClosureArgs {
    const T* __restrict__ qin;          // stage input [K][NP][NF]
    T* __restrict__ ghost;              // [nfaces][NFP][NF]
    const T* __restrict__ fgeo;         // [K][NFC][4] outward unit normal, Fscale
    const int32_t* __restrict__ iops;   // fmask
    const int64_t* __restrict__ faces;  // element * NFC + local face
    const uint16_t* __restrict__ w;     // BF16 parameters
    int64_t nfaces;
    dgk::Physics<T> ph;
};

// Forward pass of F samples by one wave (lane = hidden unit): each weight read from LDS serves
// F samples and the F accumulations are independent. Inputs sx[F][XS] (XS >= NIN, a multiple of
// 4), activations sa[F][kH], outputs sy[F][NOUT]; sx and sa2 may share storage (sx is consumed
// before sa2 is written). Per output the sum runs over the inputs in ascending order, as in
// mlpForwardWave.
template <int NIN, int NOUT, int F, int XS>
__device__ inline void mlpForwardBatch(const float* net, const float* sx, float* sa1, float* sa2, float* sy, int lane) {
    using N = NetLds<NIN, NOUT>;
    static_assert(XS % 4 == 0 && XS >= NIN && NOUT <= 16 && F % 4 == 0, "batch layout");
    typedef float F4 __attribute__((ext_vector_type(4)));
    float z[F];
    {
        const float* w = net + N::kW1 + lane * N::S1;
        const float b = net[N::kB1 + lane];
#pragma unroll
        for (int f = 0; f < F; ++f) z[f] = b;
        constexpr int NIN4 = NIN & ~3;
#pragma unroll 2
        for (int i = 0; i < NIN4; i += 4) {
            const float w0 = w[i], w1 = w[i + 1], w2 = w[i + 2], w3 = w[i + 3];
#pragma unroll
            for (int f = 0; f < F; ++f) {
                const F4 x = *reinterpret_cast<const F4*>(sx + f * XS + i);
                z[f] = fmaf(w3, x[3], fmaf(w2, x[2], fmaf(w1, x[1], fmaf(w0, x[0], z[f]))));
            }
        }
#pragma unroll
        for (int i = NIN4; i < NIN; ++i) {
            const float wi = w[i];
#pragma unroll
            for (int f = 0; f < F; ++f) z[f] = fmaf(wi, sx[f * XS + i], z[f]);
        }
#pragma unroll
        for (int f = 0; f < F; ++f) sa1[f * kH + lane] = tanhf(z[f]);
    }
    waveFence();
    {
        const float* w = net + N::kW2 + lane * N::S2;
        const float b = net[N::kB2 + lane];
#pragma unroll
        for (int f = 0; f < F; ++f) z[f] = b;
#pragma unroll 2
        for (int k = 0; k < kH; k += 4) {
            const float w0 = w[k], w1 = w[k + 1], w2 = w[k + 2], w3 = w[k + 3];
#pragma unroll
            for (int f = 0; f < F; ++f) {
                const F4 x = *reinterpret_cast<const F4*>(sa1 + f * kH + k);
                z[f] = fmaf(w3, x[3], fmaf(w2, x[2], fmaf(w1, x[1], fmaf(w0, x[0], z[f]))));
            }
        }
#pragma unroll
        for (int f = 0; f < F; ++f) sa2[f * kH + lane] = tanhf(z[f]);
    }
    waveFence();
    {
        // lane = output j + 16 g: samples g, g + 4, ...
        const int j = lane & 15, g = lane >> 4;
        if (j < NOUT) {
            constexpr int FG = F / 4;
            const float* w = net + N::kW3 + j * N::S2;
            const float b = net[N::kB3 + j];
            float y[FG];
#pragma unroll
            for (int c = 0; c < FG; ++c) y[c] = b;
#pragma unroll 2
            for (int k = 0; k < kH; k += 4) {
                const float w0 = w[k], w1 = w[k + 1], w2 = w[k + 2], w3 = w[k + 3];
#pragma unroll
                for (int c = 0; c < FG; ++c) {
                    const F4 x = *reinterpret_cast<const F4*>(sa2 + (g + 4 * c) * kH + k);
                    y[c] = fmaf(w3, x[3], fmaf(w2, x[2], fmaf(w1, x[1], fmaf(w0, x[0], y[c]))));
                }
            }
#pragma unroll
            for (int c = 0; c < FG; ++c) sy[(g + 4 * c) * NOUT + j] = y[c];
        }
    }
    waveFence();
}

// FW faces per wave (4 or 8), kClosureWaves waves per block: the network is staged in LDS
// once per block and every weight read serves a wave's faces together. With few faces (a 2D
// truncation boundary) 4 faces per wave spread the work over twice the waves.
constexpr int kClosureWaves = 4;
constexpr int kClosureBlock = 64 * kClosureWaves;
constexpr int kClosureSmall = 4096;  // at most this many faces: 4 faces per wave, else 8

// Exterior states of the learned faces for one stage. Face-node work runs with lane =
// (face slot, node): 16 lanes per face (NFP <= 16), 4 faces per pass.
template <int DIM, int P, int NF, typename T, int FW>
__global__ __launch_bounds__(kClosureBlock) void closureKernel(ClosureArgs<T> a) {
    using S = dgk::Sizes<DIM, P>;
    using L = dgk::OpLayout<DIM, P>;
    constexpr int NP = S::NP, NFP = S::NFP, NFC = S::NFC;
    constexpr int NT = DIM - 1;  // tangential directions
    constexpr int NIN = DIM * NFP + 2, NOUT = NFP;
    constexpr int NPASS = FW / 4;
    static_assert(FW % 4 == 0, "4 faces per pass");
    constexpr int XS = (NIN + 3) & ~3;
    static_assert(NFP <= 16 && XS <= kH, "closure lane layout");
    using N = NetLds<NIN, NOUT>;
    __shared__ float net[N::kTotal];
    __shared__ __align__(16) float bufX[kClosureWaves][FW * kH];   // inputs, then second activations
    __shared__ __align__(16) float bufA[kClosureWaves][FW * kH];   // first activations
    __shared__ float bufY[kClosureWaves][FW * NOUT];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int slot = lane >> 4, node = lane & 15;
    const bool nodeLane = node < NFP;
    const int64_t f0 = static_cast<int64_t>(blockIdx.x) * (kClosureWaves * FW) + wave * FW;
    // face records of this lane's faces, issued with the network loads
    int64_t ef[NPASS];
#pragma unroll
    for (int ps = 0; ps < NPASS; ++ps) {
        const int64_t idx = f0 + ps * 4 + slot;
        ef[ps] = a.faces[idx < a.nfaces ? idx : a.nfaces - 1];
    }
    loadNet<NIN, NOUT, kClosureBlock>(a.w, net, tid);
    __syncthreads();
    float* sx = bufX[wave];
    const T c0 = a.ph.c0;
    // per pass: this lane's face frame and node characteristics
    T n[NPASS][3], t[NPASS][NT][3], vn0[NPASS], wo[NPASS], wi[NPASS], tt[NPASS][NT], qm[NPASS][NF];
    float sc[NPASS];
#pragma unroll
    for (int ps = 0; ps < NPASS; ++ps) {
        const int64_t e = ef[ps] / NFC;
        const int f = static_cast<int>(ef[ps] - e * NFC);
        n[ps][0] = a.fgeo[(e * NFC + f) * 4];
        n[ps][1] = a.fgeo[(e * NFC + f) * 4 + 1];
        n[ps][2] = DIM == 3 ? a.fgeo[(e * NFC + f) * 4 + 2] : T(0);
        const int nd = a.iops[L::kMask + f * NFP + (nodeLane ? node : 0)];
        const T* q = a.qin + (e * NP + nd) * NF;
#pragma unroll
        for (int k = 0; k < NF; ++k) qm[ps][k] = q[k];
    }
#pragma unroll
    for (int ps = 0; ps < NPASS; ++ps) {
        T vt0;
        if constexpr (DIM == 2) {
            t[ps][0][0] = -n[ps][1];
            t[ps][0][1] = n[ps][0];
            t[ps][0][2] = T(0);
            vn0[ps] = a.ph.v0[0] * n[ps][0] + a.ph.v0[1] * n[ps][1];
            vt0 = a.ph.v0[0] * t[ps][0][0] + a.ph.v0[1] * t[ps][0][1];
        } else {
            vn0[ps] = a.ph.v0[0] * n[ps][0] + a.ph.v0[1] * n[ps][1] + a.ph.v0[2] * n[ps][2];
            // t1 along the tangential mean flow, else along the axis least aligned with n
            T u[3];
            for (int c = 0; c < 3; ++c) u[c] = a.ph.v0[c] - vn0[ps] * n[ps][c];
            vt0 = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
            if (!(vt0 > T(1e-6) * c0)) {
                const T ax = fabs(n[ps][0]), ay = fabs(n[ps][1]), az = fabs(n[ps][2]);
                const int k = (ax <= ay && ax <= az) ? 0 : (ay <= az ? 1 : 2);
                for (int c = 0; c < 3; ++c) u[c] = (c == k ? T(1) : T(0)) - n[ps][k] * n[ps][c];
                vt0 = T(0);
            }
            const T iu = T(1) / sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
            for (int c = 0; c < 3; ++c) t[ps][0][c] = u[c] * iu;
            t[ps][1][0] = n[ps][1] * t[ps][0][2] - n[ps][2] * t[ps][0][1];
            t[ps][1][1] = n[ps][2] * t[ps][0][0] - n[ps][0] * t[ps][0][2];
            t[ps][1][2] = n[ps][0] * t[ps][0][1] - n[ps][1] * t[ps][0][0];
        }
        const T Pv = qm[ps][0] * a.ph.irc0;
        T Nn = qm[ps][1] * n[ps][0] + qm[ps][2] * n[ps][1];
        if constexpr (DIM == 3) Nn += qm[ps][3] * n[ps][2];
        wo[ps] = nodeLane ? Pv + Nn : T(0);
        wi[ps] = Pv - Nn;
        for (int c = 0; c < NT; ++c) {
            T v = qm[ps][1] * t[ps][c][0] + qm[ps][2] * t[ps][c][1];
            if constexpr (DIM == 3) v += qm[ps][3] * t[ps][c][2];
            tt[ps][c] = nodeLane ? v : T(0);
        }
        // scale: RMS over the face's nodes of the outgoing and tangential characteristics
        const float wof = static_cast<float>(wo[ps]);
        float ss = wof * wof;
        for (int c = 0; c < NT; ++c) {
            const float ttf = static_cast<float>(tt[ps][c]);
            ss += ttf * ttf;
        }
#pragma unroll
        for (int off = 8; off > 0; off >>= 1) ss += __shfl_xor(ss, off, 64);
        sc[ps] = sqrtf(ss / NFP) + 1e-30f;
        const float is = 1.0f / sc[ps];
        float* x = sx + (ps * 4 + slot) * XS;
        if (nodeLane) {
            x[node] = wof * is;
            for (int c = 0; c < NT; ++c) x[(1 + c) * NFP + node] = static_cast<float>(tt[ps][c]) * is;
        } else if (node == NFP) {
            const float ic0 = static_cast<float>(T(1) / c0);
            x[DIM * NFP] = static_cast<float>(vn0[ps]) * ic0;
            x[DIM * NFP + 1] = static_cast<float>(vt0) * ic0;
        } else if (node == NFP + 1) {
            for (int i = NIN; i < XS; ++i) x[i] = 0.0f;
        }
    }
    waveFence();
    mlpForwardBatch<NIN, NOUT, FW, XS>(net, sx, bufA[wave], sx, bufY[wave], lane);
#pragma unroll
    for (int ps = 0; ps < NPASS; ++ps) {
        const int64_t idx = f0 + ps * 4 + slot;
        if (!nodeLane || idx >= a.nfaces) continue;
        const T what = static_cast<T>(sc[ps] * bufY[wave][(ps * 4 + slot) * NOUT + node]);
        const T lam = vn0[ps] - c0;
        const T wip = (T(2) * lam * what - (lam + c0) * wi[ps]) / (lam - c0);
        T tp[NT];
        for (int c = 0; c < NT; ++c)
            tp[c] = (vn0[ps] >= T(0)) ? tt[ps][c] : -(vn0[ps] + c0) * tt[ps][c] / (vn0[ps] - c0);
        const T Pp = T(0.5) * (wo[ps] + wip), Np_ = T(0.5) * (wo[ps] - wip);
        T qp[NF];
        qp[0] = a.ph.rho0 * c0 * Pp;
        for (int x = 0; x < DIM; ++x) {
            T v = Np_ * n[ps][x];
            for (int c = 0; c < NT; ++c) v += tp[c] * t[ps][c][x];
            qp[1 + x] = v;
        }
        if constexpr (DIM == 2) qp[3] = qm[ps][3];
        if constexpr (NF == 5) qp[4] = qm[ps][4] + (qp[0] - qm[ps][0]) * a.ph.ic2;
        T* g = a.ghost + (idx * NFP + node) * NF;
#pragma unroll
        for (int k = 0; k < NF; ++k) g[k] = qp[k];
    }
}

// Interior and exterior traces (exterior in the interior element's face-node order) and
// outward unit normals of the listed faces: out [n][2][NFP][NF], normals [n][3].
template <int DIM, int P, int NF, typename T>
__global__ __launch_bounds__(256) void traceKernel(const T* __restrict__ q, const T* __restrict__ qlo,
                                                   const T* __restrict__ geo, const T* __restrict__ ops,
                                                   const int32_t* __restrict__ iops, const int32_t* __restrict__ nbr,
                                                   const uint8_t* __restrict__ code, const int64_t* __restrict__ elem,
                                                   const int32_t* __restrict__ face, int64_t n, double* __restrict__ out,
                                                   double* __restrict__ normals) {
    using S = dgk::Sizes<DIM, P>;
    using L = dgk::OpLayout<DIM, P>;
    constexpr int NP = S::NP, NFP = S::NFP, NFC = S::NFC;
    const int64_t it = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (it >= n * NFP) return;
    const int64_t i = it / NFP;
    const int an = static_cast<int>(it - i * NFP);
    const int64_t e = elem[i];
    const int f = face[i];
    const int own = iops[L::kMask + f * NFP + an];
    double* o = out + (i * 2 * NFP + an) * NF;
    for (int k = 0; k < NF; ++k) {
        const int64_t x = (e * NP + own) * NF + k;
        o[k] = static_cast<double>(q[x]) + (qlo ? static_cast<double>(qlo[x]) : 0.0);
    }
    const int nb = nbr[e * NFC + f];
    double* ox = out + (i * 2 * NFP + NFP + an) * NF;
    if (nb >= 0) {
        const int cd = code[e * NFC + f];
        const int fp = cd & 3, pc = cd >> 2;
        const int nn = iops[L::kMask + fp * NFP + iops[L::kPerm + pc * NFP + an]];
        for (int k = 0; k < NF; ++k) {
            const int64_t x = (static_cast<int64_t>(nb) * NP + nn) * NF + k;
            ox[k] = static_cast<double>(q[x]) + (qlo ? static_cast<double>(qlo[x]) : 0.0);
        }
    } else {
        for (int k = 0; k < NF; ++k) ox[k] = 0.0;
    }
    if (an == 0) {
        const T* J = geo + e * DIM * DIM;
        double nv[3] = {0, 0, 0}, nn2 = 0;
        for (int c = 0; c < DIM; ++c) {
            double s = 0;
            for (int r = 0; r < DIM; ++r) s += static_cast<double>(J[r * DIM + c]) * static_cast<double>(ops[L::kN + f * 3 + r]);
            nv[c] = s;
            nn2 += s * s;
        }
        const double inv = 1.0 / sqrt(nn2);
        for (int c = 0; c < 3; ++c) normals[i * 3 + c] = nv[c] * inv;
    }
}

// ---------------------------------------------------------------- training kernels ----

// Inference of n samples: X [n][NIN] -> Y [n][NOUT]; one wave per sample, grid-stride.
template <int NIN, int NOUT>
__global__ __launch_bounds__(64) void forwardKernel(const uint16_t* __restrict__ w, const float* __restrict__ X,
                                                    float* __restrict__ Y, int64_t n) {
    __shared__ float net[NetLds<NIN, NOUT>::kTotal];
    __shared__ float sx[NIN], sa1[kH], sa2[kH], sy[NOUT];
    const int lane = threadIdx.x;
    loadNet<NIN, NOUT, 64>(w, net, lane);
    __syncthreads();
    for (int64_t s = blockIdx.x; s < n; s += gridDim.x) {
        if (lane < NIN) sx[lane] = X[s * NIN + lane];
        waveFence();
        mlpForwardWave<NIN, NOUT>(net, sx, sa1, sa2, sy, lane);
        if (lane < NOUT) Y[s * NOUT + lane] = sy[lane];
        waveFence();
    }
}

// Loss L = sum over the batch of |y - y*|^2 / (batch * NOUT) and its gradient. Each block
// (one wave) takes samples [b*per, (b+1)*per) of the batch, in the order of idx (a
// permutation of the data set, or null for identity), and writes its partial gradient
// (canonical parameter order) and partial loss; partials are summed in block order.
template <int NIN, int NOUT>
__global__ __launch_bounds__(64) void gradKernel(const uint16_t* __restrict__ w, const float* __restrict__ X,
                                                 const float* __restrict__ Yt, const int64_t* __restrict__ idx,
                                                 int64_t first, int64_t count, int64_t per, float scale,
                                                 float* __restrict__ gpart, double* __restrict__ lpart) {
    using N = NetLds<NIN, NOUT>;
    constexpr int P = numParams(NIN, NOUT);
    __shared__ float net[N::kTotal];
    __shared__ float sx[NIN], sa1[kH], sa2[kH], sy[NOUT], sd3[NOUT], sd2[kH];
    const int lane = threadIdx.x;
    loadNet<NIN, NOUT, 64>(w, net, lane);
    __syncthreads();
    float g1[NIN], g2[kH], g3[NOUT];
    float gb1 = 0.f, gb2 = 0.f, gb3 = 0.f;
#pragma unroll
    for (int i = 0; i < NIN; ++i) g1[i] = 0.f;
#pragma unroll
    for (int k = 0; k < kH; ++k) g2[k] = 0.f;
#pragma unroll
    for (int o = 0; o < NOUT; ++o) g3[o] = 0.f;
    double loss = 0.0;
    const int64_t b0 = static_cast<int64_t>(blockIdx.x) * per;
    const int64_t b1 = (b0 + per < count) ? b0 + per : count;
    for (int64_t j = b0; j < b1; ++j) {
        const int64_t s = idx ? idx[first + j] : first + j;
        if (lane < NIN) sx[lane] = X[s * NIN + lane];
        waveFence();
        mlpForwardWave<NIN, NOUT>(net, sx, sa1, sa2, sy, lane);
        if (lane < NOUT) {
            const float e = sy[lane] - Yt[s * NOUT + lane];
            loss += static_cast<double>(e) * static_cast<double>(e);
            const float d3 = 2.0f * e * scale;
            sd3[lane] = d3;
            gb3 += d3;
        }
        waveFence();
        // layer 3 weights (column `lane`) and the error at the second hidden layer
        const float a2 = sa2[lane];
        float gz2 = 0.f;
#pragma unroll
        for (int o = 0; o < NOUT; ++o) {
            const float d3 = sd3[o];
            g3[o] = fmaf(d3, a2, g3[o]);
            gz2 = fmaf(net[N::kW3 + o * N::S2 + lane], d3, gz2);
        }
        const float d2 = gz2 * (1.0f - a2 * a2);
        sd2[lane] = d2;
        gb2 += d2;
        waveFence();
        // layer 2 weights (row `lane`) and the error at the first hidden layer
        const float a1 = sa1[lane];
        float gz1 = 0.f;
#pragma unroll
        for (int k = 0; k < kH; ++k) {
            g2[k] = fmaf(d2, sa1[k], g2[k]);
            gz1 = fmaf(net[N::kW2 + k * N::S2 + lane], sd2[k], gz1);
        }
        const float d1 = gz1 * (1.0f - a1 * a1);
        gb1 += d1;
#pragma unroll
        for (int i = 0; i < NIN; ++i) g1[i] = fmaf(d1, sx[i], g1[i]);
        waveFence();
    }
    float* gp = gpart + static_cast<int64_t>(blockIdx.x) * P;
    constexpr int oB1 = kH * NIN, oW2 = oB1 + kH, oB2 = oW2 + kH * kH, oW3 = oB2 + kH, oB3 = oW3 + NOUT * kH;
#pragma unroll
    for (int i = 0; i < NIN; ++i) gp[lane * NIN + i] = g1[i];
    gp[oB1 + lane] = gb1;
#pragma unroll
    for (int k = 0; k < kH; ++k) gp[oW2 + lane * kH + k] = g2[k];
    gp[oB2 + lane] = gb2;
#pragma unroll
    for (int o = 0; o < NOUT; ++o) gp[oW3 + o * kH + lane] = g3[o];
    if (lane < NOUT) gp[oB3 + lane] = gb3;
    // wave-level loss sum (lanes >= NOUT hold 0)
    for (int off = 32; off > 0; off >>= 1) loss += __shfl_xor(loss, off, 64);
    if (lane == 0) lpart[blockIdx.x] = loss;
}

// ------------------------------------------------------ matrix-core training kernel ----
//
// The same loss and partial gradient as gradKernel for one block of kMfmaTile = 32 samples,
// with every product of the forward and backward passes on v_mfma_f32_16x16x32_bf16 (FP32
// accumulation): activations are matrices [unit][sample] (a 16x16 tile = 16 units x 16
// samples), the weights (BF16 values, exact) are the A operands read straight from the BF16
// parameters, the FP32 activations and errors are split into BF16 terms (three for the
// forward products: FP32 operand precision; two for the backward products: ~2^-17), and the
// weight gradients contract over the block's 32 samples (one K step):
//   Z1 = W1 X + b1, A1 = tanh Z1; Z2 = W2 A1 + b2, A2 = tanh Z2; Y = W3 A2 + b3
//   D3 = 2 (Y - Y*) scale;  D2 = (W3^T D3) (1 - A2^2);  D1 = (W2^T D2) (1 - A1^2)
//   dW3 = D3 A2^T, dW2 = D2 A1^T, dW1 = D1 X^T, db = row sums of D.
// Matrix layouts of the instruction (checked on gfx950): A lane l row l%16, k 8(l/16)+j;
// B lane l column l%16, k 8(l/16)+j; D lane l column l%16, rows 4(l/16)+r.
constexpr int kMfmaTile = 32;   // samples per block
constexpr int kMfmaSS = 36;     // LDS row stride (floats) of [unit][sample] matrices

// 8 BF16 weights w[(r0 + l%16) * rs + (k0 + 8(l/16) + j) * ks] (zero outside rows < R, k < K)
// as an A fragment.
__device__ inline dgk::u32x4_t weightFrag(const uint16_t* __restrict__ w, int lane, int r0, int rs, int R, int k0,
                                          int ks, int K) {
    const int row = r0 + (lane & 15), kb = k0 + 8 * (lane >> 4);
    uint32_t h[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) h[j] = (row < R && kb + j < K) ? w[row * rs + (kb + j) * ks] : 0u;
    return dgk::u32x4_t{h[0] | (h[1] << 16), h[2] | (h[3] << 16), h[4] | (h[5] << 16), h[6] | (h[7] << 16)};
}

// B fragment values with k along the units: m[(u0 + 8(l/16) + j) * SS + c0 + l%16].
__device__ inline void unitFrag(const float* m, int lane, int u0, int c0, float (&v)[8]) {
    const float* p = m + (u0 + 8 * (lane >> 4)) * kMfmaSS + c0 + (lane & 15);
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = p[j * kMfmaSS];
}

// A or B fragment values with k along the samples: m[(u0 + l%16) * SS + 8(l/16) + j].
__device__ inline void sampleFrag(const float* m, int lane, int u0, float (&v)[8]) {
    const float4* p = reinterpret_cast<const float4*>(m + (u0 + (lane & 15)) * kMfmaSS + 8 * (lane >> 4));
    const float4 a = p[0], b = p[1];
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
}

// D = W . B for B split in NS terms (W exact): leading term and the rest in separate sums.
template <int NS>
__device__ inline void mmaSplitB(dgk::u32x4_t wf, const dgk::u32x4_t (&b)[NS], dgk::f32x4_t& main,
                                 dgk::f32x4_t& corr) {
    main = dgk::mfmaBf16(wf, b[0], main);
#pragma unroll
    for (int s = 1; s < NS; ++s) corr = dgk::mfmaBf16(wf, b[s], corr);
}

// Row sums over the 16 columns of a D tile (the 16 lanes of a lane group).
__device__ inline float rowSum16(float x) {
#pragma unroll
    for (int off = 1; off < 16; off <<= 1) x += __shfl_xor(x, off, 64);
    return x;
}

template <int NIN, int NOUT>
__global__ __launch_bounds__(64) void gradMfmaKernel(const uint16_t* __restrict__ w, const float* __restrict__ X,
                                                     const float* __restrict__ Yt, const int64_t* __restrict__ idx,
                                                     int64_t first, int64_t count, float scale,
                                                     float* __restrict__ gpart, double* __restrict__ lpart) {
    static_assert(NIN <= 32 && NOUT <= 16, "network shape of the matrix-core training kernel");
    constexpr int P = numParams(NIN, NOUT), SS = kMfmaSS, S = kMfmaTile;
    constexpr int oB1 = kH * NIN, oW2 = oB1 + kH, oB2 = oW2 + kH * kH, oW3 = oB2 + kH, oB3 = oW3 + NOUT * kH;
    using dgk::f32x4_t;
    using dgk::u32x4_t;
    __shared__ __align__(16) float sX[32 * SS];    // [input][sample]
    __shared__ __align__(16) float sA1[kH * SS];   // [hidden 1][sample]
    __shared__ __align__(16) float sA2[kH * SS];   // [hidden 2][sample]
    __shared__ __align__(16) float sD3[16 * SS];   // [output][sample]
    __shared__ __align__(16) float sD2[kH * SS];
    __shared__ __align__(16) float sD1[kH * SS];
    __shared__ int64_t sIdx[S];
    const int lane = threadIdx.x, g = lane >> 4, cl = lane & 15;
    const int64_t b0 = static_cast<int64_t>(blockIdx.x) * S;
    const int nS = static_cast<int>((count - b0) < S ? (count - b0) : S);
    if (lane < S) sIdx[lane] = lane < nS ? (idx ? idx[first + b0 + lane] : first + b0 + lane) : -1;
    __syncthreads();
    // inputs (zero for padding inputs and absent samples)
    for (int x = lane; x < 32 * S; x += 64) {
        const int c = x / 32, i = x - c * 32;
        const int64_t sm = sIdx[c];
        sX[i * SS + c] = (sm >= 0 && i < NIN) ? X[sm * NIN + i] : 0.f;
    }
    __syncthreads();
    const f32x4_t z4 = {0.f, 0.f, 0.f, 0.f};
    // ---- forward: layer 1 ----
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) {
        const u32x4_t wf = weightFrag(w, lane, 16 * rt, NIN, kH, 0, 1, NIN);
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) {
            float v[8];
            u32x4_t bf[3];
            unitFrag(sX, lane, 0, 16 * ct, v);
            dgk::splitFrag<3>(v, bf);
            f32x4_t m = z4, c = z4;
            mmaSplitB<3>(wf, bf, m, c);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int h = 16 * rt + 4 * g + r;
                sA1[h * SS + 16 * ct + cl] = tanhf(m[r] + c[r] + bf16ToFloat(w[oB1 + h]));
            }
        }
    }
    __syncthreads();
    // ---- layer 2 ----
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) {
        const u32x4_t wf0 = weightFrag(w + oW2, lane, 16 * rt, kH, kH, 0, 1, kH);
        const u32x4_t wf1 = weightFrag(w + oW2, lane, 16 * rt, kH, kH, 32, 1, kH);
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) {
            f32x4_t m = z4, c = z4;
            float v[8];
            u32x4_t bf[3];
            unitFrag(sA1, lane, 0, 16 * ct, v);
            dgk::splitFrag<3>(v, bf);
            mmaSplitB<3>(wf0, bf, m, c);
            unitFrag(sA1, lane, 32, 16 * ct, v);
            dgk::splitFrag<3>(v, bf);
            mmaSplitB<3>(wf1, bf, m, c);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int h = 16 * rt + 4 * g + r;
                sA2[h * SS + 16 * ct + cl] = tanhf(m[r] + c[r] + bf16ToFloat(w[oB2 + h]));
            }
        }
    }
    __syncthreads();
    // ---- layer 3, output error, loss, b3 ----
    double loss = 0.0;
    float gb3[4] = {0.f, 0.f, 0.f, 0.f};
    {
        const u32x4_t wf0 = weightFrag(w + oW3, lane, 0, kH, NOUT, 0, 1, kH);
        const u32x4_t wf1 = weightFrag(w + oW3, lane, 0, kH, NOUT, 32, 1, kH);
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) {
            f32x4_t m = z4, c = z4;
            float v[8];
            u32x4_t bf[3];
            unitFrag(sA2, lane, 0, 16 * ct, v);
            dgk::splitFrag<3>(v, bf);
            mmaSplitB<3>(wf0, bf, m, c);
            unitFrag(sA2, lane, 32, 16 * ct, v);
            dgk::splitFrag<3>(v, bf);
            mmaSplitB<3>(wf1, bf, m, c);
            const int64_t sm = sIdx[16 * ct + cl];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int o = 4 * g + r;
                float d = 0.f;
                if (o < NOUT && sm >= 0) {
                    const float e = (m[r] + c[r] + bf16ToFloat(w[oB3 + o])) - Yt[sm * NOUT + o];
                    loss += static_cast<double>(e) * static_cast<double>(e);
                    d = 2.0f * e * scale;
                }
                sD3[o * SS + 16 * ct + cl] = d;
                gb3[r] += d;
            }
        }
    }
    __syncthreads();
    float* gp = gpart + static_cast<int64_t>(blockIdx.x) * P;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const float t = rowSum16(gb3[r]);
        if (cl == 0 && 4 * g + r < NOUT) gp[oB3 + 4 * g + r] = t;
    }
    // ---- dW3 = D3 A2^T (rows: outputs, columns: hidden 2) ----
    {
        float v[8];
        u32x4_t af[2];
        sampleFrag(sD3, lane, 0, v);
        dgk::splitFrag<2>(v, af);
#pragma unroll
        for (int ht = 0; ht < 4; ++ht) {
            u32x4_t bf[2];
            sampleFrag(sA2, lane, 16 * ht, v);
            dgk::splitFrag<2>(v, bf);
            f32x4_t m = dgk::mfmaBf16(af[0], bf[0], z4);
            f32x4_t c = dgk::mfmaBf16(af[0], bf[1], z4);
            c = dgk::mfmaBf16(af[1], bf[0], c);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int o = 4 * g + r;
                if (o < NOUT) gp[oW3 + o * kH + 16 * ht + cl] = m[r] + c[r];
            }
        }
    }
    // ---- D2 = (W3^T D3) (1 - A2^2), b2 ----
    {
        // A = W3^T: rows hidden 2, k outputs (W3 is [NOUT][kH]: element (h, o) at o * kH + h)
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) {
            const u32x4_t wf = weightFrag(w + oW3, lane, 16 * rt, 1, kH, 0, kH, NOUT);
            float gsum[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ct = 0; ct < 2; ++ct) {
                float v[8];
                u32x4_t bf[2];
                // k = outputs 8(l/16)+j < 32: rows of sD3 beyond 16 read as zero
                {
                    const int o0 = 8 * g;
#pragma unroll
                    for (int j = 0; j < 8; ++j) v[j] = (o0 + j < 16) ? sD3[(o0 + j) * SS + 16 * ct + cl] : 0.f;
                }
                dgk::splitFrag<2>(v, bf);
                f32x4_t m = z4, c = z4;
                mmaSplitB<2>(wf, bf, m, c);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int h = 16 * rt + 4 * g + r;
                    const float a2 = sA2[h * SS + 16 * ct + cl];
                    const float d = (m[r] + c[r]) * (1.0f - a2 * a2);
                    sD2[h * SS + 16 * ct + cl] = d;
                    gsum[r] += d;
                }
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const float t = rowSum16(gsum[r]);
                if (cl == 0) gp[oB2 + 16 * rt + 4 * g + r] = t;
            }
        }
    }
    __syncthreads();
    // ---- dW2 = D2 A1^T ----
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) {
        float v[8];
        u32x4_t af[2];
        sampleFrag(sD2, lane, 16 * rt, v);
        dgk::splitFrag<2>(v, af);
#pragma unroll
        for (int ht = 0; ht < 4; ++ht) {
            u32x4_t bf[2];
            sampleFrag(sA1, lane, 16 * ht, v);
            dgk::splitFrag<2>(v, bf);
            f32x4_t m = dgk::mfmaBf16(af[0], bf[0], z4);
            f32x4_t c = dgk::mfmaBf16(af[0], bf[1], z4);
            c = dgk::mfmaBf16(af[1], bf[0], c);
#pragma unroll
            for (int r = 0; r < 4; ++r) gp[oW2 + (16 * rt + 4 * g + r) * kH + 16 * ht + cl] = m[r] + c[r];
        }
    }
    // ---- D1 = (W2^T D2) (1 - A1^2), b1 ----
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) {
        // A = W2^T: rows hidden 1, k hidden 2 (element (h1, h2) of W2^T at h2 * kH + h1)
        const u32x4_t wf0 = weightFrag(w + oW2, lane, 16 * rt, 1, kH, 0, kH, kH);
        const u32x4_t wf1 = weightFrag(w + oW2, lane, 16 * rt, 1, kH, 32, kH, kH);
        float gsum[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) {
            f32x4_t m = z4, c = z4;
            float v[8];
            u32x4_t bf[2];
            unitFrag(sD2, lane, 0, 16 * ct, v);
            dgk::splitFrag<2>(v, bf);
            mmaSplitB<2>(wf0, bf, m, c);
            unitFrag(sD2, lane, 32, 16 * ct, v);
            dgk::splitFrag<2>(v, bf);
            mmaSplitB<2>(wf1, bf, m, c);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int h = 16 * rt + 4 * g + r;
                const float a1 = sA1[h * SS + 16 * ct + cl];
                const float d = (m[r] + c[r]) * (1.0f - a1 * a1);
                sD1[h * SS + 16 * ct + cl] = d;
                gsum[r] += d;
            }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const float t = rowSum16(gsum[r]);
            if (cl == 0) gp[oB1 + 16 * rt + 4 * g + r] = t;
        }
    }
    __syncthreads();
    // ---- dW1 = D1 X^T ----
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) {
        float v[8];
        u32x4_t af[2];
        sampleFrag(sD1, lane, 16 * rt, v);
        dgk::splitFrag<2>(v, af);
#pragma unroll
        for (int it = 0; it < 2; ++it) {
            if (16 * it >= NIN) break;
            u32x4_t bf[2];
            sampleFrag(sX, lane, 16 * it, v);
            dgk::splitFrag<2>(v, bf);
            f32x4_t m = dgk::mfmaBf16(af[0], bf[0], z4);
            f32x4_t c = dgk::mfmaBf16(af[0], bf[1], z4);
            c = dgk::mfmaBf16(af[1], bf[0], c);
            const int i = 16 * it + cl;
#pragma unroll
            for (int r = 0; r < 4; ++r)
                if (i < NIN) gp[(16 * rt + 4 * g + r) * NIN + i] = m[r] + c[r];
        }
    }
    for (int off = 32; off > 0; off >>= 1) loss += __shfl_xor(loss, off, 64);
    if (lane == 0) lpart[blockIdx.x] = loss;
}

// Sum of the block partials -> grad, and optionally one Adam step on the FP32 master
// parameters with a BF16 copy for the next forward pass. A 256-thread block takes 16
// parameters; its 16 slices of threads sum the partials b = slice, slice + 16, ... in
// ascending order, then the 16 slice sums are added in slice order (a fixed order: the result
// does not depend on timing).
constexpr int kReduceParams = 16, kReduceSlices = 16;
[[maybe_unused]] static __global__ __launch_bounds__(256) void reduceAdamKernel(const float* __restrict__ gpart, int nparts, int P,
                                                        float* __restrict__ grad, float* __restrict__ master,
                                                        float* __restrict__ m1, float* __restrict__ m2,
                                                        uint16_t* __restrict__ wbf, float lr, float beta1, float beta2,
                                                        float eps, float bc1, float bc2, int doUpdate) {
    __shared__ float part[kReduceSlices][kReduceParams];
    const int pl = threadIdx.x % kReduceParams, slice = threadIdx.x / kReduceParams;
    const int x = blockIdx.x * kReduceParams + pl;
    float g = 0.f;
    if (x < P)
        for (int b = slice; b < nparts; b += kReduceSlices) g += gpart[static_cast<int64_t>(b) * P + x];
    part[slice][pl] = g;
    __syncthreads();
    if (slice != 0 || x >= P) return;
    g = 0.f;
#pragma unroll
    for (int k = 0; k < kReduceSlices; ++k) g += part[k][pl];
    if (grad) grad[x] = g;
    if (!doUpdate) return;
    const float a = beta1 * m1[x] + (1.f - beta1) * g;
    const float v = beta2 * m2[x] + (1.f - beta2) * g * g;
    m1[x] = a;
    m2[x] = v;
    const float mh = a / bc1, vh = v / bc2;
    const float p = master[x] - lr * mh / (sqrtf(vh) + eps);
    master[x] = p;
    wbf[x] = floatToBf16(p);
}

}  // namespace dgc
