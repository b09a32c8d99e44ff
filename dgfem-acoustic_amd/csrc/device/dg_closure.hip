// Training and inference entry points of the learned absorbing-boundary closure (network
// and feature map described in dg_closure.hpp). No counterpart in the reference solver.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/dgfem_amd.h"
#include "../host/mesh.hpp"
#include "dg_closure.hpp"

namespace {

thread_local std::string g_closureError;

#define CHK(call)                                                                       \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            throw std::runtime_error(std::string(#call) + ": " + hipGetErrorString(e_)); \
        }                                                                               \
    } while (0)

// Device buffers of one call, released in reverse order on scope exit.
struct Scope {
    std::vector<void*> ptrs;
    hipStream_t stream = nullptr;
    explicit Scope(int device) {
        CHK(hipSetDevice(device));
        CHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    }
    template <typename T>
    T* alloc(size_t n) {
        void* p = nullptr;
        CHK(hipMalloc(&p, (n ? n : 1) * sizeof(T)));
        ptrs.push_back(p);
        return static_cast<T*>(p);
    }
    template <typename T>
    T* upload(const T* h, size_t n) {
        T* d = alloc<T>(n);
        if (n) CHK(hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, stream));
        return d;
    }
    ~Scope() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto it = ptrs.rbegin(); it != ptrs.rend(); ++it) (void)hipFree(*it);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

constexpr int kPer = 32;  // samples per gradient block
static_assert(kPer == dgc::kMfmaTile, "one matrix-core tile per gradient block");

// Loss and gradient kernel: the matrix-core one (default) or, with DGFEM_CLOSURE_TRAIN=valu,
// the one-sample-per-wave VALU kernel.
bool trainOnMatrixCores() {
    const char* e = std::getenv("DGFEM_CLOSURE_TRAIN");
    if (!e || !e[0] || std::string(e) == "mfma") return true;
    if (std::string(e) == "valu") return false;
    throw std::runtime_error(std::string("DGFEM_CLOSURE_TRAIN=") + e + " is not a training kernel (mfma, valu)");
}

template <int NIN_, int NOUT_>
struct Net {
    static constexpr int NIN = NIN_, NOUT = NOUT_;
    static constexpr int P = dgc::numParams(NIN, NOUT);
};

// Network shapes built: 2D faces of order 1..5 (n_in = 2 Nfp + 2, Nfp = p + 1) and 3D faces
// of order 1..3 (n_in = 3 Nfp + 2, Nfp = 3, 6, 10). Index into the list, or -1.
constexpr int kShapes[][2] = {{6, 2}, {8, 3}, {10, 4}, {12, 5}, {14, 6}, {11, 3}, {20, 6}, {32, 10}};
int netShape(int nin, int nout) {
    for (int i = 0; i < static_cast<int>(sizeof(kShapes) / sizeof(kShapes[0])); ++i)
        if (kShapes[i][0] == nin && kShapes[i][1] == nout) return i;
    return -1;
}

#define DG_NET_DISPATCH(shape, FN, ...)          \
    switch (shape) {                             \
        case 0: FN<6, 2>(__VA_ARGS__); break;    \
        case 1: FN<8, 3>(__VA_ARGS__); break;    \
        case 2: FN<10, 4>(__VA_ARGS__); break;   \
        case 3: FN<12, 5>(__VA_ARGS__); break;   \
        case 4: FN<14, 6>(__VA_ARGS__); break;   \
        case 5: FN<11, 3>(__VA_ARGS__); break;   \
        case 6: FN<20, 6>(__VA_ARGS__); break;   \
        case 7: FN<32, 10>(__VA_ARGS__); break;  \
        default: break;                          \
    }

const char* kSizeError =
    "closure sizes must be n_in = 2 Nfp + 2 with Nfp in 2..6 (2D) or n_in = 3 Nfp + 2 with Nfp in {3, 6, 10} "
    "(3D), n_out = Nfp";

// Loss (sum of squares, unscaled) and block partial gradients of samples
// [first, first + count) in the order of idx, then their (scaled) reduction.
template <int NIN, int NOUT>
void lossGrad(Scope& s, const uint16_t* w, const float* X, const float* Y, const int64_t* idx, int64_t first,
                int64_t count, float* gpart, double* lpart, float* grad, float* master, float* m1, float* m2,
                uint16_t* wbf, const dg_train_params* prm, int64_t step, bool update, double* lossHost) {
    using N = Net<NIN, NOUT>;
    const int nparts = static_cast<int>((count + kPer - 1) / kPer);
    const float scale = 1.0f / (static_cast<float>(count) * N::NOUT);
    if (trainOnMatrixCores())
        hipLaunchKernelGGL((dgc::gradMfmaKernel<N::NIN, N::NOUT>), dim3(nparts), dim3(64), 0, s.stream, w, X, Y, idx,
                           first, count, scale, gpart, lpart);
    else
        hipLaunchKernelGGL((dgc::gradKernel<N::NIN, N::NOUT>), dim3(nparts), dim3(64), 0, s.stream, w, X, Y, idx,
                           first, count, static_cast<int64_t>(kPer), scale, gpart, lpart);
    float lr = 0, b1 = 0, b2 = 0, eps = 0, bc1 = 1, bc2 = 1;
    if (update) {
        lr = static_cast<float>(prm->lr);
        b1 = static_cast<float>(prm->beta1);
        b2 = static_cast<float>(prm->beta2);
        eps = static_cast<float>(prm->eps);
        bc1 = static_cast<float>(1.0 - std::pow(prm->beta1, static_cast<double>(step)));
        bc2 = static_cast<float>(1.0 - std::pow(prm->beta2, static_cast<double>(step)));
    }
    hipLaunchKernelGGL(dgc::reduceAdamKernel, dim3((N::P + dgc::kReduceParams - 1) / dgc::kReduceParams), dim3(256), 0,
                       s.stream, gpart, nparts, N::P,
                       grad, master, m1, m2, wbf, lr, b1, b2, eps, bc1, bc2, update ? 1 : 0);
    CHK(hipGetLastError());
    if (lossHost) {
        std::vector<double> lp(nparts);
        CHK(hipMemcpyAsync(lp.data(), lpart, nparts * sizeof(double), hipMemcpyDeviceToHost, s.stream));
        CHK(hipStreamSynchronize(s.stream));
        double t = 0;
        for (double v : lp) t += v;
        *lossHost = t;
    }
}

uint64_t splitmix(uint64_t& st) {
    uint64_t z = (st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

template <int NIN, int NOUT>
void forwardImpl(const uint16_t* params, const float* X, int64_t n, float* Y, int device) {
    using N = Net<NIN, NOUT>;
    Scope s(device);
    const uint16_t* w = s.upload(params, N::P);
    const float* dx = s.upload(X, static_cast<size_t>(n) * N::NIN);
    float* dy = s.alloc<float>(static_cast<size_t>(n) * N::NOUT);
    if (n > 0) {
        const int64_t grid = n < 65536 ? n : 65536;
        hipLaunchKernelGGL((dgc::forwardKernel<N::NIN, N::NOUT>), dim3(static_cast<unsigned>(grid)), dim3(64), 0,
                           s.stream, w, dx, dy, n);
        CHK(hipGetLastError());
        CHK(hipMemcpyAsync(Y, dy, static_cast<size_t>(n) * N::NOUT * sizeof(float), hipMemcpyDeviceToHost, s.stream));
    }
    CHK(hipStreamSynchronize(s.stream));
}

template <int NIN, int NOUT>
void lossGradImpl(const uint16_t* params, const float* X, const float* Y, int64_t n, double* loss, float* grad,
                  int device) {
    using N = Net<NIN, NOUT>;
    Scope s(device);
    const uint16_t* w = s.upload(params, N::P);
    const float* dx = s.upload(X, static_cast<size_t>(n) * N::NIN);
    const float* dy = s.upload(Y, static_cast<size_t>(n) * N::NOUT);
    const int64_t nparts = (n + kPer - 1) / kPer;
    float* gpart = s.alloc<float>(static_cast<size_t>(nparts) * N::P);
    double* lpart = s.alloc<double>(nparts);
    float* dg = s.alloc<float>(N::P);
    double lsum = 0;
    lossGrad<NIN, NOUT>(s, w, dx, dy, nullptr, 0, n, gpart, lpart, dg, nullptr, nullptr, nullptr, nullptr, nullptr, 0,
                        false, &lsum);
    *loss = lsum / (static_cast<double>(n) * N::NOUT);
    if (grad) CHK(hipMemcpyAsync(grad, dg, N::P * sizeof(float), hipMemcpyDeviceToHost, s.stream));
    CHK(hipStreamSynchronize(s.stream));
}

template <int NIN, int NOUT>
void trainImpl(float* master, uint16_t* params, const float* X, const float* Y, int64_t n, const dg_train_params* prm,
               double* epochLoss, int device) {
    using N = Net<NIN, NOUT>;
    Scope s(device);
    std::vector<uint16_t> wb(N::P);
    for (int i = 0; i < N::P; ++i) wb[i] = dgc::floatToBf16(master[i]);
    uint16_t* w = s.upload(wb.data(), N::P);
    float* dm = s.upload(master, N::P);
    std::vector<float> zeros(N::P, 0.0f);
    float* m1 = s.upload(zeros.data(), N::P);
    float* m2 = s.upload(zeros.data(), N::P);
    const float* dx = s.upload(X, static_cast<size_t>(n) * N::NIN);
    const float* dy = s.upload(Y, static_cast<size_t>(n) * N::NOUT);
    const int64_t batch = prm->batch < n ? prm->batch : n;
    const int64_t nparts = (batch + kPer - 1) / kPer;
    float* gpart = s.alloc<float>(static_cast<size_t>(nparts) * N::P);
    const int64_t nb = (n + batch - 1) / batch;
    double* lpart = s.alloc<double>(static_cast<size_t>(nb) * nparts);
    int64_t* didx = s.alloc<int64_t>(n);
    std::vector<int64_t> perm(n);
    std::vector<double> lp(static_cast<size_t>(nb) * nparts);
    uint64_t st = prm->seed;
    int64_t step = 0;
    for (int ep = 0; ep < prm->epochs; ++ep) {
        for (int64_t i = 0; i < n; ++i) perm[i] = i;
        for (int64_t i = n - 1; i > 0; --i) {
            const int64_t j = static_cast<int64_t>(splitmix(st) % static_cast<uint64_t>(i + 1));
            std::swap(perm[i], perm[j]);
        }
        CHK(hipMemcpyAsync(didx, perm.data(), n * sizeof(int64_t), hipMemcpyHostToDevice, s.stream));
        for (int64_t b = 0; b < nb; ++b) {
            const int64_t first = b * batch;
            const int64_t count = (first + batch <= n) ? batch : n - first;
            ++step;
            lossGrad<NIN, NOUT>(s, w, dx, dy, didx, first, count, gpart, lpart + b * nparts, nullptr, dm, m1, m2, w, prm,
                                step, true, nullptr);
        }
        // blocks of a short last batch leave their slots of lpart as written in earlier epochs:
        // only the first ceil(count/32) partials of each batch are summed
        CHK(hipMemcpyAsync(lp.data(), lpart, lp.size() * sizeof(double), hipMemcpyDeviceToHost, s.stream));
        CHK(hipStreamSynchronize(s.stream));
        if (epochLoss) {
            double t = 0;
            for (int64_t b = 0; b < nb; ++b) {
                const int64_t count = (b * batch + batch <= n) ? batch : n - b * batch;
                const int64_t np = (count + kPer - 1) / kPer;
                for (int64_t q = 0; q < np; ++q) t += lp[b * nparts + q];
            }
            epochLoss[ep] = t / (static_cast<double>(n) * N::NOUT);
        }
    }
    CHK(hipMemcpyAsync(master, dm, N::P * sizeof(float), hipMemcpyDeviceToHost, s.stream));
    CHK(hipMemcpyAsync(params, w, N::P * sizeof(uint16_t), hipMemcpyDeviceToHost, s.stream));
    CHK(hipStreamSynchronize(s.stream));
}

}  // namespace

extern "C" {

DG_API const char* dg_closure_last_error(void) { return g_closureError.c_str(); }

DG_API int64_t dg_closure_num_params(int32_t n_in, int32_t n_out) {
    return netShape(n_in, n_out) < 0 ? -1 : dgc::numParams(n_in, n_out);
}

DG_API void dg_closure_to_bf16(const float* x, int64_t n, uint16_t* out) {
    for (int64_t i = 0; i < n; ++i) out[i] = dgc::floatToBf16(x[i]);
}

DG_API dg_status dg_closure_forward(int32_t n_in, int32_t n_out, const uint16_t* params, const float* X, int64_t n,
                                    float* Y, int32_t device) {
    if (!params || !X || !Y || n < 0) return DG_ERR_INVALID;
    const int shape = netShape(n_in, n_out);
    if (shape < 0) {
        g_closureError = kSizeError;
        return DG_ERR_UNSUPPORTED;
    }
    try {
        DG_NET_DISPATCH(shape, forwardImpl, params, X, n, Y, device)
    } catch (const std::exception& ex) {
        g_closureError = ex.what();
        return DG_ERR_DEVICE;
    }
    return DG_OK;
}

DG_API dg_status dg_closure_loss_grad(int32_t n_in, int32_t n_out, const uint16_t* params, const float* X,
                                      const float* Y, int64_t n, double* loss, float* grad, int32_t device) {
    if (!params || !X || !Y || n <= 0 || !loss) return DG_ERR_INVALID;
    const int shape = netShape(n_in, n_out);
    if (shape < 0) {
        g_closureError = kSizeError;
        return DG_ERR_UNSUPPORTED;
    }
    try {
        DG_NET_DISPATCH(shape, lossGradImpl, params, X, Y, n, loss, grad, device)
    } catch (const std::exception& ex) {
        g_closureError = ex.what();
        return DG_ERR_DEVICE;
    }
    return DG_OK;
}

DG_API dg_status dg_closure_train(int32_t n_in, int32_t n_out, float* master, uint16_t* params, const float* X,
                                  const float* Y, int64_t n, const dg_train_params* prm, double* epoch_loss,
                                  int32_t device) {
    if (!master || !params || !X || !Y || n <= 0 || !prm || prm->epochs < 0 || prm->batch <= 0) return DG_ERR_INVALID;
    const int shape = netShape(n_in, n_out);
    if (shape < 0) {
        g_closureError = kSizeError;
        return DG_ERR_UNSUPPORTED;
    }
    try {
        DG_NET_DISPATCH(shape, trainImpl, master, params, X, Y, n, prm, epoch_loss, device)
    } catch (const std::exception& ex) {
        g_closureError = ex.what();
        return DG_ERR_DEVICE;
    }
    return DG_OK;
}
DG_API dg_status dg_closure_features(int32_t dim, int32_t nfp, int32_t nf, int64_t n, const double* traces,
                                     const double* normals, const dg_physics* ph, float* X, float* Y, float* S) {
    if ((dim != 2 && dim != 3) || nfp < 2 || nfp > 10 || (nf != 4 && nf != 5) || n < 0 || !traces || !normals || !ph ||
        !X)
        return DG_ERR_INVALID;
    const int nt = dim - 1, nin = dim * nfp + 2, nout = nfp;
    const double irc = 1.0 / (ph->rho0 * ph->c0);
    std::vector<double> wo(nfp), tt(static_cast<size_t>(nt) * nfp);
    for (int64_t i = 0; i < n; ++i) {
        const double* nv = normals + 3 * i;
        double nn[3] = {nv[0], nv[1], dim == 3 ? nv[2] : 0.0}, t[2][3] = {{0, 0, 0}, {0, 0, 0}};
        double vn0 = 0.0, vt0 = 0.0;
        for (int c = 0; c < dim; ++c) vn0 += ph->v0[c] * nn[c];
        if (dim == 2) {
            t[0][0] = -nn[1];
            t[0][1] = nn[0];
            vt0 = ph->v0[0] * t[0][0] + ph->v0[1] * t[0][1];
        } else {
            // t1 along the tangential mean flow, else along the axis least aligned with n
            double u[3];
            for (int c = 0; c < 3; ++c) u[c] = ph->v0[c] - vn0 * nn[c];
            vt0 = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
            if (!(vt0 > 1e-6 * ph->c0)) {
                const double ax = std::fabs(nn[0]), ay = std::fabs(nn[1]), az = std::fabs(nn[2]);
                const int k = (ax <= ay && ax <= az) ? 0 : (ay <= az ? 1 : 2);
                for (int c = 0; c < 3; ++c) u[c] = (c == k ? 1.0 : 0.0) - nn[k] * nn[c];
                vt0 = 0.0;
            }
            const double iu = 1.0 / std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
            for (int c = 0; c < 3; ++c) t[0][c] = u[c] * iu;
            t[1][0] = nn[1] * t[0][2] - nn[2] * t[0][1];
            t[1][1] = nn[2] * t[0][0] - nn[0] * t[0][2];
            t[1][2] = nn[0] * t[0][1] - nn[1] * t[0][0];
        }
        float ss = 0.f;
        for (int j = 0; j < nfp; ++j) {
            const double* q = traces + ((i * 2) * nfp + j) * nf;
            double vn = 0.0;
            for (int c = 0; c < dim; ++c) vn += q[1 + c] * nn[c];
            wo[j] = q[0] * irc + vn;
            const float a = static_cast<float>(wo[j]);
            ss += a * a;
            for (int k = 0; k < nt; ++k) {
                double v = 0.0;
                for (int c = 0; c < dim; ++c) v += q[1 + c] * t[k][c];
                tt[static_cast<size_t>(k) * nfp + j] = v;
                const float b = static_cast<float>(v);
                ss += b * b;
            }
        }
        const float s = std::sqrt(ss / nfp) + 1e-30f;
        if (S) S[i] = s;
        float* x = X + i * nin;
        for (int j = 0; j < nfp; ++j) {
            x[j] = static_cast<float>(wo[j]) / s;
            for (int k = 0; k < nt; ++k)
                x[(1 + k) * nfp + j] = static_cast<float>(tt[static_cast<size_t>(k) * nfp + j]) / s;
        }
        x[dim * nfp] = static_cast<float>(vn0 / ph->c0);
        x[dim * nfp + 1] = static_cast<float>(vt0 / ph->c0);
        if (Y) {
            // target: the incoming characteristic of the neighbour's trace
            float* y = Y + i * nout;
            for (int j = 0; j < nfp; ++j) {
                const double* qe = traces + ((i * 2 + 1) * nfp + j) * nf;
                double vn = 0.0;
                for (int c = 0; c < dim; ++c) vn += qe[1 + c] * nn[c];
                y[j] = static_cast<float>((qe[0] * irc - vn) / s);
            }
        }
    }
    return DG_OK;
}

}  // extern "C"
