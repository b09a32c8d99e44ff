// Reference-simplex nodal operators (FP64). See refelem.hpp for the contract.
#include "refelem.hpp"

#include <cmath>
#include <map>
#include <mutex>
#include <stdexcept>
#include <utility>

namespace dg {

const int kFaceVerts2D[3][2] = {{0, 1}, {1, 2}, {2, 0}};
const int kFaceVerts3D[4][3] = {{0, 2, 1}, {0, 1, 3}, {0, 3, 2}, {3, 1, 2}};

int numNodes(int dim, int p) {
    if (dim == 2) return (p + 1) * (p + 2) / 2;
    if (dim == 3) return (p + 1) * (p + 2) * (p + 3) / 6;
    return p + 1;
}

bool invertMatrix(std::vector<double>& A, int n) {
    std::vector<double> inv(static_cast<size_t>(n) * n, 0.0);
    for (int i = 0; i < n; ++i) inv[i * n + i] = 1.0;
    for (int c = 0; c < n; ++c) {
        int piv = c;
        double best = std::fabs(A[c * n + c]);
        for (int r = c + 1; r < n; ++r) {
            double v = std::fabs(A[r * n + c]);
            if (v > best) { best = v; piv = r; }
        }
        if (best == 0.0) return false;
        if (piv != c) {
            for (int k = 0; k < n; ++k) {
                std::swap(A[c * n + k], A[piv * n + k]);
                std::swap(inv[c * n + k], inv[piv * n + k]);
            }
        }
        double d = 1.0 / A[c * n + c];
        for (int k = 0; k < n; ++k) { A[c * n + k] *= d; inv[c * n + k] *= d; }
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            double f = A[r * n + c];
            if (f == 0.0) continue;
            for (int k = 0; k < n; ++k) {
                A[r * n + k] -= f * A[c * n + k];
                inv[r * n + k] -= f * inv[c * n + k];
            }
        }
    }
    A.swap(inv);
    return true;
}

void matmul(const double* A, const double* B, double* C, int n, int k, int m) {
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < m; ++j) {
            double s = 0.0;
            for (int l = 0; l < k; ++l) s += A[i * k + l] * B[l * m + j];
            C[i * m + j] = s;
        }
}

namespace {

// Orthonormal Jacobi polynomial P_N^{(a,b)}(x) on [-1,1].
double jacobiP(double x, double a, double b, int N) {
    double g0 = std::pow(2.0, a + b + 1) / (a + b + 1) * std::tgamma(a + 1) * std::tgamma(b + 1) /
                std::tgamma(a + b + 1);
    double p0 = 1.0 / std::sqrt(g0);
    if (N == 0) return p0;
    double g1 = (a + 1) * (b + 1) / (a + b + 3) * g0;
    double p1 = ((a + b + 2) * x / 2 + (a - b) / 2) / std::sqrt(g1);
    if (N == 1) return p1;
    double aold = 2.0 / (2 + a + b) * std::sqrt((a + 1) * (b + 1) / (a + b + 3));
    double pm1 = p0, pc = p1;
    for (int i = 1; i < N; ++i) {
        double h1 = 2 * i + a + b;
        double anew = 2.0 / (h1 + 2) *
                      std::sqrt((i + 1) * (i + 1 + a + b) * (i + 1 + a) * (i + 1 + b) / ((h1 + 1) * (h1 + 3)));
        double bnew = -(a * a - b * b) / (h1 * (h1 + 2));
        double pn = (-aold * pm1 + (x - bnew) * pc) / anew;
        pm1 = pc;
        pc = pn;
        aold = anew;
    }
    return pc;
}

double gradJacobiP(double x, double a, double b, int N) {
    if (N == 0) return 0.0;
    return std::sqrt(N * (N + a + b + 1)) * jacobiP(x, a + 1, b + 1, N - 1);
}

// PKDO basis on the bi-unit triangle (-1,-1),(1,-1),(-1,1).
void simplex2D(double r, double s, int i, int j, double& v, double& vr, double& vs) {
    double a = (std::fabs(s - 1.0) > 1e-14) ? 2.0 * (1.0 + r) / (1.0 - s) - 1.0 : -1.0;
    double b = s;
    double fa = jacobiP(a, 0, 0, i), dfa = gradJacobiP(a, 0, 0, i);
    double gb = jacobiP(b, 2 * i + 1, 0, j), dgb = gradJacobiP(b, 2 * i + 1, 0, j);
    v = std::sqrt(2.0) * fa * gb * std::pow(1.0 - b, i);
    double dr = dfa * gb;
    if (i > 0) dr *= std::pow(0.5 * (1 - b), i - 1);
    double ds = dfa * (gb * (0.5 * (1 + a)));
    if (i > 0) ds *= std::pow(0.5 * (1 - b), i - 1);
    double tmp = dgb * std::pow(0.5 * (1 - b), i);
    if (i > 0) tmp -= 0.5 * i * gb * std::pow(0.5 * (1 - b), i - 1);
    ds += fa * tmp;
    double sc = std::pow(2.0, i + 0.5);
    vr = dr * sc;
    vs = ds * sc;
}

// PKDO basis on the bi-unit tetrahedron.
void simplex3D(double r, double s, double t, int i, int j, int k, double& v, double& vr, double& vs,
               double& vt) {
    double a = (std::fabs(s + t) > 1e-14) ? 2.0 * (1.0 + r) / (-s - t) - 1.0 : -1.0;
    double b = (std::fabs(t - 1.0) > 1e-14) ? 2.0 * (1.0 + s) / (1.0 - t) - 1.0 : -1.0;
    double c = t;
    double fa = jacobiP(a, 0, 0, i), dfa = gradJacobiP(a, 0, 0, i);
    double gb = jacobiP(b, 2 * i + 1, 0, j), dgb = gradJacobiP(b, 2 * i + 1, 0, j);
    double hc = jacobiP(c, 2 * (i + j) + 2, 0, k), dhc = gradJacobiP(c, 2 * (i + j) + 2, 0, k);
    v = 2.0 * std::sqrt(2.0) * fa * gb * std::pow(1.0 - b, i) * hc * std::pow(1.0 - c, i + j);
    double dr = dfa * (gb * hc);
    if (i > 0) dr *= std::pow(0.5 * (1 - b), i - 1);
    if (i + j > 0) dr *= std::pow(0.5 * (1 - c), i + j - 1);
    double ds = 0.5 * (1 + a) * dr;
    double tmp = dgb * std::pow(0.5 * (1 - b), i);
    if (i > 0) tmp += (-0.5 * i) * (gb * std::pow(0.5 * (1 - b), i - 1));
    if (i + j > 0) tmp *= std::pow(0.5 * (1 - c), i + j - 1);
    tmp = fa * (tmp * hc);
    ds += tmp;
    double dt = 0.5 * (1 + a) * dr + 0.5 * (1 + b) * tmp;
    tmp = dhc * std::pow(0.5 * (1 - c), i + j);
    if (i + j > 0) tmp -= 0.5 * (i + j) * (hc * std::pow(0.5 * (1 - c), i + j - 1));
    tmp = fa * (gb * tmp);
    tmp *= std::pow(0.5 * (1 - b), i);
    dt += tmp;
    double sc = std::pow(2.0, 2 * i + j + 1.5);
    vr = dr * sc;
    vs = ds * sc;
    vt = dt * sc;
}

// Orthonormal basis (value + xi-gradient) at a point given in xi coordinates.
void orthoBasis(int dim, int p, const double* xi3, std::vector<double>& val, std::vector<double>& grad) {
    int Np = numNodes(dim, p);
    val.assign(Np, 0.0);
    grad.assign(static_cast<size_t>(Np) * 3, 0.0);
    int m = 0;
    if (dim == 1) {
        double r = 2 * xi3[0] - 1;
        for (int i = 0; i <= p; ++i, ++m) {
            val[m] = jacobiP(r, 0, 0, i);
            grad[m * 3] = 2.0 * gradJacobiP(r, 0, 0, i);
        }
    } else if (dim == 2) {
        double r = 2 * xi3[0] - 1, s = 2 * xi3[1] - 1;
        for (int i = 0; i <= p; ++i)
            for (int j = 0; j <= p - i; ++j, ++m) {
                double v, vr, vs;
                simplex2D(r, s, i, j, v, vr, vs);
                val[m] = v;
                grad[m * 3 + 0] = 2.0 * vr;
                grad[m * 3 + 1] = 2.0 * vs;
            }
    } else {
        double r = 2 * xi3[0] - 1, s = 2 * xi3[1] - 1, t = 2 * xi3[2] - 1;
        for (int i = 0; i <= p; ++i)
            for (int j = 0; j <= p - i; ++j)
                for (int k = 0; k <= p - i - j; ++k, ++m) {
                    double v, vr, vs, vt;
                    simplex3D(r, s, t, i, j, k, v, vr, vs, vt);
                    val[m] = v;
                    grad[m * 3 + 0] = 2.0 * vr;
                    grad[m * 3 + 1] = 2.0 * vs;
                    grad[m * 3 + 2] = 2.0 * vt;
                }
    }
}

// Canonical lattice on a dim-simplex: loops l_dim (outer) ... l_1 (inner), l_0 = p - sum.
std::vector<std::array<int, 4>> lattice(int dim, int p) {
    std::vector<std::array<int, 4>> out;
    if (dim == 1) {
        for (int i = 0; i <= p; ++i) out.push_back({p - i, i, 0, 0});
    } else if (dim == 2) {
        for (int j = 0; j <= p; ++j)
            for (int i = 0; i <= p - j; ++i) out.push_back({p - i - j, i, j, 0});
    } else {
        for (int k = 0; k <= p; ++k)
            for (int j = 0; j <= p - k; ++j)
                for (int i = 0; i <= p - j - k; ++i) out.push_back({p - i - j - k, i, j, k});
    }
    return out;
}

RefElement build(int dim, int p) {
    if (dim < 2 || dim > 3 || p < 1 || p > 8) throw std::invalid_argument("unsupported (dim, p)");
    RefElement R;
    R.dim = dim;
    R.p = p;
    R.Np = numNodes(dim, p);
    R.Nfp = numNodes(dim - 1, p);
    R.Nfaces = dim + 1;
    const int Np = R.Np, Nfp = R.Nfp;
    R.bary = lattice(dim, p);
    R.xi.assign(static_cast<size_t>(Np) * 3, 0.0);
    for (int n = 0; n < Np; ++n)
        for (int k = 0; k < dim; ++k) R.xi[n * 3 + k] = static_cast<double>(R.bary[n][k + 1]) / p;

    // Vandermonde and its xi-gradients.
    R.V.assign(static_cast<size_t>(Np) * Np, 0.0);
    std::vector<double> Vg[3];
    for (int r = 0; r < dim; ++r) Vg[r].assign(static_cast<size_t>(Np) * Np, 0.0);
    std::vector<double> val, grad;
    for (int n = 0; n < Np; ++n) {
        orthoBasis(dim, p, &R.xi[n * 3], val, grad);
        for (int m = 0; m < Np; ++m) {
            R.V[n * Np + m] = val[m];
            for (int r = 0; r < dim; ++r) Vg[r][n * Np + m] = grad[m * 3 + r];
        }
    }
    std::vector<double> Vinv = R.V;
    if (!invertMatrix(Vinv, Np)) throw std::runtime_error("singular Vandermonde");
    // M = (V V^T)^-1 in bi-unit coordinates; xi coordinates scale volumes by 2^-dim.
    std::vector<double> VVt(static_cast<size_t>(Np) * Np);
    for (int i = 0; i < Np; ++i)
        for (int j = 0; j < Np; ++j) {
            double s = 0;
            for (int m = 0; m < Np; ++m) s += R.V[i * Np + m] * R.V[j * Np + m];
            VVt[i * Np + j] = s;
        }
    R.Minv = VVt;
    double vs = std::pow(0.5, dim);
    for (auto& x : R.Minv) x /= vs;  // Minv_xi = (V V^T) / 2^-dim
    R.M = R.Minv;
    if (!invertMatrix(R.M, Np)) throw std::runtime_error("singular mass");

    R.Dr.assign(static_cast<size_t>(dim) * Np * Np, 0.0);
    R.Dw.assign(static_cast<size_t>(dim) * Np * Np, 0.0);
    std::vector<double> tmp(static_cast<size_t>(Np) * Np), DrT(static_cast<size_t>(Np) * Np);
    for (int r = 0; r < dim; ++r) {
        double* D = &R.Dr[static_cast<size_t>(r) * Np * Np];
        matmul(Vg[r].data(), Vinv.data(), D, Np, Np, Np);
        for (int i = 0; i < Np; ++i)
            for (int j = 0; j < Np; ++j) DrT[i * Np + j] = D[j * Np + i];
        // Dw = Minv * Dr^T * M
        matmul(DrT.data(), R.M.data(), tmp.data(), Np, Np, Np);
        matmul(R.Minv.data(), tmp.data(), &R.Dw[static_cast<size_t>(r) * Np * Np], Np, Np, Np);
    }

    // Faces.
    R.fbary.clear();
    for (auto& b : lattice(dim - 1, p)) R.fbary.push_back({b[0], b[1], b[2]});
    R.fmask.assign(static_cast<size_t>(R.Nfaces) * Nfp, -1);
    std::map<std::array<int, 4>, int> index;
    for (int n = 0; n < Np; ++n) index[R.bary[n]] = n;
    for (int f = 0; f < R.Nfaces; ++f) {
        const int* fv = (dim == 2) ? kFaceVerts2D[f] : kFaceVerts3D[f];
        for (int a = 0; a < Nfp; ++a) {
            std::array<int, 4> b = {0, 0, 0, 0};
            for (int k = 0; k < dim; ++k) b[fv[k]] = R.fbary[a][k];
            R.fmask[f * Nfp + a] = index.at(b);
        }
    }
    // Face permutations: neighbour face vertex k = own face vertex permVerts[c][k].
    if (dim == 2) {
        R.permVerts = {{0, 1, 0}, {1, 0, 0}};
    } else {
        R.permVerts = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    }
    R.nFacePerms = static_cast<int>(R.permVerts.size());
    std::map<std::array<int, 3>, int> findex;
    for (int a = 0; a < Nfp; ++a) findex[R.fbary[a]] = a;
    R.facePerm.assign(static_cast<size_t>(R.nFacePerms) * Nfp, -1);
    for (int c = 0; c < R.nFacePerms; ++c)
        for (int a = 0; a < Nfp; ++a) {
            // own barycentric mu[own vertex]; neighbour's k-th vertex is own vertex permVerts[c][k]
            std::array<int, 3> nb = {0, 0, 0};
            for (int k = 0; k < dim; ++k) nb[k] = R.fbary[a][R.permVerts[c][k]];
            R.facePerm[c * Nfp + a] = findex.at(nb);
        }

    // Face mass in face parameter coordinates (line [0,1]: measure 1; triangle: 1/2).
    {
        int fd = dim - 1;
        std::vector<double> Vf(static_cast<size_t>(Nfp) * Nfp);
        for (int a = 0; a < Nfp; ++a) {
            double x3[3] = {0, 0, 0};
            for (int k = 0; k < fd; ++k) x3[k] = static_cast<double>(R.fbary[a][k + 1]) / p;
            orthoBasis(fd, p, x3, val, grad);
            for (int m = 0; m < Nfp; ++m) Vf[a * Nfp + m] = val[m];
        }
        std::vector<double> Mf(static_cast<size_t>(Nfp) * Nfp);
        for (int i = 0; i < Nfp; ++i)
            for (int j = 0; j < Nfp; ++j) {
                double s = 0;
                for (int m = 0; m < Nfp; ++m) s += Vf[i * Nfp + m] * Vf[j * Nfp + m];
                Mf[i * Nfp + j] = s;
            }
        if (!invertMatrix(Mf, Nfp)) throw std::runtime_error("singular face mass");
        double fs = std::pow(0.5, fd);
        for (auto& x : Mf) x *= fs;
        R.Mface = Mf;
        R.faceParamMeasure = (dim == 2) ? 1.0 : 0.5;
    }
    // E and LIFT.
    const int NF = R.Nfaces * Nfp;
    std::vector<double> E(static_cast<size_t>(Np) * NF, 0.0);
    for (int f = 0; f < R.Nfaces; ++f)
        for (int b = 0; b < Nfp; ++b)
            for (int a = 0; a < Nfp; ++a) E[R.fmask[f * Nfp + b] * NF + f * Nfp + a] = R.Mface[b * Nfp + a];
    R.LIFT.assign(static_cast<size_t>(Np) * NF, 0.0);
    matmul(R.Minv.data(), E.data(), R.LIFT.data(), Np, Np, NF);

    // Area-weighted outward reference normals divided by the face parameter measure.
    R.Nref.assign(static_cast<size_t>(R.Nfaces) * 3, 0.0);
    if (dim == 2) {
        const double N2[3][2] = {{0, -1}, {1, 1}, {-1, 0}};
        for (int f = 0; f < 3; ++f) {
            R.Nref[f * 3 + 0] = N2[f][0];
            R.Nref[f * 3 + 1] = N2[f][1];
        }
    } else {
        const double N3[4][3] = {{0, 0, -1}, {0, -1, 0}, {-1, 0, 0}, {1, 1, 1}};
        for (int f = 0; f < 4; ++f)
            for (int k = 0; k < 3; ++k) R.Nref[f * 3 + k] = N3[f][k];
    }
    return R;
}

}  // namespace

std::vector<double> RefElement::basisAt(const double* xi3) const {
    std::vector<double> val, grad;
    orthoBasis(dim, p, xi3, val, grad);
    // phi_n(x) = sum_m Vinv^T ... : phi = V^-T psi
    std::vector<double> Vinv = V;
    invertMatrix(Vinv, Np);
    std::vector<double> out(Np, 0.0);
    for (int n = 0; n < Np; ++n) {
        double s = 0;
        for (int m = 0; m < Np; ++m) s += Vinv[m * Np + n] * val[m];
        out[n] = s;
    }
    return out;
}

const RefElement& refElement(int dim, int p) {
    static std::mutex mu;
    static std::map<std::pair<int, int>, RefElement> cache;
    std::lock_guard<std::mutex> lock(mu);
    auto key = std::make_pair(dim, p);
    auto it = cache.find(key);
    if (it == cache.end()) it = cache.emplace(key, build(dim, p)).first;
    return it->second;
}

}  // namespace dg
