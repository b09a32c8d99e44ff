/*
 * CPU oracle (test infrastructure only): restatement of the reference solver's algorithm.
 * See dg_oracle.h. Reference citations are given per function.
 */
#define _GNU_SOURCE
#include "dg_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXNP 120

static char g_err[512];
const char* oracle_error(void) { return g_err; }
static void set_err(const char* msg) { snprintf(g_err, sizeof g_err, "%s", msg); }

/* Gmsh local face tables (tri edges / tet faces); see src/Mesh.cpp:178-181. */
static const int FV2[3][2] = {{0, 1}, {1, 2}, {2, 0}};
static const int FV3[4][3] = {{0, 2, 1}, {0, 1, 3}, {0, 3, 2}, {3, 1, 2}};

static int face_vert(int dim, int f, int k) { return dim == 2 ? FV2[f][k] : FV3[f][k]; }

struct oracle_ctx {
    int dim, p, Np, Nfp, Nfc;
    int64_t K;
    oracle_physics ph;
    /* per element */
    double* coords;     /* [K][Np][3] caller order */
    int* canon;         /* [K][Np] caller position -> lattice index */
    int* pos;           /* [K][nlat] lattice index -> caller position */
    double* Jinv;       /* [K][9]  dxi_r/dx_c */
    double* detJ;       /* [K] */
    double* bary;       /* [K][3] barycenter of vertices */
    double* Minv;       /* [K][Np][Np] inverse mass (caller order) */
    /* lattice */
    int nlat;
    int (*lat)[4];      /* [Np] barycentric counts */
    int (*flat)[3];     /* [Nfp] face barycentric counts */
    /* volume quadrature (reference coordinates) */
    int G;
    double* qx;         /* [G][3] */
    double* qw;         /* [G] */
    double* phi;        /* [G][Np] lattice order */
    double* dphi;       /* [G][Np][3] xi-gradients, lattice order */
    /* face quadrature (face parameter coordinates) */
    int Gf;
    double* fqw;        /* [Gf] */
    double* psi;        /* [Gf][Nfp] */
    /* unique faces */
    int64_t F;
    int64_t* fEl;       /* [F][2] up, dn (-1) */
    int* fLoc;          /* [F][2] local face index in up / dn */
    int* fNode;         /* [F][Nfp][2] caller node position in up / dn */
    double* fN;         /* [F][3] unit normal (outward of up; outward on the boundary) */
    double* fT;         /* [F][3] tangent */
    double* fB;         /* [F][3] bitangent */
    double* fDet;       /* [F] face Jacobian determinant (parameter measure -> physical) */
    int8_t* fBnd;       /* [F] 1 on the boundary */
    int8_t* fBC;        /* [F] 0 absorbing, 1 reflecting */
    double* RKR;        /* [F][16] */
    int64_t* elF;       /* [K][Nfc] unique face id */
    int* elFO;          /* [K][Nfc] orientation s(e,f) */
    /* work arrays */
    double* Flux;       /* [4][K*Np][3] */
    double* ghostF;     /* [4][F][Gf] normal flux at boundary quadrature points */
    double* fFlux;      /* [F][Nfp] */
};

/* ---------------------------------------------------------------- quadrature ---- */
/* Gauss-Legendre nodes/weights on [0,1] (Newton on P_n). */
static void gauss_legendre01(int n, double* x, double* w) {
    for (int i = 0; i < n; ++i) {
        double z = cos(M_PI * (i + 0.75) / (n + 0.5)), pp = 0;
        for (int it = 0; it < 100; ++it) {
            double p1 = 1, p2 = 0;
            for (int j = 1; j <= n; ++j) {
                double p3 = p2;
                p2 = p1;
                p1 = ((2.0 * j - 1) * z * p2 - (j - 1.0) * p3) / j;
            }
            pp = n * (z * p1 - p2) / (z * z - 1);
            double z1 = z;
            z = z1 - p1 / pp;
            if (fabs(z - z1) < 1e-16) break;
        }
        x[i] = 0.5 * (1 - z);
        w[i] = 1.0 / ((1 - z * z) * pp * pp);  /* 2/((1-z^2)pp^2) on [-1,1], halved */
    }
}

/* Collapsed (Duffy) product rule on the reference simplex of dimension d with
   n0 x n1 (x n2) Gauss-Legendre points: exact for the degree-2p integrands here. */
static int simplex_rule(int d, int p, double** X, double** W) {
    int n[3] = {p + 1, p + 1, p + 2};
    double gx[3][16], gw[3][16];
    for (int k = 0; k < d; ++k) gauss_legendre01(n[k], gx[k], gw[k]);
    int G = 1;
    for (int k = 0; k < d; ++k) G *= n[k];
    *X = calloc((size_t)G * 3, sizeof(double));
    *W = calloc((size_t)G, sizeof(double));
    int g = 0;
    if (d == 1) {
        for (int a = 0; a < n[0]; ++a, ++g) { (*X)[g * 3] = gx[0][a]; (*W)[g] = gw[0][a]; }
    } else if (d == 2) {
        for (int b = 0; b < n[1]; ++b)
            for (int a = 0; a < n[0]; ++a, ++g) {
                double u = gx[0][a], v = gx[1][b];
                (*X)[g * 3 + 0] = u * (1 - v);
                (*X)[g * 3 + 1] = v;
                (*W)[g] = gw[0][a] * gw[1][b] * (1 - v);
            }
    } else {
        for (int c = 0; c < n[2]; ++c)
            for (int b = 0; b < n[1]; ++b)
                for (int a = 0; a < n[0]; ++a, ++g) {
                    double u = gx[0][a], v = gx[1][b], w = gx[2][c];
                    (*X)[g * 3 + 0] = u * (1 - v) * (1 - w);
                    (*X)[g * 3 + 1] = v * (1 - w);
                    (*X)[g * 3 + 2] = w;
                    (*W)[g] = gw[0][a] * gw[1][b] * gw[2][c] * (1 - v) * (1 - w) * (1 - w);
                }
    }
    return G;
}

/* ------------------------------------------------------------- Lagrange basis ---- */
/* Silvester's product formula for the equispaced Lagrange basis on a d-simplex:
   phi_a(l) = prod_k prod_{m<a_k} (p l_k - m)/(m+1). */
static double silvester(int d, int p, const int* a, const double* l) {
    double v = 1;
    for (int k = 0; k <= d; ++k)
        for (int m = 0; m < a[k]; ++m) v *= (p * l[k] - m) / (m + 1.0);
    return v;
}
static double silvester_dl(int d, int p, const int* a, const double* l, int kd) {
    double rest = 1;
    for (int k = 0; k <= d; ++k) {
        if (k == kd) continue;
        for (int m = 0; m < a[k]; ++m) rest *= (p * l[k] - m) / (m + 1.0);
    }
    double s = 0;
    for (int m = 0; m < a[kd]; ++m) {
        double t = p / (m + 1.0);
        for (int mm = 0; mm < a[kd]; ++mm)
            if (mm != m) t *= (p * l[kd] - mm) / (mm + 1.0);
        s += t;
    }
    return s * rest;
}

static int build_lattice(int d, int p, int (*out)[4]) {
    int n = 0;
    if (d == 1) {
        for (int i = 0; i <= p; ++i) { out[n][0] = p - i; out[n][1] = i; out[n][2] = 0; out[n][3] = 0; ++n; }
    } else if (d == 2) {
        for (int j = 0; j <= p; ++j)
            for (int i = 0; i <= p - j; ++i) { out[n][0] = p - i - j; out[n][1] = i; out[n][2] = j; out[n][3] = 0; ++n; }
    } else {
        for (int k = 0; k <= p; ++k)
            for (int j = 0; j <= p - k; ++j)
                for (int i = 0; i <= p - j - k; ++i) {
                    out[n][0] = p - i - j - k; out[n][1] = i; out[n][2] = j; out[n][3] = k; ++n;
                }
    }
    return n;
}

static int lattice_index(int d, int p, const int (*lat)[4], int n, const int* b) {
    (void)p;
    for (int i = 0; i < n; ++i) {
        int ok = 1;
        for (int k = 0; k <= d; ++k)
            if (lat[i][k] != b[k]) ok = 0;
        if (ok) return i;
    }
    return -1;
}

/* Gauss-Jordan inverse with partial pivoting (Eigen's role in src/utils.cpp:391-396). */
static int invert(double* A, int n) {
    double* inv = calloc((size_t)n * n, sizeof(double));
    for (int i = 0; i < n; ++i) inv[i * n + i] = 1;
    for (int c = 0; c < n; ++c) {
        int piv = c;
        double best = fabs(A[c * n + c]);
        for (int r = c + 1; r < n; ++r)
            if (fabs(A[r * n + c]) > best) { best = fabs(A[r * n + c]); piv = r; }
        if (best == 0) { free(inv); return -1; }
        if (piv != c)
            for (int k = 0; k < n; ++k) {
                double t = A[c * n + k]; A[c * n + k] = A[piv * n + k]; A[piv * n + k] = t;
                t = inv[c * n + k]; inv[c * n + k] = inv[piv * n + k]; inv[piv * n + k] = t;
            }
        double dd = 1.0 / A[c * n + c];
        for (int k = 0; k < n; ++k) { A[c * n + k] *= dd; inv[c * n + k] *= dd; }
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            double f = A[r * n + c];
            if (f == 0) continue;
            for (int k = 0; k < n; ++k) { A[r * n + k] -= f * A[c * n + k]; inv[r * n + k] -= f * inv[c * n + k]; }
        }
    }
    memcpy(A, inv, sizeof(double) * n * n);
    free(inv);
    return 0;
}

/* ------------------------------------------------------------------ faces ---- */
typedef struct { int64_t k[3]; int64_t ef; } fkey;
static int fkey_cmp(const void* a, const void* b) {
    const fkey* x = a;
    const fkey* y = b;
    for (int i = 0; i < 3; ++i)
        if (x->k[i] != y->k[i]) return x->k[i] < y->k[i] ? -1 : 1;
    return x->ef < y->ef ? -1 : (x->ef > y->ef);
}
static void sort3(int64_t* v, int n) {
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j)
            if (v[j] < v[i]) { int64_t t = v[i]; v[i] = v[j]; v[j] = t; }
}

static int in_sorted(const int64_t* a, int64_t n, int64_t x) {
    int64_t lo = 0, hi = n - 1;
    while (lo <= hi) {
        int64_t m = (lo + hi) / 2;
        if (a[m] == x) return 1;
        if (a[m] < x) lo = m + 1; else hi = m - 1;
    }
    return 0;
}

static void normalize3(double* v) {
    double n = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (n > 0) { v[0] /= n; v[1] /= n; v[2] /= n; }
}

/* ------------------------------------------------------------ boundary fluxes ---- */
/* R*K*R^-1 of the absorbing condition at one point (src/Mesh.cpp:628-663). */
static void rkr_matrix(const oracle_physics* ph, const double* n, const double* t, const double* s, double* R) {
    double c0 = ph->c0, rho0 = ph->rho0;
    double vn0 = ph->v0[0] * n[0] + ph->v0[1] * n[1] + ph->v0[2] * n[2];
    double lambda = (vn0 < 0) ? 0 : -1;
    R[0] = 0.25 * (c0 + vn0);
    for (int q = 0; q < 3; ++q) R[1 + q] = 0.25 * (c0 * rho0 * (c0 + vn0) * n[q]);
    for (int r = 0; r < 3; ++r) {
        R[(r + 1) * 4] = 0.25 * (n[r] * (c0 + vn0) / (rho0 * c0));
        for (int q = 0; q < 3; ++q)
            R[(r + 1) * 4 + 1 + q] = 0.25 * ((c0 + vn0) * n[r] * n[q] - vn0 * lambda * (t[q] * t[r] + s[q] * s[r]));
    }
}

/* Reflecting (rigid wall) flux n . F(p, v - (v.n)n) at one point (src/Mesh.cpp:911-949). */
static void reflect_flux(const oracle_physics* ph, const double* n, const double* q, double* out) {
    const double* v0 = ph->v0;
    double rho0 = ph->rho0, c0 = ph->c0;
    double ug[4] = {q[0], q[1], q[2], q[3]};
    double dot = n[0] * ug[1] + n[1] * ug[2] + n[2] * ug[3];
    ug[1] -= dot * n[0]; ug[2] -= dot * n[1]; ug[3] -= dot * n[2];
    double FG[4][3] = {
        {v0[0] * ug[0] + rho0 * c0 * c0 * ug[1], v0[1] * ug[0] + rho0 * c0 * c0 * ug[2], v0[2] * ug[0] + rho0 * c0 * c0 * ug[3]},
        {v0[0] * ug[1] + ug[0] / rho0, v0[1] * ug[1], v0[2] * ug[1]},
        {v0[0] * ug[2], v0[1] * ug[2] + ug[0] / rho0, v0[2] * ug[2]},
        {v0[0] * ug[3], v0[1] * ug[3], v0[2] * ug[3] + ug[0] / rho0}};
    for (int k = 0; k < 4; ++k) out[k] = n[0] * FG[k][0] + n[1] * FG[k][1] + n[2] * FG[k][2];
}

void oracle_bc_flux(const oracle_physics* ph, int bc, const double n[3], const double t[3], const double s[3],
                    const double q[4], double out[4]) {
    if (bc == 1) {
        reflect_flux(ph, n, q, out);
    } else {
        double R[16];
        rkr_matrix(ph, n, t, s, R);
        for (int k = 0; k < 4; ++k) out[k] = R[k * 4] * q[0] + R[k * 4 + 1] * q[1] + R[k * 4 + 2] * q[2] + R[k * 4 + 3] * q[3];
    }
}

/* -------------------------------------------------------------- construction ---- */
/* Mesh precompute: src/Mesh.cpp:20-686 (geometry, faces, orientation, BCs, RKR) and
   precomputeMassMatrix / getElMassMatrix (src/Mesh.cpp:691-724). */
oracle_ctx* oracle_create(const oracle_mesh* m, const oracle_physics* ph) {
    if (!m || !ph || (m->dim != 2 && m->dim != 3) || m->order < 1 || m->order > 8) {
        set_err("invalid mesh description");
        return NULL;
    }
    oracle_ctx* c = calloc(1, sizeof *c);
    int d = m->dim, p = m->order;
    c->dim = d; c->p = p; c->K = m->num_elements; c->ph = *ph;
    c->Nfc = d + 1;
    c->lat = calloc(MAXNP, sizeof *c->lat);
    c->Np = build_lattice(d, p, c->lat);
    c->nlat = c->Np;
    int (*fl)[4] = calloc(MAXNP, sizeof *fl);
    c->Nfp = build_lattice(d - 1, p, fl);
    c->flat = calloc(c->Nfp, sizeof *c->flat);
    for (int a = 0; a < c->Nfp; ++a) for (int k = 0; k < 3; ++k) c->flat[a][k] = fl[a][k];
    free(fl);
    if (m->nodes_per_element != c->Np) { set_err("nodes_per_element does not match order"); oracle_destroy(c); return NULL; }
    const int Np = c->Np, Nfp = c->Nfp, Nfc = c->Nfc;
    const int64_t K = c->K;

    /* reference quadrature and basis tables */
    c->G = simplex_rule(d, p, &c->qx, &c->qw);
    c->phi = calloc((size_t)c->G * Np, sizeof(double));
    c->dphi = calloc((size_t)c->G * Np * 3, sizeof(double));
    for (int g = 0; g < c->G; ++g) {
        double l[4] = {1, 0, 0, 0};
        for (int k = 0; k < d; ++k) { l[k + 1] = c->qx[g * 3 + k]; l[0] -= l[k + 1]; }
        for (int n = 0; n < Np; ++n) {
            c->phi[g * Np + n] = silvester(d, p, c->lat[n], l);
            double d0 = silvester_dl(d, p, c->lat[n], l, 0);
            for (int k = 0; k < d; ++k)
                c->dphi[(g * Np + n) * 3 + k] = silvester_dl(d, p, c->lat[n], l, k + 1) - d0;
        }
    }
    double* fqx;
    c->Gf = simplex_rule(d - 1, p, &fqx, &c->fqw);
    c->psi = calloc((size_t)c->Gf * Nfp, sizeof(double));
    for (int g = 0; g < c->Gf; ++g) {
        double l[4] = {1, 0, 0, 0};
        for (int k = 0; k < d - 1; ++k) { l[k + 1] = fqx[g * 3 + k]; l[0] -= l[k + 1]; }
        for (int a = 0; a < Nfp; ++a) {
            int aa[4] = {c->flat[a][0], c->flat[a][1], c->flat[a][2], 0};
            c->psi[g * Nfp + a] = silvester(d - 1, p, aa, l);
        }
    }
    free(fqx);

    /* elements: Jacobians, lattice identification, mass matrices */
    c->coords = malloc(sizeof(double) * K * Np * 3);
    memcpy(c->coords, m->node_coords, sizeof(double) * K * Np * 3);
    c->canon = calloc((size_t)K * Np, sizeof(int));
    c->pos = calloc((size_t)K * Np, sizeof(int));
    c->Jinv = calloc((size_t)K * 9, sizeof(double));
    c->detJ = calloc((size_t)K, sizeof(double));
    c->bary = calloc((size_t)K * 3, sizeof(double));
    c->Minv = calloc((size_t)K * Np * Np, sizeof(double));
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(+ : bad)
    for (int64_t e = 0; e < K; ++e) {
        const double* X = c->coords + e * Np * 3;
        double J[3][3] = {{0}}, Ji[3][3] = {{0}}, det;
        for (int r = 0; r < d; ++r)
            for (int q = 0; q < 3; ++q) J[q][r] = X[(r + 1) * 3 + q] - X[q];
        if (d == 2) {
            det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
            Ji[0][0] = J[1][1] / det; Ji[0][1] = -J[0][1] / det;
            Ji[1][0] = -J[1][0] / det; Ji[1][1] = J[0][0] / det;
        } else {
            double A[9];
            for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) A[i * 3 + j] = J[i][j];
            det = A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) + A[2] * (A[3] * A[7] - A[4] * A[6]);
            if (invert(A, 3) != 0) { bad++; continue; }
            for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) Ji[i][j] = A[i * 3 + j];
        }
        if (!(fabs(det) > 0)) { bad++; continue; }
        c->detJ[e] = fabs(det);
        for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) c->Jinv[e * 9 + i * 3 + j] = Ji[i][j];
        for (int q = 0; q < 3; ++q) {
            double s = 0;
            for (int v = 0; v <= d; ++v) s += X[v * 3 + q];
            c->bary[e * 3 + q] = s / (d + 1);
        }
        for (int a = 0; a < Np; ++a) {
            double l[4] = {1, 0, 0, 0};
            int b[4] = {0, 0, 0, 0};
            for (int r = 0; r < d; ++r) {
                double xi = 0;
                for (int q = 0; q < d; ++q) xi += Ji[r][q] * (X[a * 3 + q] - X[q]);
                l[r + 1] = xi;
                l[0] -= xi;
            }
            for (int k = 0; k <= d; ++k) b[k] = (int)lround(l[k] * p);
            int li = lattice_index(d, p, (const int(*)[4])c->lat, Np, b);
            if (li < 0) { bad++; break; }
            c->canon[e * Np + a] = li;
            c->pos[e * Np + li] = a;
        }
        /* mass matrix M_ij = sum_g phi_i phi_j w_g detJ  (src/Mesh.cpp:708-724) */
        double* M = c->Minv + e * Np * Np;
        for (int i = 0; i < Np; ++i)
            for (int j = 0; j < Np; ++j) {
                double s = 0;
                int ci = c->canon[e * Np + i], cj = c->canon[e * Np + j];
                for (int g = 0; g < c->G; ++g) s += c->phi[g * Np + ci] * c->phi[g * Np + cj] * c->qw[g] * c->detJ[e];
                M[i * Np + j] = s;
            }
        if (invert(M, Np) != 0) bad++;
    }
    if (bad) { set_err("degenerate or non-lattice element"); oracle_destroy(c); return NULL; }

    /* unique faces; up = lowest element index (src/Mesh.cpp:981-1130, :491-509) */
    fkey* keys = malloc(sizeof(fkey) * K * Nfc);
    for (int64_t e = 0; e < K; ++e)
        for (int f = 0; f < Nfc; ++f) {
            fkey* k = &keys[e * Nfc + f];
            k->k[0] = k->k[1] = k->k[2] = -1;
            for (int v = 0; v < d; ++v) k->k[v] = m->elem_verts[e * (d + 1) + face_vert(d, f, v)];
            sort3(k->k, d);
            k->ef = e * Nfc + f;
        }
    qsort(keys, (size_t)(K * Nfc), sizeof(fkey), fkey_cmp);
    int64_t F = 0;
    for (int64_t i = 0; i < K * Nfc;) {
        int64_t j = i + 1;
        while (j < K * Nfc && memcmp(keys[j].k, keys[i].k, sizeof keys[i].k) == 0) ++j;
        if (j - i > 2) { set_err("non-manifold face"); free(keys); oracle_destroy(c); return NULL; }
        ++F;
        i = j;
    }
    c->F = F;
    c->fEl = malloc(sizeof(int64_t) * F * 2);
    c->fLoc = malloc(sizeof(int) * F * 2);
    c->fNode = malloc(sizeof(int) * F * Nfp * 2);
    c->fN = calloc((size_t)F * 3, sizeof(double));
    c->fT = calloc((size_t)F * 3, sizeof(double));
    c->fB = calloc((size_t)F * 3, sizeof(double));
    c->fDet = calloc((size_t)F, sizeof(double));
    c->fBnd = calloc((size_t)F, 1);
    c->fBC = calloc((size_t)F, 1);
    c->RKR = calloc((size_t)F * 16, sizeof(double));
    c->elF = malloc(sizeof(int64_t) * K * Nfc);
    c->elFO = malloc(sizeof(int) * K * Nfc);
    int64_t fi = 0;
    for (int64_t i = 0; i < K * Nfc; ++fi) {
        int64_t j = i + 1;
        while (j < K * Nfc && memcmp(keys[j].k, keys[i].k, sizeof keys[i].k) == 0) ++j;
        c->fEl[fi * 2] = keys[i].ef / Nfc;
        c->fLoc[fi * 2] = (int)(keys[i].ef % Nfc);
        c->fEl[fi * 2 + 1] = (j - i == 2) ? keys[i + 1].ef / Nfc : -1;
        c->fLoc[fi * 2 + 1] = (j - i == 2) ? (int)(keys[i + 1].ef % Nfc) : -1;
        c->fBnd[fi] = (j - i == 1);
        for (int64_t q = i; q < j; ++q) c->elF[keys[q].ef] = fi;
        i = j;
    }
    free(keys);

    int failed = 0;
#pragma omp parallel for schedule(static) reduction(+ : failed)
    for (int64_t f = 0; f < F; ++f) {
        int64_t up = c->fEl[f * 2], dn = c->fEl[f * 2 + 1];
        int lf = c->fLoc[f * 2];
        int64_t gv[3];
        double P[3][3];
        for (int v = 0; v < d; ++v) {
            int lv = face_vert(d, lf, v);
            gv[v] = m->elem_verts[up * (d + 1) + lv];
            for (int q = 0; q < 3; ++q) P[v][q] = c->coords[(up * Np + lv) * 3 + q];
        }
        /* face nodes in the owner's face order and their positions in both elements */
        for (int a = 0; a < Nfp; ++a) {
            int b[4] = {0, 0, 0, 0};
            for (int v = 0; v < d; ++v) b[face_vert(d, lf, v)] = c->flat[a][v];
            int li = lattice_index(d, p, (const int(*)[4])c->lat, Np, b);
            c->fNode[(f * Nfp + a) * 2] = c->pos[up * Np + li];
            c->fNode[(f * Nfp + a) * 2 + 1] = -1;
            if (dn >= 0) {
                int bd[4] = {0, 0, 0, 0};
                for (int v = 0; v < d; ++v) {
                    int found = -1;
                    for (int w = 0; w <= d; ++w)
                        if (m->elem_verts[dn * (d + 1) + w] == gv[v]) found = w;
                    if (found < 0) { failed++; break; }
                    bd[found] = c->flat[a][v];
                }
                int ld = lattice_index(d, p, (const int(*)[4])c->lat, Np, bd);
                if (ld < 0) { failed++; continue; }
                c->fNode[(f * Nfp + a) * 2 + 1] = c->pos[dn * Np + ld];
            }
        }
        /* normal, tangents, face Jacobian (src/Mesh.cpp:273-364) */
        double n[3] = {0, 0, 0}, t[3] = {0, 0, 0}, s[3] = {0, 0, 0};
        if (d == 2) {
            double tx = P[1][0] - P[0][0], ty = P[1][1] - P[0][1];
            n[0] = ty; n[1] = -tx;
            c->fDet[f] = sqrt(tx * tx + ty * ty);
        } else {
            double a1[3], a2[3];
            for (int q = 0; q < 3; ++q) { a1[q] = P[1][q] - P[0][q]; a2[q] = P[2][q] - P[0][q]; }
            n[0] = a1[1] * a2[2] - a1[2] * a2[1];
            n[1] = a1[2] * a2[0] - a1[0] * a2[2];
            n[2] = a1[0] * a2[1] - a1[1] * a2[0];
            c->fDet[f] = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);  /* 2 * area */
        }
        normalize3(n);
        /* Flux-sign contract: the face normal points out of the owner (sigma = +1). */
        double fc[3] = {0, 0, 0};
        for (int v = 0; v < d; ++v) for (int q = 0; q < 3; ++q) fc[q] += P[v][q] / d;
        double dot = 0;
        for (int q = 0; q < 3; ++q) dot += n[q] * (fc[q] - c->bary[up * 3 + q]);
        if (dot < 0) for (int q = 0; q < 3; ++q) n[q] = -n[q];
        if (d == 2) {
            t[0] = -n[1]; t[1] = n[0]; t[2] = 0;
        } else {
            for (int q = 0; q < 3; ++q) t[q] = P[1][q] - P[0][q];
            normalize3(t);
            s[0] = n[1] * t[2] - n[2] * t[1];
            s[1] = n[2] * t[0] - n[0] * t[2];
            s[2] = n[0] * t[1] - n[1] * t[0];
            normalize3(s);
        }
        memcpy(c->fN + f * 3, n, sizeof n);
        memcpy(c->fT + f * 3, t, sizeof t);
        memcpy(c->fB + f * 3, s, sizeof s);
    }
    if (failed) { set_err("face node matching failed"); oracle_destroy(c); return NULL; }

    /* orientation s(e,f) = sign(n . (x_facenode0 - barycenter)) (src/Mesh.cpp:452-477);
       boundary faces keep their (outward) normal and s = 1 (src/Mesh.cpp:536-568). */
    for (int64_t e = 0; e < K; ++e)
        for (int lf = 0; lf < Nfc; ++lf) {
            int64_t f = c->elF[e * Nfc + lf];
            int lv0 = face_vert(d, lf, 0);
            double dot = 0;
            for (int q = 0; q < 3; ++q) dot += (c->coords[(e * Np + lv0) * 3 + q] - c->bary[e * 3 + q]) * c->fN[f * 3 + q];
            c->elFO[e * Nfc + lf] = (dot >= 0) ? 1 : -1;
            if (c->fBnd[f]) c->elFO[e * Nfc + lf] = 1;
        }

    /* BC type by the first-node rule (src/Mesh.cpp:579-605) */
    for (int64_t f = 0; f < F; ++f) {
        if (!c->fBnd[f]) continue;
        int64_t up = c->fEl[f * 2];
        int64_t first = m->elem_verts[up * (d + 1) + face_vert(d, c->fLoc[f * 2], 0)];
        int8_t bc = 0;
        for (int g = 0; g < m->num_groups; ++g) {
            int64_t o0 = m->group_offsets[g], o1 = m->group_offsets[g + 1];
            if (in_sorted(m->group_nodes + o0, o1 - o0, first)) bc = m->group_type[g];
        }
        c->fBC[f] = bc;
    }

    /* RKR absorbing matrices (src/Mesh.cpp:619-666) */
    for (int64_t f = 0; f < F; ++f) {
        if (!c->fBnd[f] || c->fBC[f] != 0) continue;
        rkr_matrix(ph, c->fN + f * 3, c->fT + f * 3, c->fB + f * 3, c->RKR + f * 16);
    }

    c->Flux = calloc((size_t)4 * K * Np * 3, sizeof(double));
    c->ghostF = calloc((size_t)4 * F * c->Gf, sizeof(double));
    c->fFlux = calloc((size_t)F * Nfp, sizeof(double));
    return c;
}

void oracle_destroy(oracle_ctx* c) {
    if (!c) return;
    free(c->coords); free(c->canon); free(c->pos); free(c->Jinv); free(c->detJ); free(c->bary);
    free(c->Minv); free(c->lat); free(c->flat); free(c->qx); free(c->qw); free(c->phi); free(c->dphi);
    free(c->fqw); free(c->psi); free(c->fEl); free(c->fLoc); free(c->fNode); free(c->fN); free(c->fT);
    free(c->fB); free(c->fDet); free(c->fBnd); free(c->fBC); free(c->RKR); free(c->elF); free(c->elFO);
    free(c->Flux); free(c->ghostF); free(c->fFlux);
    free(c);
}

int oracle_num_volume_points(const oracle_ctx* c) { return c->G; }
int oracle_num_face_points(const oracle_ctx* c) { return c->Gf; }
int64_t oracle_num_faces(const oracle_ctx* c) { return c->F; }

/* Boundary type of the local face lf of element e (src/Mesh.cpp:579-605): 0 absorbing,
   1 reflecting, -1 interior. */
int oracle_face_bc(const oracle_ctx* c, int64_t e, int lf) {
    if (e < 0 || e >= c->K || lf < 0 || lf >= c->Nfc) return -2;
    const int64_t f = c->elF[e * c->Nfc + lf];
    return c->fBnd[f] ? c->fBC[f] : -1;
}

/* ------------------------------------------------------------ RHS operators ---- */
/* Mesh::updateFlux (src/Mesh.cpp:851-975): nodal physical fluxes and boundary ghosts. */
static void update_flux(oracle_ctx* c, const double* u) {
    const int Np = c->Np, Nfp = c->Nfp, Gf = c->Gf;
    const int64_t K = c->K, N = K * Np, F = c->F;
    const double* v0 = c->ph.v0;
    const double rho0 = c->ph.rho0, c0 = c->ph.c0;
    const double *p = u, *vx = u + N, *vy = u + 2 * N, *vz = u + 3 * N;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
        double* F0 = c->Flux + (0 * N + i) * 3;
        double* F1 = c->Flux + (1 * N + i) * 3;
        double* F2 = c->Flux + (2 * N + i) * 3;
        double* F3 = c->Flux + (3 * N + i) * 3;
        F0[0] = v0[0] * p[i] + rho0 * c0 * c0 * vx[i];
        F0[1] = v0[1] * p[i] + rho0 * c0 * c0 * vy[i];
        F0[2] = v0[2] * p[i] + rho0 * c0 * c0 * vz[i];
        F1[0] = v0[0] * vx[i] + p[i] / rho0; F1[1] = v0[1] * vx[i]; F1[2] = v0[2] * vx[i];
        F2[0] = v0[0] * vy[i]; F2[1] = v0[1] * vy[i] + p[i] / rho0; F2[2] = v0[2] * vy[i];
        F3[0] = v0[0] * vz[i]; F3[1] = v0[1] * vz[i]; F3[2] = v0[2] * vz[i] + p[i] / rho0;
    }
#pragma omp parallel for schedule(static)
    for (int64_t f = 0; f < F; ++f) {
        if (!c->fBnd[f]) continue;
        int64_t el = c->fEl[f * 2];
        const double* n = c->fN + f * 3;
        for (int g = 0; g < Gf; ++g) {
            double ug[4] = {0, 0, 0, 0};
            for (int a = 0; a < Nfp; ++a) {
                int64_t nid = el * Np + c->fNode[(f * Nfp + a) * 2];
                double w = c->psi[g * Nfp + a];
                for (int q = 0; q < 4; ++q) ug[q] += u[q * N + nid] * w;
            }
            double out[4];
            if (c->fBC[f] == 1) {
                reflect_flux(&c->ph, n, ug, out);
            } else {
                const double* R = c->RKR + f * 16;
                for (int q = 0; q < 4; ++q)
                    out[q] = R[q * 4] * ug[0] + R[q * 4 + 1] * ug[1] + R[q * 4 + 2] * ug[2] + R[q * 4 + 3] * ug[3];
            }
            for (int q = 0; q < 4; ++q) c->ghostF[(q * F + f) * Gf + g] = out[q];
        }
    }
}

/* Mesh::precomputeFlux (src/Mesh.cpp:761-813), race-free: face integrals of the numerical
   flux for one field. */
static void precompute_flux(oracle_ctx* c, const double* ueq, const double* Feq, int eq) {
    const int Np = c->Np, Nfp = c->Nfp, Gf = c->Gf;
    const int64_t F = c->F;
    const double c0 = c->ph.c0;
    const int fcoef = 1;
#pragma omp parallel for schedule(static)
    for (int64_t f = 0; f < F; ++f) {
        double FI[Gf];
        for (int g = 0; g < Gf; ++g) FI[g] = 0;
        const double* n = c->fN + f * 3;
        if (c->fBnd[f]) {
            for (int g = 0; g < Gf; ++g) FI[g] = c->ghostF[((int64_t)eq * F + f) * Gf + g];
        } else {
            int64_t up = c->fEl[f * 2], dn = c->fEl[f * 2 + 1];
            for (int i = 0; i < Nfp; ++i) {
                int64_t iu = up * Np + c->fNode[(f * Nfp + i) * 2];
                int64_t id = dn * Np + c->fNode[(f * Nfp + i) * 2 + 1];
                for (int g = 0; g < Gf; ++g) {
                    double Fnum[3];
                    for (int x = 0; x < 3; ++x)
                        Fnum[x] = 0.5 * ((Feq[iu * 3 + x] + Feq[id * 3 + x]) + fcoef * c0 * n[x] * (ueq[iu] - ueq[id]));
                    FI[g] += (n[0] * Fnum[0] + n[1] * Fnum[1] + n[2] * Fnum[2]) * c->psi[g * Nfp + i];
                }
            }
        }
        for (int a = 0; a < Nfp; ++a) {
            double s = 0;
            for (int g = 0; g < Gf; ++g) s += c->fqw[g] * c->psi[g * Nfp + a] * FI[g] * c->fDet[f];
            c->fFlux[f * Nfp + a] = s;
        }
    }
}

/* solver::numStep (src/solver.cpp:38-57) with getElFlux (src/Mesh.cpp:822-839),
   getElStiffVector (src/Mesh.cpp:734-751), eigen::minus and eigen::linEq
   (src/utils.cpp:406-427): k = beta*k + dt*Minv*(S - F). */
static void num_step(oracle_ctx* c, double* k, double dt, double beta) {
    const int Np = c->Np, Nfp = c->Nfp, Nfc = c->Nfc, G = c->G, d = c->dim;
    const int64_t K = c->K, N = K * Np;
    for (int eq = 0; eq < 4; ++eq) {
        double* keq = k + eq * N;
        const double* Feq = c->Flux + (int64_t)eq * N * 3;
        precompute_flux(c, keq, Feq, eq);
#pragma omp parallel for schedule(static)
        for (int64_t el = 0; el < K; ++el) {
            double Fel[MAXNP], S[MAXNP], gp[G * 3];
            for (int i = 0; i < Np; ++i) Fel[i] = 0;
            for (int lf = 0; lf < Nfc; ++lf) {
                int64_t f = c->elF[el * Nfc + lf];
                int side = (c->fEl[f * 2] == el) ? 0 : 1;
                for (int nf = 0; nf < Nfp; ++nf)
                    Fel[c->fNode[(f * Nfp + nf) * 2 + side]] += c->elFO[el * Nfc + lf] * c->fFlux[f * Nfp + nf];
            }
            const double* Ji = c->Jinv + el * 9;
            const double detJ = c->detJ[el];
            for (int i = 0; i < Np; ++i) {
                double s = 0;
                int ci = c->canon[el * Np + i];
                for (int g = 0; g < G; ++g) {
                    /* physical gradient of phi_i at g: J^-T grad_xi phi_i */
                    const double* dg = c->dphi + ((int64_t)g * Np + ci) * 3;
                    for (int x = 0; x < 3; ++x) {
                        double v = 0;
                        if (x < d)
                            for (int r = 0; r < d; ++r) v += Ji[r * 3 + x] * dg[r];
                        gp[g * 3 + x] = v;
                    }
                }
                for (int j = 0; j < Np; ++j) {
                    int64_t jId = el * Np + j;
                    int cj = c->canon[jId];
                    const double* Fj = Feq + jId * 3;
                    for (int g = 0; g < G; ++g)
                        s += (Fj[0] * gp[g * 3] + Fj[1] * gp[g * 3 + 1] + Fj[2] * gp[g * 3 + 2]) * c->phi[g * Np + cj] * c->qw[g] * detJ;
                }
                S[i] = s - Fel[i];
            }
            const double* Mi = c->Minv + el * Np * Np;
            double* ke = keq + el * Np;
            double out[MAXNP];
            for (int i = 0; i < Np; ++i) {
                double s = 0;
                for (int j = 0; j < Np; ++j) s += Mi[i * Np + j] * S[j];
                out[i] = beta * ke[i] + dt * s;
            }
            for (int i = 0; i < Np; ++i) ke[i] = out[i];
        }
    }
}

int oracle_rhs(oracle_ctx* c, const double* u, double* k, double dt, int num_threads) {
#ifdef _OPENMP
    if (num_threads > 0) omp_set_num_threads(num_threads);
#else
    (void)num_threads;
#endif
    const int64_t N = c->K * c->Np;
    memcpy(k, u, sizeof(double) * 4 * N);
    update_flux(c, k);
    num_step(c, k, dt, 0.0);
    return 0;
}

/* ---------------------------------------------------------------- time loop ---- */
/* solver::rungeKutta (src/solver.cpp:311-584) and solver::forwardEuler (:66-302). */
int oracle_run(oracle_ctx* c, double* u, const oracle_run_params* prm, oracle_run_output* out) {
#ifdef _OPENMP
    if (prm->num_threads > 0) omp_set_num_threads(prm->num_threads);
#endif
    const int Np = c->Np;
    const int64_t K = c->K, N = K * Np;
    const double c0 = c->ph.c0;
    const int ns = prm->num_sources, no = prm->num_observers;
    /* source node sets: ||x - xs||^2 < size^2 (src/solver.cpp:340-358) */
    int64_t** src = calloc(ns > 0 ? ns : 1, sizeof(int64_t*));
    int64_t* nsrc = calloc(ns > 0 ? ns : 1, sizeof(int64_t));
    for (int s = 0; s < ns; ++s) {
        const oracle_source* S = &prm->sources[s];
        src[s] = malloc(sizeof(int64_t) * (N > 0 ? N : 1));
        for (int64_t n = 0; n < N; ++n) {
            const double* x = c->coords + n * 3;
            if (pow(x[0] - S->x, 2) + pow(x[1] - S->y, 2) + pow(x[2] - S->z, 2) < pow(S->size, 2)) src[s][nsrc[s]++] = n;
        }
    }
    /* observer node sets: distance < size, IDW weights (src/solver.cpp:361-383, :552-573) */
    int64_t** obs = calloc(no > 0 ? no : 1, sizeof(int64_t*));
    double** obsd = calloc(no > 0 ? no : 1, sizeof(double*));
    int64_t* nobs = calloc(no > 0 ? no : 1, sizeof(int64_t));
    for (int o = 0; o < no; ++o) {
        const oracle_observer* O = &prm->observers[o];
        obs[o] = malloc(sizeof(int64_t) * (N > 0 ? N : 1));
        obsd[o] = malloc(sizeof(double) * (N > 0 ? N : 1));
        for (int64_t n = 0; n < N; ++n) {
            const double* x = c->coords + n * 3;
            double dist = sqrt(pow(x[0] - O->x, 2) + pow(x[1] - O->y, 2) + pow(x[2] - O->z, 2));
            if (dist < O->size) { obs[o][nobs[o]] = n; obsd[o][nobs[o]] = dist; nobs[o]++; }
        }
    }
    double* snap = malloc(sizeof(double) * 4 * N);
    double* k1 = malloc(sizeof(double) * 4 * N);
    double *k2 = NULL, *k3 = NULL, *k4 = NULL;
    if (prm->integrator == 0) {
        k2 = malloc(sizeof(double) * 4 * N);
        k3 = malloc(sizeof(double) * 4 * N);
        k4 = malloc(sizeof(double) * 4 * N);
    }
    int64_t istep = 0;
    out->num_saves = 0;
    for (double t = prm->t_start, step = 0, tDisplay = 0; t <= prm->t_end; t += prm->dt, tDisplay += prm->dt, ++step) {
        if (istep >= prm->max_steps) break;
        if (tDisplay >= prm->t_rate || step == 0) {
            tDisplay = 0;
            memcpy(snap, u, sizeof(double) * 4 * N);
            if (out->save_steps) out->save_steps[out->num_saves] = istep;
            out->num_saves++;
        }
        for (int s = 0; s < ns; ++s) {
            const oracle_source* S = &prm->sources[s];
            if (S->kind == 0) {
                if (t < S->duration)
                    for (int64_t q = 0; q < nsrc[s]; ++q) u[src[s][q]] = S->amp * sin(2 * M_PI * S->freq * t + S->phase);
            } else if (prm->value_fn) {
                double v;
                if (prm->value_fn(prm->value_user, s, t, &v))
                    for (int64_t q = 0; q < nsrc[s]; ++q) u[src[s][q]] = v;
            }
        }
        if (prm->integrator == 0) {
            memcpy(k1, u, sizeof(double) * 4 * N);
            memcpy(k2, u, sizeof(double) * 4 * N);
            memcpy(k3, u, sizeof(double) * 4 * N);
            memcpy(k4, u, sizeof(double) * 4 * N);
            update_flux(c, k1);
            num_step(c, k1, prm->dt, 0);
            for (int64_t i = 0; i < 4 * N; ++i) k2[i] += 0.5 * k1[i];
            update_flux(c, k2);
            num_step(c, k2, prm->dt, 0);
            for (int64_t i = 0; i < 4 * N; ++i) k3[i] += 0.5 * k2[i];
            update_flux(c, k3);
            num_step(c, k3, prm->dt, 0);
            for (int64_t i = 0; i < 4 * N; ++i) k4[i] += 1.0 * k3[i];
            update_flux(c, k4);
            num_step(c, k4, prm->dt, 0);
            for (int64_t i = 0; i < 4 * N; ++i) u[i] += (k1[i] + 2 * k2[i] + 2 * k3[i] + k4[i]) / 6.0;
        } else {
            update_flux(c, u);
            num_step(c, u, prm->dt, 1.0);
        }
        /* residuals vs snapshot (src/solver.cpp:518-547) */
        double r[5] = {0, 0, 0, 0, 0};
        for (int64_t i = 0; i < N; ++i) {
            r[0] += pow(snap[i] - u[i], 2);
            r[1] += pow(snap[i] / (c0 * c0) - u[i] / (c0 * c0), 2);
            r[2] += pow(snap[N + i] - u[N + i], 2);
            r[3] += pow(snap[2 * N + i] - u[2 * N + i], 2);
            r[4] += pow(snap[3 * N + i] - u[3 * N + i], 2);
        }
        if (out->times) out->times[istep] = t;
        if (out->residuals)
            for (int q = 0; q < 5; ++q) out->residuals[istep * 5 + q] = r[q] / (double)N;
        if (out->observers)
            for (int o = 0; o < no; ++o) {
                double pv[4] = {0, 0, 0, 0}, ws = 0;
                for (int64_t q = 0; q < nobs[o]; ++q) {
                    double w = 1.0 / (pow(obsd[o][q], 2) + 1.0e-12);
                    for (int f = 0; f < 4; ++f) pv[f] += u[f * N + obs[o][q]] * w;
                    ws += w;
                }
                for (int f = 0; f < 4; ++f) out->observers[(istep * no + o) * 4 + f] = pv[f] / ws;
            }
        istep++;
    }
    out->steps = istep;
    for (int s = 0; s < ns; ++s) free(src[s]);
    for (int o = 0; o < no; ++o) { free(obs[o]); free(obsd[o]); }
    free(src); free(nsrc); free(obs); free(obsd); free(nobs);
    free(snap); free(k1); free(k2); free(k3); free(k4);
    return 0;
}
