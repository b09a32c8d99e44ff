// Reference-simplex operators for the nodal DG form of the reference solver.
//
// The reference (src/Mesh.cpp:40-137, :691-751) integrates its weak form with Gmsh
// "Gauss{2p}" rules over the Lagrange basis on Gmsh's equispaced nodes. On straight-sided
// simplices every integrand there has degree <= 2p, so the same semi-discretisation is
// given exactly by nodal operators on that node set:
//   M^-1 S   = sum_r Dw_r G_r            (G_r = sum_x (dxi_r/dx) F_x)
//   M^-1 Fel = sum_f Fscale_f LIFT_f pi_f
// This file builds Dw_r = M^-1 Dr^T M and LIFT = M^-1 E from an orthonormal (PKDO)
// Vandermonde matrix, in FP64, in the coordinates of Gmsh's reference simplex
// (v0 = 0, v1 = e1, v2 = e2, v3 = e3).
#pragma once
#include <array>
#include <cstdint>
#include <vector>

namespace dg {

// Gmsh local face tables (MTriangle edges / MTetrahedron faces). Face f of an element is
// formed by local vertices kFaceVerts[dim][f][0..dim-1]. The reference picks up these
// orderings through gmsh::model::mesh::getElementEdgeNodes/getElementFaceNodes
// (src/Mesh.cpp:178-181); they fix the "first node" used by its BC rule (src/Mesh.cpp:593).
extern const int kFaceVerts2D[3][2];
extern const int kFaceVerts3D[4][3];

struct RefElement {
    int dim = 0;     // 2 (triangle) or 3 (tetrahedron)
    int p = 0;       // polynomial order
    int Np = 0;      // nodes per element
    int Nfp = 0;     // nodes per face
    int Nfaces = 0;  // faces per element
    // Canonical lattice: node n <-> integer barycentric (l0..ld), sum = p, xi_k = l_k / p.
    std::vector<std::array<int, 4>> bary;   // [Np]
    std::vector<double> xi;                 // [Np][3] reference coordinates
    std::vector<int> fmask;                 // [Nfaces][Nfp] element node of face node a
    // Face-lattice order: face node a <-> barycentric (mu_0..mu_{d-1}) w.r.t. the face's
    // vertex order in the local face table; canonical order loops mu_2 (outer) then mu_1.
    std::vector<std::array<int, 3>> fbary;  // [Nfp]
    int nFacePerms = 0;                     // 2 (edges) or 6 (triangles)
    std::vector<int> facePerm;              // [nFacePerms][Nfp]
    std::vector<std::array<int, 3>> permVerts;  // [nFacePerms]: neighbour vertex k = own vertex permVerts[c][k]

    std::vector<double> V;      // [Np][Np] orthonormal Vandermonde (row = node)
    std::vector<double> M;      // [Np][Np] reference mass matrix (xi coordinates)
    std::vector<double> Minv;   // [Np][Np]
    std::vector<double> Dr;     // [dim][Np][Np] strong derivative d/dxi_r (nodal)
    std::vector<double> Dw;     // [dim][Np][Np] weak derivative  M^-1 Dr^T M
    std::vector<double> Mface;  // [Nfp][Nfp] face mass in face parameter coordinates
    std::vector<double> LIFT;   // [Np][Nfaces*Nfp]  M^-1 E
    std::vector<double> Nref;   // [Nfaces][3] area-weighted outward normal / parameter measure
    double faceParamMeasure = 1.0;  // 1 (line [0,1]) or 1/2 (triangle)

    // Lagrange basis values at an arbitrary reference point xi (size Np).
    std::vector<double> basisAt(const double* xi3) const;
};

// Build (and cache) the reference element for a given (dim, p). Supported: dim 2 or 3,
// 1 <= p <= 8.
const RefElement& refElement(int dim, int p);

int numNodes(int dim, int p);

// Small dense helpers (row-major).
bool invertMatrix(std::vector<double>& A, int n);  // in place, partial pivoting
void matmul(const double* A, const double* B, double* C, int n, int k, int m);  // C[n][m] = A[n][k] B[k][m]

}  // namespace dg
