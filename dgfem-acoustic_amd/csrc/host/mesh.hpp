// Host-side mesh ingest and FP64 precompute for the device-resident DG path.
//
// Mirrors the reference's Mesh construction (src/Mesh.cpp:20-686) for straight-sided
// simplices: element Jacobians, face pairing, face-node correspondence, outward normals,
// boundary flags and the reference's BC assignment rule. Face pairing uses sorted vertex
// keys (O(F log F)) instead of the reference's O(F^2) searches (src/Mesh.cpp:981-1130);
// the pairing itself is the same relation.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "refelem.hpp"

namespace dg {

// Deterministic generator: SplitMix64.
struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return static_cast<double>(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
};

// A straight-sided simplex mesh as the solver receives it: vertices, element->vertex
// connectivity (the first dim+1 nodes of each element) and the DG node coordinates of
// every element in the caller's node order.
struct SimplexMesh {
    int dim = 0;
    int order = 0;
    int64_t K = 0;
    int Np = 0;
    std::vector<double> vertexCoords;     // [nv][3]
    std::vector<int64_t> elemVerts;       // [K][dim+1]
    std::vector<double> nodeCoords;       // [K][Np][3] (caller order)
    std::vector<int8_t> faceSide;         // [K][dim+1] generator boundary side id, -1 interior
};

// Jittered structured generators on [lo,hi]^dim (seeded, reproducible).
//  2D: n x n quads, each split along a Bernoulli(1/2) diagonal; interior vertices jittered
//      by U(-jitter*h, jitter*h) per coordinate; counter-clockwise triangles.
//  3D: n^3 hexes, each split into 6 Kuhn tetrahedra sharing the (0,0,0)-(1,1,1) diagonal;
//      interior vertices jittered likewise; positively oriented.
// Nodes are placed on the canonical equispaced lattice (see refelem.hpp).
SimplexMesh generateBoxMesh(int dim, int n, int order, double jitter, uint64_t seed, double lo, double hi);

// The reference's BC rule (src/Mesh.cpp:579-605): a boundary face is absorbing (0) by
// default; for each physical group in ascending tag order, a boundary face whose FIRST
// node (node 0 of its local face, Gmsh face tables) belongs to the group's node set gets
// the group's type. Groups are given as masks over generator sides; the group's node set
// is every vertex of every boundary face on those sides.
std::vector<int8_t> resolveBoundaryTypes(const SimplexMesh& m, const std::vector<uint32_t>& groupSideMask,
                                         const std::vector<int8_t>& groupType);

// Same rule for explicit node-tag sets (mesh files): groupNodes[g] sorted vertex ids.
std::vector<int8_t> resolveBoundaryTypesFromNodeSets(int dim, int64_t K, const int64_t* elemVerts,
                                                     const std::vector<uint8_t>& isBoundaryFace,
                                                     const std::vector<std::vector<int64_t>>& groupNodes,
                                                     const std::vector<int8_t>& groupType);

// 1 for every (element, local face) that no other element shares (src/Mesh.cpp:536-568).
std::vector<uint8_t> boundaryFaceFlags(int dim, int64_t K, const int64_t* elemVerts);

// Device-ready FP64 precompute in canonical node order.
struct DGMesh {
    int dim = 0, p = 0, Np = 0, Nfp = 0, Nfaces = 0;
    int64_t K = 0;
    std::vector<double> nodes;       // [K][Np][3] canonical order
    std::vector<int32_t> toCanon;    // [K][Np]: caller position -> canonical index
    std::vector<double> Jinv;        // [K][3][3] dxi_r/dx_c (row r, column c); 2D uses the 2x2 block
    std::vector<double> detJ;        // [K]
    std::vector<double> inradius;    // [K]
    std::vector<int32_t> nbr;        // [K][Nfaces]: neighbour element; boundary: -1 absorbing,
                                     // -2 reflecting, -3-i learned-closure face i
    std::vector<int64_t> learnedFaces;  // element * Nfaces + local face of each learned-closure face
    std::vector<uint8_t> nbrCode;    // [K][Nfaces]: neighbour local face | (perm code << 2)
    std::vector<uint8_t> isBoundary; // [K][Nfaces]
    int64_t numFaces = 0, numBoundaryFaces = 0;
};

// Builds the precompute. faceBC (may be empty => all absorbing) is per (element, local
// face) and only read on boundary faces. Throws std::runtime_error with a message on
// malformed input (non-conforming faces, curved/non-equispaced nodes, degenerate elements).
DGMesh buildDGMesh(int dim, int order, int64_t K, const double* nodeCoords, const int64_t* elemVerts,
                   const int8_t* faceBC);

// Outward unit normal and Fscale = sJ/|detJ| of local face f of element e.
void faceGeometry(const DGMesh& m, int64_t e, int f, double n[3], double& fscale);

}  // namespace dg
