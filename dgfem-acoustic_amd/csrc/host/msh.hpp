// Gmsh MSH 4.0 / 4.1 ASCII reader for the solver's meshes (the reference opens them through
// the gmsh API: gmsh::open, src/configParser.cpp:197/260, and reads the elements of the first
// element type of the top dimension, src/Mesh.cpp:40-47).
#pragma once
#include <array>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

namespace dg {

struct PhysicalGroup {
    int dim = 0;
    int tag = 0;
    std::string name;
    std::vector<int64_t> nodeTags;  // sorted node tags of the group's elements
};

struct MshMesh {
    int dim = 0;                         // top dimension
    int elementType = 0;                 // Gmsh element type number of the solver elements
    int order = 0;
    int nodesPerElement = 0;
    std::vector<int64_t> elementTags;    // [K]
    std::vector<int64_t> elementNodes;   // [K][Np] node tags, file order
    std::map<int64_t, std::array<double, 3>> nodes;
    std::vector<PhysicalGroup> groups;   // sorted by (dim, tag), as gmsh lists them
    int64_t numElements() const { return static_cast<int64_t>(elementTags.size()); }
    // Coordinates [K][Np][3] and vertex ids [K][dim+1] (the first dim+1 nodes).
    void solverArrays(std::vector<double>& coords, std::vector<int64_t>& verts) const;
};

// Throws std::runtime_error with a message for missing files / unsupported content.
MshMesh readMsh(const std::string& path);

// Gmsh element type -> (dimension, order) for the simplices and lines used here.
bool gmshElementInfo(int type, int& dim, int& order, int& numNodes);

}  // namespace dg
