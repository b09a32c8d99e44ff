// Host-only entry points of the C ABI (mesh generation and the BC rule).
#include <cstring>
#include <stdexcept>

#include "../../../include/dgfem_amd.h"
#include "mesh.hpp"

extern "C" {

dg_status dg_box_mesh_size(int32_t dim, int32_t n, int32_t order, int64_t* num_elements, int32_t* nodes_per_element) {
    if ((dim != 2 && dim != 3) || n < 1 || order < 1 || order > 8) return DG_ERR_INVALID;
    if (num_elements) *num_elements = dim == 2 ? int64_t(2) * n * n : int64_t(6) * n * n * n;
    if (nodes_per_element) *nodes_per_element = dg::numNodes(dim, order);
    return DG_OK;
}

dg_status dg_box_mesh_fill(int32_t dim, int32_t n, int32_t order, double jitter, uint64_t seed, double lo, double hi,
                           double* node_coords, int64_t* element_vertices, int8_t* face_side) {
    try {
        dg::SimplexMesh m = dg::generateBoxMesh(dim, n, order, jitter, seed, lo, hi);
        if (node_coords) std::memcpy(node_coords, m.nodeCoords.data(), m.nodeCoords.size() * sizeof(double));
        if (element_vertices)
            std::memcpy(element_vertices, m.elemVerts.data(), m.elemVerts.size() * sizeof(int64_t));
        if (face_side) std::memcpy(face_side, m.faceSide.data(), m.faceSide.size());
    } catch (const std::exception&) {
        return DG_ERR_INVALID;
    }
    return DG_OK;
}

dg_status dg_resolve_bc_sides(int32_t dim, int64_t num_elements, const int64_t* element_vertices,
                              const int8_t* face_side, int32_t num_groups, const uint32_t* group_side_mask,
                              const int8_t* group_type, int8_t* face_bc) {
    if ((dim != 2 && dim != 3) || !element_vertices || !face_side || !face_bc || num_groups < 0) return DG_ERR_INVALID;
    dg::SimplexMesh m;
    m.dim = dim;
    m.K = num_elements;
    m.elemVerts.assign(element_vertices, element_vertices + num_elements * (dim + 1));
    m.faceSide.assign(face_side, face_side + num_elements * (dim + 1));
    std::vector<uint32_t> masks(group_side_mask, group_side_mask + num_groups);
    std::vector<int8_t> types(group_type, group_type + num_groups);
    std::vector<int8_t> bc = dg::resolveBoundaryTypes(m, masks, types);
    std::memcpy(face_bc, bc.data(), bc.size());
    return DG_OK;
}

dg_status dg_ref_operators(int32_t dim, int32_t order, double* Dw, double* LIFT, int32_t* fmask, double* Nref,
                           int32_t* lattice) {
    try {
        const dg::RefElement& R = dg::refElement(dim, order);
        if (Dw) std::memcpy(Dw, R.Dw.data(), R.Dw.size() * sizeof(double));
        if (LIFT) std::memcpy(LIFT, R.LIFT.data(), R.LIFT.size() * sizeof(double));
        if (fmask)
            for (size_t i = 0; i < R.fmask.size(); ++i) fmask[i] = R.fmask[i];
        if (Nref) std::memcpy(Nref, R.Nref.data(), R.Nref.size() * sizeof(double));
        if (lattice)
            for (int n = 0; n < R.Np; ++n)
                for (int k = 0; k < 4; ++k) lattice[n * 4 + k] = R.bary[n][k];
    } catch (const std::exception&) {
        return DG_ERR_UNSUPPORTED;
    }
    return DG_OK;
}

dg_status dg_ref_mass(int32_t dim, int32_t order, double* M) {
    if (!M) return DG_ERR_INVALID;
    try {
        const dg::RefElement& R = dg::refElement(dim, order);
        std::memcpy(M, R.M.data(), R.M.size() * sizeof(double));
    } catch (const std::exception&) {
        return DG_ERR_UNSUPPORTED;
    }
    return DG_OK;
}

dg_status dg_mesh_precompute(const dg_mesh_desc* m, double* Jinv, int32_t* nbr, uint8_t* code, int32_t* face_perm,
                             int32_t* to_canonical) {
    if (!m || !m->node_coords || !m->element_vertices) return DG_ERR_INVALID;
    try {
        dg::DGMesh g = dg::buildDGMesh(m->dim, m->order, m->num_elements, m->node_coords, m->element_vertices,
                                       m->face_bc);
        if (Jinv) std::memcpy(Jinv, g.Jinv.data(), g.Jinv.size() * sizeof(double));
        if (nbr) std::memcpy(nbr, g.nbr.data(), g.nbr.size() * sizeof(int32_t));
        if (code) std::memcpy(code, g.nbrCode.data(), g.nbrCode.size());
        if (to_canonical) std::memcpy(to_canonical, g.toCanon.data(), g.toCanon.size() * sizeof(int32_t));
        if (face_perm) {
            const dg::RefElement& R = dg::refElement(m->dim, m->order);
            for (size_t i = 0; i < R.facePerm.size(); ++i) face_perm[i] = R.facePerm[i];
        }
    } catch (const std::exception&) {
        return DG_ERR_MESH;
    }
    return DG_OK;
}

}  // extern "C"

// ---- front-end helpers ----
#include <sstream>

#include "config.hpp"
#include "formula.hpp"
#include "sources_io.hpp"

// Meshes with at most this many boundary faces list each of them in the case summary.
constexpr int64_t kSummaryMaxBoundaryFaces = 4096;

extern "C" {

double dg_eval_formula(const char* expression, double t) {
    return expression ? dg::evalFormula(expression, t) : -1.0;
}

dg_status dg_read_source_table(const char* path, double* rows, int64_t capacity, int64_t* num_rows) {
    if (!path) return DG_ERR_INVALID;
    try {
        const std::string p(path);
        const std::string ext = p.substr(p.find_last_of('.') + 1);
        std::vector<std::vector<double>> t =
            ext == "wav" ? dg::parseWavTable(p) : dg::parseCsvTable(p, ';');
        if (num_rows) *num_rows = static_cast<int64_t>(t.size());
        for (int64_t i = 0; rows && i < capacity && i < static_cast<int64_t>(t.size()); ++i) {
            rows[2 * i] = t[i].size() > 0 ? t[i][0] : 0.0;
            rows[2 * i + 1] = t[i].size() > 1 ? t[i][1] : 0.0;
        }
    } catch (const std::exception&) {
        return DG_ERR_INVALID;
    }
    return DG_OK;
}

dg_status dg_case_summary(const char* config_path, char* out, int64_t capacity) {
    if (!config_path || !out || capacity < 2) return DG_ERR_INVALID;
    std::ostringstream js;
    dg_status st = DG_OK;
    try {
        dg::Case cs = dg::loadCase(config_path);
        const dg::Config& c = cs.config;
        std::vector<int8_t> bc = dg::boundaryTypes(cs);
        std::vector<int64_t> verts(cs.mesh.numElements() * (cs.mesh.dim + 1));
        for (int64_t e = 0; e < cs.mesh.numElements(); ++e)
            for (int v = 0; v <= cs.mesh.dim; ++v)
                verts[e * (cs.mesh.dim + 1) + v] = cs.mesh.elementNodes[e * cs.mesh.nodesPerElement + v];
        std::vector<uint8_t> isB = dg::boundaryFaceFlags(cs.mesh.dim, cs.mesh.numElements(), verts.data());
        int64_t nAbs = 0, nRef = 0;
        for (size_t i = 0; i < isB.size(); ++i)
            if (isB[i]) (bc[i] == 1 ? nRef : nAbs)++;
        js.precision(17);
        js << "{\"mesh\":\"" << c.meshFileName << "\",\"dim\":" << cs.mesh.dim << ",\"order\":" << cs.mesh.order
           << ",\"elements\":" << cs.mesh.numElements() << ",\"nodes_per_element\":" << cs.mesh.nodesPerElement
           << ",\"time\":[" << c.timeStart << "," << c.timeEnd << "," << c.timeStep << "," << c.timeRate << "]"
           << ",\"integrator\":\"" << c.timeIntMethod << "\",\"num_threads\":" << c.numThreads << ",\"mean_flow\":["
           << c.v0[0] << "," << c.v0[1] << "," << c.v0[2] << "," << c.rho0 << "," << c.c0 << "]"
           << ",\"bc_counts\":{\"absorbing\":" << nAbs << ",\"reflecting\":" << nRef << "},\"physical_bcs\":[";
        bool first = true;
        for (const auto& kv : c.physBCs) {
            js << (first ? "" : ",") << "[" << kv.first << ",\"" << kv.second.first << "\"]";
            first = false;
        }
        js << "],\"sources\":[";
        for (size_t i = 0; i < c.sources.size(); ++i) {
            const dg::SourceDef& s = c.sources[i];
            js << (i ? "," : "") << "{\"kind\":" << static_cast<int>(s.kind) << ",\"pole\":" << s.pole << ",\"x\":" << s.x
               << ",\"y\":" << s.y << ",\"z\":" << s.z << ",\"size\":" << s.size << ",\"amp\":" << s.amp
               << ",\"freq\":" << s.freq << ",\"phase\":" << s.phase << ",\"duration\":" << s.duration
               << ",\"formula\":\"" << s.formula << "\",\"rows\":" << s.table.size() << "}";
        }
        js << "],\"observers\":[";
        for (size_t i = 0; i < c.observers.size(); ++i)
            js << (i ? "," : "") << "[" << c.observers[i][0] << "," << c.observers[i][1] << "," << c.observers[i][2] << ","
               << c.observers[i][3] << "]";
        js << "],\"boundary_faces\":[";
        // per boundary face (element in file order, Gmsh local face, type), small meshes only
        int64_t nB = 0;
        for (size_t i = 0; i < isB.size(); ++i) nB += isB[i] ? 1 : 0;
        if (nB <= kSummaryMaxBoundaryFaces) {
            bool firstFace = true;
            for (size_t i = 0; i < isB.size(); ++i) {
                if (!isB[i]) continue;
                js << (firstFace ? "" : ",") << "[" << i / (cs.mesh.dim + 1) << "," << i % (cs.mesh.dim + 1) << ","
                   << static_cast<int>(bc[i]) << "]";
                firstFace = false;
            }
        }
        js << "],\"initial_conditions\":[";
        for (size_t i = 0; i < c.initConditions.size(); ++i) {
            js << (i ? "," : "") << "[";
            for (int k = 0; k < 6; ++k) js << (k ? "," : "") << c.initConditions[i][k];
            js << "]";
        }
        js << "]}";
    } catch (const std::exception& ex) {
        js.str("");
        js << ex.what();
        st = DG_ERR_INVALID;
    }
    const std::string s = js.str();
    const size_t n = std::min<size_t>(s.size(), static_cast<size_t>(capacity - 1));
    std::memcpy(out, s.data(), n);
    out[n] = 0;
    return st;
}

}  // extern "C"
