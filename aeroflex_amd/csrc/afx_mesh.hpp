// Host-side unstructured mesh in the reference's ("rans::mesh") order.
//
// The arrays and their order are the drop-in contract with the reference
// solver (src/rans/include/rans/mesh.h:198-273): faces are created while
// scanning cells in file order (owner = creating cell, mesh.h:317-341), the
// neighbour is filled by convert_node_face_info (mesh.h:344-375), metrics
// follow mesh.h:378-453 and ghost cells are appended in boundary-segment order
// (mesh.h:744-787).  Everything the device solver does is reported back in this
// order, so a caller that holds reference-ordered vectors never sees the GPU
// numbering.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace afx {

struct Mesh {
    // physical names: key = tag + 1000*dim (mesh.h:481-487)
    std::vector<uint32_t> name_keys;
    std::vector<std::string> names;

    std::vector<double> nodesX, nodesY;

    // faces
    std::vector<uint32_t> edgesNodes;   // 2 per face
    std::vector<uint32_t> edgesCells;   // 2 per face (owner, neighbour/ghost)
    std::vector<double> edgesLengths, edgesNormalsX, edgesNormalsY, edgesCentersX, edgesCentersY;

    // boundary
    std::vector<uint32_t> boundaryEdges;          // face index per boundary segment
    std::vector<uint32_t> boundaryEdgesIntTag;    // physical key per boundary segment
    std::vector<int32_t> boundaryNameIdx;         // index into names per boundary segment

    // cells (real cells first, then ghosts)
    std::vector<uint32_t> cellsNodes;   // 4 per cell
    std::vector<uint32_t> cellsEdges;   // 4 per real cell (UINT32_MAX pad)
    std::vector<int32_t> cellsIsTriangle;
    std::vector<double> cellsAreas, cellsCentersX, cellsCentersY;

    uint32_t nRealCells = 0;

    uint32_t n_nodes() const { return (uint32_t)nodesX.size(); }
    uint32_t n_faces() const { return (uint32_t)edgesLengths.size(); }
    uint32_t n_cells() const { return (uint32_t)cellsAreas.size(); }
    uint32_t n_bfaces() const { return (uint32_t)boundaryEdges.size(); }
    int name_index(uint32_t key) const;
};

static const uint32_t kNoEdge = 0xFFFFFFFFu;

// Raw element description (what a mesh file or a generator provides).
struct Elements {
    std::vector<double> x, y;
    std::vector<uint32_t> cell_nodes;   // 4 per cell
    std::vector<uint8_t> cell_nv;       // 3 or 4
    std::vector<uint32_t> bseg_nodes;   // 2 per boundary segment
    std::vector<uint32_t> bseg_key;     // physical key per boundary segment
    std::vector<uint32_t> name_keys;
    std::vector<std::string> names;
};

// Parse gmsh MSH 4.1 ASCII into Elements (throws std::runtime_error).
Elements read_msh41(const std::string& path);
// Write Elements as gmsh MSH 4.1 ASCII (one entity per physical name).
void write_msh41(const Elements& el, const std::string& path);
// Faces, metrics and ghosts with the reference's construction order.
Mesh build_mesh(const Elements& el);

// Analytic NACA 4-digit section (closed trailing edge), O-grid of ni points
// around and nj points from the wall to a circle of radius r_far chords;
// each quad split into two triangles along the (i,j)-(i+1,j+1) diagonal.
Elements gen_naca_ogrid(int ni, int nj, double r_far, double thickness, double first_h);

// Mixed triangle/quad mesh around a NACA 4-digit section: an O-grid of
// jittered points (Gaussian, sigma = jitter x local spacing, seeded
// splitmix64 + Box-Muller), each quad kept with probability quad_frac or split
// along the diagonal that satisfies the Delaunay empty-circle test.
Elements gen_jittered_airfoil(int ni, int nj, double r_far, double thickness, double first_h,
                              double jitter, double quad_frac, uint64_t seed);

// Reverse Cuthill-McKee order of the real-cell face-adjacency graph:
// returns perm with perm[new] = old.
std::vector<uint32_t> rcm_order(const Mesh& m);

// splitmix64 / Box-Muller (portable, identical on every platform)
struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
    double normal();
};

}  // namespace afx
