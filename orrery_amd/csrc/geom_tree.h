// The winding-number tree of a mesh (the host tree of geom.cpp) flattened for the device
// voxeliser (geom_device.hip), and the grid checks the two share.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

struct lbm_mesh;

namespace lbm::geom {

// One tree node: its dipole (area-weighted centre, area normal in units of the tree scale,
// radius), its bounding box, and its children or its run of triangles in tree order.
struct FlatNode {
    double centre[3];
    double normal[3];
    double radius;
    double lo[3];
    double hi[3];
    std::uint32_t first;
    std::uint32_t count;
    std::uint32_t left;
    std::uint32_t right;
    std::uint32_t leaf;
    std::uint32_t pad;
};

struct FlatTree {
    std::vector<FlatNode> nodes;  // node 0 is the root
    std::vector<double> tri;      // 9 doubles (three vertices) per triangle, in tree order
    double scale = 1.0;           // power of two of the mesh extent (dipole units)
    double beta = 2.0;            // far-field opening criterion
};

// Builds the tree geom.cpp evaluates on the host and flattens it.
void flatten_tree(const lbm_mesh* m, FlatTree& out);

// The grid checks of lbm_voxelize (false with the error recorded for lbm_geom_last_error).
bool check_grid(const double* origin, double spacing, std::uint64_t nx, std::uint64_t ny, std::uint64_t nz);

// Records a geometry error; returns LBM_EINVAL.
int fail(const std::string& message);

} // namespace lbm::geom
