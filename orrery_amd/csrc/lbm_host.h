// Host-side setup of the device lattices: validation with the reference's reasons, node
// classes, boundary links, boundary-node tables, device-memory accounting and the library's
// error state. Integer outputs (classes, links, codes) are exact: they follow the reference's
// loops (src/aero/lbm3d/lbm3d_common.cpp:74-198, src/aero/lbm3d/lbm3d_internal.h:139-248,
// src/aero/lbm/lbm2d_common.cpp:31-97) in the same order.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/orrery_lbm.h"

namespace lbm::host {

// Status with message for the C ABI (stored thread-locally, read by lbm_last_error()).
int fail(int code, const std::string& message);
int ok();

constexpr std::size_t kMaxGridCells = std::size_t{1} << 27U; // geom::max_grid_cells (voxel_grid.h:29)

std::string problem3(const lbm3d_config& c, const std::uint8_t* solid, const double* distance);
std::string problem2(const lbm2d_config& c, const std::uint8_t* solid);
// Reasons the binary32 kernels refuse (lbm3d_gpu.cpp:94-127, lbm2d_gpu.cpp equivalent).
std::string binary32_problem3(const lbm3d_config& c);
std::string binary32_problem2(const lbm2d_config& c);
// check_inlet_velocity (lbm3d_common.cpp:163-174).
std::string inlet_problem3(const lbm3d_config& c, const double v[3]);

struct Link {
    std::uint32_t cell;
    std::uint32_t direction;
    double q;
};

// Everything the device lattice needs from the mask.
struct Tables {
    // 0 solid, 1 bulk, 2 boundary (listed), 3 side, 4 inlet (D3Q19: streamed in place by the
    // bulk kernel; the reference's classification counts them as boundary nodes)
    std::vector<std::uint8_t> cls;
    std::vector<std::uint32_t> bcell;    // boundary node -> cell (cell order)
    std::vector<std::uint32_t> scell;    // D3Q19: side node -> cell (cell order)
    std::vector<std::uint32_t> bcode;    // [Q][bstride] source codes
    // D3Q19: per boundary node, bit i set when its output i must also be stored in the lattice
    // (the node it streams to reads the lattice there: a bulk or side node, or the far read of a
    // short bounce-back link); boundary readers of boundary sources read the stash instead.
    std::vector<std::uint32_t> bwrite;
    std::vector<double> bq;              // [Q][bstride] wall fractions (interpolated wall only)
    std::vector<Link> links;             // boundary_links order
    std::vector<std::uint32_t> lslot;    // boundary slot of each link's node
    // D3Q19: boundary nodes before each run of 64 cells ([cells / 64 + 1]); the fused step
    // kernel finds a boundary node's slot as tile_slot[cell / 64] + its rank in the wave.
    std::vector<std::uint32_t> tile_slot;
    std::uint32_t bstride = 0;
};

std::uint32_t padded(std::size_t n, std::uint32_t to);
Tables tables3(const lbm3d_config& c, const std::uint8_t* solid, const double* distance);
Tables tables2(const lbm2d_config& c, const std::uint8_t* solid);

// Fake device-memory limit (0: none) and the device query that honours it.
void set_memory_limit(std::size_t bytes);
int device_memory(std::size_t* available, std::size_t* max_alloc);
std::string megabytes(std::size_t bytes);

} // namespace lbm::host
