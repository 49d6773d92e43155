// Internals of the D3Q19 device handle shared by the single-lattice ABI (lbm3d_abi.hip) and
// the batch engine (lbm_batch.hip): the handle itself, its device-memory accounting and the
// kernel descriptor built from it.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/orrery_lbm.h"
#include "lbm_device.h"
#include "lbm_graph.h"
#include "lbm_host.h"
#include "lbm_kernels.h"

namespace lbm::internal {
// The blocked population layout of a D3Q19 lattice (population_layout below).
struct PopulationLayout {
    std::uint32_t shift = 0;   // 2^shift cells per block
    std::uint32_t guard = 0;   // guard cells before the first cell
    std::size_t elements = 0;  // populations in the buffer, guards included
};
} // namespace lbm::internal

struct lbm3d {
    lbm3d_config cfg{};
    int storage = LBM_STORE_F32;
    std::size_t cells = 0;
    lbm::internal::PopulationLayout layout;
    lbm::host::Tables tables;
    std::vector<std::uint8_t> solid;
    hipStream_t stream = nullptr;
    // How a step is launched (LBM_SCHEDULE_*). With the two-stream schedule the boundary
    // kernel runs on `side`, concurrently with the bulk kernel on `stream`: within one step
    // every population address is read and written by one node only, and the stash is
    // double-buffered, so the two kernels are independent. `fork` and `join` order each step
    // after the previous one on both streams.
    int schedule = LBM_SCHEDULE_SERIAL;
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr;
    hipEvent_t join = nullptr;
    int device = 0;
    std::uint64_t steps = 0;
    double inlet[3] = {0, 0, 0};
    int block = 256;
    int cpt = 1;    // cells per thread of the bulk kernel, A step
    int cpt_b = 1;  // and B step (its x-shifted planes may prefer another width)
    // Cap on the bulk kernel's workgroups per compute unit, A step and B step (0: none), and
    // the device's LDS per compute unit and per workgroup that the cap is reserved from.
    int groups[2] = {0, 0};
    std::uint32_t lds_per_cu = 0;
    std::uint32_t lds_per_block = 0;
    // device memory
    void* f = nullptr;
    std::uint8_t* cls = nullptr;
    std::uint32_t* bcell = nullptr;
    std::uint32_t* bcode = nullptr;
    std::uint32_t* bwrite = nullptr;
    void* bq = nullptr;
    void* stash = nullptr;
    void* hist = nullptr;
    std::uint32_t* lcell = nullptr;
    std::uint8_t* ldir = nullptr;
    double* lq = nullptr;
    std::uint32_t* lslot = nullptr;
    std::uint32_t* tile_slot = nullptr;
    std::uint32_t* scell = nullptr;
    void* offsets = nullptr;  // storage offsets, [38] arithmetic-type values (Lattice3::shift)
    double* partial = nullptr;
    double* result = nullptr;
    // Steps since initialisation, on the device (after the reduction results): the 16-bit
    // formats draw their rounding dither from it; a one-thread kernel advances it each step.
    std::uint32_t* clock = nullptr;
    std::size_t device_bytes = 0;
    std::vector<void*> owned;
    // Replayed runs of steps of lbm3d_step (serial and fused schedules).
    lbm::StepGraph graph;
    // Held by a batch (lbm_batch_add): the batch steps the lattice on its own stream and keeps
    // its A-A parity in device memory, so the calls that step, re-initialise, retarget or free
    // the lattice refuse it; the queries and readbacks stay open.
    bool in_batch = false;
};

namespace lbm::internal {

inline std::size_t store_bytes(int storage) {
    switch (storage) {
        case LBM_STORE_F16S:
        case LBM_STORE_BF16S:
            return 2;
        case LBM_STORE_F64:
            return 8;
        default:
            return 4;
    }
}

inline std::size_t real_bytes(int storage) { return storage == LBM_STORE_F64 ? 8 : 4; }

// Default cells per thread of the bulk kernel: runs of two 16-bit populations move as one
// 32-bit access; binary32 and binary64 runs of one cell keep the most waves resident.
inline int default_cells_per_thread(int storage) {
    return storage == LBM_STORE_F16S || storage == LBM_STORE_BF16S ? 2 : 1;
}

// The blocked population layout of a D3Q19 lattice (Lattice3::f): 2^shift cells per block, at
// least 2^8 (a wave's cells, up to 64 lanes of 4, stay in one block) and at most 32 KiB per
// direction (2^13 cells of 4 or 8 bytes, 2^14 of 2) or the first power of two that holds a plane;
// `guard` cells, whole blocks, before the first cell and at least as many after the last,
// covering the farthest neighbour (nx ny + nx + 1 cells) and a run's further cells. On
// 512x256x256 the in-place pattern moves its 19 directions 9% faster in blocks of 2^13 cells
// than in one plane per direction (its A step 14%, its B step 3%); the bulk kernel's f32 step is
// fastest at 2^13 (2^12, 2^14, 2^15: 1.5% slower), its f16s step at 2^14-2^15 (1.3% faster than at
// 2^13). ORRERY_LBM_BLOCK_SHIFT (8 to 16), read from the environment at each lattice's
// creation, forces the block size for layout experiments and tests (results do not depend on it).
inline PopulationLayout population_layout(std::size_t nx, std::size_t ny, std::size_t nz, int storage) {
    const char* env = std::getenv("ORRERY_LBM_BLOCK_SHIFT");
    const long v = env != nullptr ? std::strtol(env, nullptr, 10) : 0;
    const std::uint32_t forced = v >= 8 && v <= 16 ? static_cast<std::uint32_t>(v) : 0U;
    PopulationLayout l;
    const std::size_t plane = nx * ny;
    const std::uint32_t most = store_bytes(storage) == 2 ? 14U : 13U;
    l.shift = 8;
    while (l.shift < most && (std::size_t{1} << l.shift) < plane) {
        ++l.shift;
    }
    if (forced != 0U) {
        l.shift = forced;
    }
    const std::size_t block = std::size_t{1} << l.shift;
    l.guard = static_cast<std::uint32_t>(lbm::host::padded(plane + nx + 1 + 4, block));
    const std::size_t blocks = (nx * ny * nz + 2U * l.guard + block - 1) / block;
    l.elements = 19U * block * blocks;
    return l;
}

// Length of the node class array: the cells rounded up to 128, padded with solid entries, so
// that a run of cells reads its classes at once.
inline std::size_t class_length(std::size_t cells) { return lbm::host::padded(cells, 128); }

// Device bytes of a lattice with these tables (the admission rule of the batch engine and of
// the memory check; the counterpart of gpu_memory_bytes, lbm3d_common.cpp:276-300).
inline std::size_t lattice_bytes(std::size_t cells, const PopulationLayout& layout, const lbm::host::Tables& t,
                                 int storage, bool interpolated) {
    const std::size_t sb = store_bytes(storage);
    const std::size_t rb = real_bytes(storage);
    std::size_t total = layout.elements * sb;  // populations (in place)
    total += class_length(cells);              // node classes (u8, padded)
    total += t.bcell.size() * 4U;                                       // boundary list
    total += t.tile_slot.size() * 4U;                                   // boundary slot per 64 cells
    total += t.scell.size() * 4U;                                       // side list
    total += 38U * rb;                                                  // storage offsets
    total += 19U * static_cast<std::size_t>(t.bstride) * 4U;            // source codes
    total += static_cast<std::size_t>(t.bstride) * 4U;                  // lattice write masks
    if (interpolated) {
        total += 19U * static_cast<std::size_t>(t.bstride) * rb; // wall fractions
    }
    total += 2U * 19U * static_cast<std::size_t>(t.bstride) * rb; // stash
    total += 5U * static_cast<std::size_t>(t.bstride) * rb;       // outlet history
    total += t.links.size() * (4U + 1U + 8U + 4U);                 // links
    const std::size_t red = std::max<std::size_t>(lbm::reduction_blocks(cells) + lbm::reduction_blocks(t.bcell.size()),
                                                  6U * lbm::reduction_blocks(t.links.size()));
    total += (red + 8U) * 8U;
    return total;
}

// Bytes lbm3d_create would allocate for this configuration and mask.
inline std::size_t required_bytes3(const lbm3d_config& c, const std::uint8_t* solid, const double* distance,
                                   int storage) {
    const std::size_t cells = c.nx * c.ny * c.nz;
    const lbm::host::Tables t = lbm::host::tables3(c, solid, distance);
    return lattice_bytes(cells, population_layout(c.nx, c.ny, c.nz, storage), t, storage, c.solid_wall == LBM_SOLID_INTERPOLATED);
}

// The storage offsets of a format with `offset` (Lattice3::shift): f_eq_i(1, u_ref) - w_i with
// u_ref the configured inlet velocity (the stream the lattice is started in and relaxes to; zero
// on a periodic lattice), then the same in the stored units. Zero for the other formats.
template <class T> std::vector<typename T::R> storage_offsets(const lbm3d_config& c) {
    using R = typename T::R;
    std::vector<R> t(38, R(0));
    if constexpr (T::offset) {
        const bool inflow = c.x_boundary == LBM_X_INFLOW_OUTFLOW;
        const double u[3] = {inflow ? c.inlet_ux : 0.0, inflow ? c.inlet_uy : 0.0, inflow ? c.inlet_uz : 0.0};
        const double uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
        for (int i = 0; i < 19; ++i) {
            const double cu = lbm::D3Q19::cx(i) * u[0] + lbm::D3Q19::cy(i) * u[1] + lbm::D3Q19::cz(i) * u[2];
            t[i] = static_cast<R>(lbm::D3Q19::w(i) * (3.0 * cu + 4.5 * cu * cu - 1.5 * uu));
            t[19 + i] = t[i] * T::kOffsetScale;  // exact: a power of two
        }
    }
    return t;
}

template <class T> lbm::Lattice3<typename T::R> descriptor3(const lbm3d& h) {
    using R = typename T::R;
    lbm::Lattice3<R> p;
    std::memset(static_cast<void*>(&p), 0, sizeof(p));  // padding too: the bytes key step graphs
    const lbm3d_config& c = h.cfg;
    p.f = h.f;
    p.cls = h.cls;
    p.bcell = h.bcell;
    p.bcode = h.bcode;
    p.bwrite = h.bwrite;
    p.bq = static_cast<const R*>(h.bq);
    p.stash = static_cast<R*>(h.stash);
    p.hist = static_cast<R*>(h.hist);
    p.tile_slot = h.tile_slot;
    p.scell = h.scell;
    p.clock = h.clock;
    p.nx = static_cast<std::uint32_t>(c.nx);
    p.ny = static_cast<std::uint32_t>(c.ny);
    p.nz = static_cast<std::uint32_t>(c.nz);
    p.cells = static_cast<std::uint32_t>(h.cells);
    p.block_shift = h.layout.shift;
    p.guard = h.layout.guard;
    p.nb = static_cast<std::uint32_t>(h.tables.bcell.size());
    p.ns = static_cast<std::uint32_t>(h.tables.scell.size());
    p.bstride = h.tables.bstride;
    p.div_nx = lbm::FastDiv(p.nx);
    p.div_plane = lbm::FastDiv(p.nx * p.ny);
    p.x_mode = c.x_boundary;
    p.convective = c.outlet == LBM_OUTLET_CONVECTIVE ? 1 : 0;
    p.side[0] = c.y_low.kind;
    p.side[1] = c.y_high.kind;
    p.side[2] = c.z_low.kind;
    p.side[3] = c.z_high.kind;
    p.wall_u[0] = static_cast<R>(c.y_low.velocity_x);
    p.wall_u[1] = static_cast<R>(c.y_high.velocity_x);
    p.wall_u[2] = static_cast<R>(c.z_low.velocity_x);
    p.wall_u[3] = static_cast<R>(c.z_high.velocity_x);
    for (int a = 0; a < 3; ++a) {
        p.inlet[a] = static_cast<R>(h.inlet[a]);
    }
    p.tau = static_cast<R>(c.tau);
    p.omega = static_cast<R>(1) / static_cast<R>(c.tau);
    const double k = c.smagorinsky_enabled ? c.smagorinsky_constant : 0.0;
    p.smag_c = static_cast<R>(k);
    p.smag_c2 = static_cast<R>(k * k);
    // The binary32 kernels switch the closure off when C^2 is 0, as the GLSL does
    // (lbm_d3q19_common.glsl:287); the binary64 one follows the CPU's flag.
    p.les = T::shifted ? (p.smag_c2 > R(0) ? 1 : 0) : (c.smagorinsky_enabled ? 1 : 0);
    p.force[0] = static_cast<R>(c.body_force_x);
    p.force[1] = static_cast<R>(c.body_force_y);
    p.force[2] = static_cast<R>(c.body_force_z);
    p.forced = (c.body_force_x != 0.0 || c.body_force_y != 0.0 || c.body_force_z != 0.0) ? 1 : 0;
    p.outlet_speed = static_cast<R>(c.inlet_ux);
    const bool sponge = c.sponge_strength > 0.0 && c.x_boundary == LBM_X_INFLOW_OUTFLOW;
    p.sponge_strength = sponge ? static_cast<R>(c.sponge_strength) : R(0);
    p.sponge_in = static_cast<std::uint32_t>(c.sponge_inlet);
    p.sponge_out = static_cast<std::uint32_t>(c.sponge_outlet);
    p.sponge_sides = static_cast<std::uint32_t>(c.sponge_sides);
    p.sponge_y = c.y_low.kind != LBM_WALL_PERIODIC ? 1 : 0;
    p.sponge_z = c.z_low.kind != LBM_WALL_PERIODIC ? 1 : 0;
    p.shift = static_cast<const R*>(h.offsets);
    return p;
}

struct LinkView {
    const std::uint32_t* cell;
    const std::uint8_t* dir;
    const double* q;
    const std::uint32_t* slot;
    std::uint32_t count;
};

inline LinkView link_view(const lbm3d& h) {
    return LinkView{h.lcell, h.ldir, h.lq, h.lslot, static_cast<std::uint32_t>(h.tables.links.size())};
}

inline int storage_of(const lbm3d& h) { return h.storage; }

// The batch engine advanced the lattice on its own stream.
inline void advance_steps(lbm3d& h, std::uint64_t steps) { h.steps += steps; }

} // namespace lbm::internal
