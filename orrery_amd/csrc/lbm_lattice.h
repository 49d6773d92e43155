// Velocity sets and the boundary-rule resolution shared by the host setup and the HIP kernels.
//
// D3Q19 order (src/aero/lbm3d/include/orrery/aero/lbm3d/d3q19.h:6-34): 0 rest; 1..6 = +x -x +y -y
// +z -z; 7..18 the face diagonals in opposite pairs. D2Q9 order (src/aero/lbm/include/orrery/
// aero/lbm/d2q9.h:9-12): 0 rest, 1 E, 2 N, 3 W, 4 S, 5 NE, 6 NW, 7 SW, 8 SE. The tables are the
// lattice's definition and must be identical to the reference's; everything else here is this
// library's own code.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#define LBM_HD __host__ __device__ __forceinline__

namespace lbm {

struct D3Q19 {
    static constexpr int Q = 19;
    static LBM_HD constexpr int cx(int i) {
        constexpr int t[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
        return t[i];
    }
    static LBM_HD constexpr int cy(int i) {
        constexpr int t[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
        return t[i];
    }
    static LBM_HD constexpr int cz(int i) {
        constexpr int t[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
        return t[i];
    }
    static LBM_HD constexpr double w(int i) {
        return i == 0 ? 1.0 / 3.0 : (i < 7 ? 1.0 / 18.0 : 1.0 / 36.0);
    }
    static LBM_HD constexpr int opp(int i) {
        constexpr int t[19] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
        return t[i];
    }
    static LBM_HD constexpr int mirror_y(int i) {
        constexpr int t[19] = {0, 1, 2, 4, 3, 5, 6, 9, 10, 7, 8, 11, 12, 13, 14, 18, 17, 16, 15};
        return t[i];
    }
    static LBM_HD constexpr int mirror_z(int i) {
        constexpr int t[19] = {0, 1, 2, 3, 4, 6, 5, 7, 8, 9, 10, 13, 14, 11, 12, 17, 18, 15, 16};
        return t[i];
    }
    // Slot of a population entering through the outlet (c_x = -1) in the outlet history:
    // 2, 8, 10, 12, 14 -> 0..4 (src/aero/lbm3d/lbm3d_internal.h:32-46).
    static LBM_HD constexpr int outlet_slot(int i) {
        return i == 2 ? 0 : (i == 8 ? 1 : (i == 10 ? 2 : (i == 12 ? 3 : 4)));
    }
};

struct D2Q9 {
    static constexpr int Q = 9;
    static LBM_HD constexpr int cx(int i) {
        constexpr int t[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
        return t[i];
    }
    static LBM_HD constexpr int cy(int i) {
        constexpr int t[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
        return t[i];
    }
    static LBM_HD constexpr double w(int i) {
        return i == 0 ? 4.0 / 9.0 : (i < 5 ? 1.0 / 9.0 : 1.0 / 36.0);
    }
    static LBM_HD constexpr int opp(int i) {
        constexpr int t[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
        return t[i];
    }
    static LBM_HD constexpr int mirror_y(int i) {
        constexpr int t[9] = {0, 1, 4, 3, 2, 8, 7, 6, 5};
        return t[i];
    }
    // 3, 6, 7 -> 0, 1, 2 (src/aero/lbm/lbm2d_internal.h:25-27).
    static LBM_HD constexpr int outlet_slot(int i) { return i == 3 ? 0 : (i == 6 ? 1 : 2); }
};

// Kinds of population source in the pull formulation of one step (the reference's SourceKind,
// src/aero/lbm3d/lbm3d_internal.h:20-28).
enum SourceKind : int { kNode = 0, kBounce = 1, kWall = 2, kInlet = 3, kOutlet = 4 };
// Device code only (D3Q19): a natural pull whose source is itself a boundary node, read from
// that node's stash slot instead of the lattice.
constexpr int kStash = 5;

enum WallCode : int { kPeriodic = 0, kSlip = 1, kNoSlip = 2 };

struct Resolved {
    int kind;
    int direction; // population index at the source
    int sx, sy, sz; // source node (kNode, kBounce: the solid node; kOutlet: outlet face node)
    int wall_side; // kWall: 0 y low, 1 y high, 2 z low, 3 z high
    int mirror;    // kNode/kOutlet through slip walls: bit 0 mirror_y applied, bit 1 mirror_z
};

// Where population i of node (x, y, z) comes from in the pull step, before the solid test:
// the y sides first, then the z sides, then x (the order of resolve_source,
// src/aero/lbm3d/lbm3d_internal.h:87-130). `solid_at(sx, sy, sz)` decides bounce-back.
template <typename SolidAt>
LBM_HD Resolved resolve_d3q19(int x, int y, int z, int i, int nx, int ny, int nz, int x_mode,
                              const int* side_kind /* y lo, y hi, z lo, z hi */,
                              const SolidAt& solid_at) {
    Resolved r{kNode, i, x - D3Q19::cx(i), y - D3Q19::cy(i), z - D3Q19::cz(i), -1, 0};
    if (r.sy < 0 || r.sy >= ny) {
        const int side = r.sy < 0 ? 0 : 1;
        const int kind = side_kind[side];
        if (kind == kPeriodic) {
            r.sy = (r.sy + ny) % ny;
        } else if (kind == kSlip) {
            r.sy = y;
            r.direction = D3Q19::mirror_y(r.direction);
            r.mirror |= 1;
        } else {
            r.kind = kWall;
            r.direction = D3Q19::opp(i);
            r.wall_side = side;
            return r;
        }
    }
    if (r.sz < 0 || r.sz >= nz) {
        const int side = r.sz < 0 ? 2 : 3;
        const int kind = side_kind[side];
        if (kind == kPeriodic) {
            r.sz = (r.sz + nz) % nz;
        } else if (kind == kSlip) {
            r.sz = z;
            r.direction = D3Q19::mirror_z(r.direction);
            r.mirror |= 2;
        } else {
            r.kind = kWall;
            r.direction = D3Q19::opp(i);
            r.wall_side = side;
            return r;
        }
    }
    if (r.sx < 0 || r.sx >= nx) {
        if (x_mode == 0) {
            r.sx = (r.sx + nx) % nx;
        } else if (r.sx < 0) {
            r.kind = kInlet;
            r.direction = i;
            return r;
        } else {
            r.kind = kOutlet;
            r.sx = nx - 1;
            return r;
        }
    }
    if (solid_at(r.sx, r.sy, r.sz)) {
        r.kind = kBounce;
        r.direction = D3Q19::opp(i);
    }
    return r;
}

// D2Q9 counterpart (src/aero/lbm/lbm2d_internal.h:37-75): y sides, then x.
template <typename SolidAt>
LBM_HD Resolved resolve_d2q9(int x, int y, int i, int nx, int ny, int x_mode,
                             const int* side_kind /* bottom, top */, const SolidAt& solid_at) {
    Resolved r{kNode, i, x - D2Q9::cx(i), y - D2Q9::cy(i), 0, -1, 0};
    if (r.sy < 0 || r.sy >= ny) {
        const int side = r.sy < 0 ? 0 : 1;
        const int kind = side_kind[side];
        if (kind == kPeriodic) {
            r.sy = (r.sy + ny) % ny;
        } else if (kind == kSlip) {
            r.sy = y;
            r.direction = D2Q9::mirror_y(i);
            r.mirror |= 1;
        } else {
            r.kind = kWall;
            r.direction = D2Q9::opp(i);
            r.wall_side = side;
            return r;
        }
    }
    if (r.sx < 0 || r.sx >= nx) {
        if (x_mode == 0) {
            r.sx = (r.sx + nx) % nx;
        } else if (r.sx < 0) {
            r.kind = kInlet;
            r.direction = i;
            return r;
        } else {
            r.kind = kOutlet;
            r.sx = nx - 1;
            return r;
        }
    }
    if (solid_at(r.sx, r.sy, 0)) {
        r.kind = kBounce;
        r.direction = D2Q9::opp(i);
    }
    return r;
}

// Quotient by a fixed divisor d >= 2 with one 64-bit high multiply: m = ceil(2^64 / d) makes
// floor(n m / 2^64) exact for every 32-bit n.
struct FastDiv {
    std::uint64_t m = 0;
    std::uint32_t d = 1;
    FastDiv() = default;
    explicit FastDiv(std::uint32_t divisor) : m(~std::uint64_t{0} / divisor + 1), d(divisor) {}
    __device__ __forceinline__ std::uint32_t div(std::uint32_t n) const {
        return static_cast<std::uint32_t>(__umul64hi(static_cast<std::uint64_t>(n), m));
    }
};

} // namespace lbm
