// Host setup of the device lattices (see lbm_host.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <iomanip>
#include <limits>
#include <sstream>

#include "lbm_device.h"
#include "lbm_host.h"
#include "lbm_lattice.h"

namespace lbm::host {

namespace {

thread_local std::string g_error;
std::size_t g_memory_limit = 0;

constexpr double kMaxLatticeMach = 0.1; // src/aero/lbm/include/orrery/aero/lbm/lattice_units.h:475

double lattice_mach(double u) { return std::abs(u) / std::sqrt(1.0 / 3.0); } // lattice_units.cpp:18-20

// check_mach (lbm3d_common.cpp:21-30, lbm2d_common.cpp:18-27)
std::string check_mach(const char* prefix, const char* what, double speed) {
    if (!(lattice_mach(speed) < kMaxLatticeMach)) {
        std::ostringstream message;
        message << prefix << what << " " << speed << " has lattice Mach number " << lattice_mach(speed)
                << ", not below " << kMaxLatticeMach << " (APPROXIMATION A3)";
        return message.str();
    }
    return {};
}

bool finite(double v) { return std::isfinite(v); }

} // namespace

int fail(int code, const std::string& message) {
    g_error = message;
    return code;
}

int ok() { return LBM_OK; }

const char* last_error() { return g_error.c_str(); }

std::uint32_t padded(std::size_t n, std::uint32_t to) {
    const std::size_t m = n == 0 ? to : (n + to - 1) / to * to;
    return static_cast<std::uint32_t>(m);
}

// lbm3d_common.cpp:74-155 (problem_of)
std::string problem3(const lbm3d_config& c, const std::uint8_t* solid, const double* distance) {
    if (c.nx < 3 || c.ny < 3 || c.nz < 3) {
        return "lbm3d: the lattice needs at least 3 nodes along each axis";
    }
    const std::size_t max = std::numeric_limits<std::size_t>::max();
    if (c.ny > max / c.nx || c.nx * c.ny > max / c.nz || c.nx * c.ny * c.nz > kMaxGridCells) {
        return "lbm3d: more than geom::max_grid_cells nodes";
    }
    if (solid == nullptr) {
        return "lbm3d: the solid mask does not match the lattice size";
    }
    if (!finite(c.tau) || !(c.tau > 0.5)) {
        return "lbm3d: tau must be finite and above 1/2 (positive viscosity)";
    }
    for (const double v : {c.inlet_ux, c.inlet_uy, c.inlet_uz, c.body_force_x, c.body_force_y, c.body_force_z,
                           c.y_low.velocity_x, c.y_high.velocity_x, c.z_low.velocity_x, c.z_high.velocity_x}) {
        if (!finite(v)) {
            return "lbm3d: velocities and body force must be finite";
        }
    }
    if (c.smagorinsky_enabled != 0 && (!finite(c.smagorinsky_constant) || c.smagorinsky_constant < 0.0)) {
        return "lbm3d: the Smagorinsky constant must be finite and non-negative";
    }
    for (const lbm_wall* w : {&c.y_low, &c.y_high, &c.z_low, &c.z_high}) {
        if (w->kind < LBM_WALL_PERIODIC || w->kind > LBM_WALL_NO_SLIP) {
            return "lbm3d: unknown wall kind";
        }
    }
    if ((c.y_low.kind == LBM_WALL_PERIODIC) != (c.y_high.kind == LBM_WALL_PERIODIC) ||
        (c.z_low.kind == LBM_WALL_PERIODIC) != (c.z_high.kind == LBM_WALL_PERIODIC)) {
        return "lbm3d: an axis is periodic on both sides or on neither";
    }
    const char* names[4] = {"y = 0 wall speed", "y = ny - 1 wall speed", "z = 0 wall speed", "z = nz - 1 wall speed"};
    const lbm_wall* walls[4] = {&c.y_low, &c.y_high, &c.z_low, &c.z_high};
    for (int k = 0; k < 4; ++k) {
        std::string bad = check_mach("lbm3d: ", names[k], walls[k]->velocity_x);
        if (!bad.empty()) {
            return bad;
        }
    }
    if (!finite(c.sponge_strength) || c.sponge_strength < 0.0 || c.sponge_strength > 1.0) {
        return "lbm3d: the absorbing-layer strength must be in [0, 1]";
    }
    if (c.sponge_strength > 0.0 &&
        (c.x_boundary != LBM_X_INFLOW_OUTFLOW || c.sponge_inlet > c.nx || c.sponge_outlet > c.nx ||
         c.sponge_sides > std::min(c.ny, c.nz))) {
        return "lbm3d: absorbing layers need the inflow/outflow x boundary and must fit in the lattice";
    }
    const std::size_t cells = c.nx * c.ny * c.nz;
    if (c.solid_wall == LBM_SOLID_INTERPOLATED) {
        if (distance == nullptr) {
            return "lbm3d: the interpolated wall needs the mask's signed distance";
        }
        for (std::size_t k = 0; k < cells; ++k) {
            if (std::isnan(distance[k])) {
                return "lbm3d: the mask's signed distance must not be NaN";
            }
        }
    }
    if (c.x_boundary == LBM_X_INFLOW_OUTFLOW) {
        const double speed = std::sqrt(c.inlet_ux * c.inlet_ux + c.inlet_uy * c.inlet_uy + c.inlet_uz * c.inlet_uz);
        std::string bad = check_mach("lbm3d: ", "inlet speed", speed);
        if (!bad.empty()) {
            return bad;
        }
        for (std::size_t z = 0; z < c.nz; ++z) {
            for (std::size_t y = 0; y < c.ny; ++y) {
                if (solid[c.nx * (y + c.ny * z)] != 0 || solid[c.nx - 1 + c.nx * (y + c.ny * z)] != 0) {
                    return "lbm3d: the inlet and outlet faces must be fluid";
                }
            }
        }
        if (c.outlet == LBM_OUTLET_CONVECTIVE && !(c.inlet_ux > 0.0)) {
            return "lbm3d: the convective outlet needs a positive inlet speed";
        }
    }
    return {};
}

// lbm2d_common.cpp:31-77 (validate)
std::string problem2(const lbm2d_config& c, const std::uint8_t* solid) {
    if (c.nx < 3 || c.ny < 3) {
        return "lbm2d: the lattice needs at least 3 x 3 nodes";
    }
    if (c.nx > kMaxGridCells / c.ny) {
        return "lbm2d: more than geom::max_grid_cells nodes";
    }
    if (solid == nullptr) {
        return "lbm2d: the solid mask does not match the lattice size";
    }
    if (!finite(c.tau) || !(c.tau > 0.5)) {
        return "lbm2d: tau must be finite and above 1/2 (positive viscosity)";
    }
    if (!finite(c.inlet_ux) || !finite(c.inlet_uy) || !finite(c.body_force_x) || !finite(c.body_force_y) ||
        !finite(c.bottom.velocity_x) || !finite(c.top.velocity_x)) {
        return "lbm2d: velocities and body force must be finite";
    }
    if (c.smagorinsky_enabled != 0 && (!finite(c.smagorinsky_constant) || c.smagorinsky_constant < 0.0)) {
        return "lbm2d: the Smagorinsky constant must be finite and non-negative";
    }
    for (const lbm_wall* w : {&c.bottom, &c.top}) {
        if (w->kind < LBM_WALL_PERIODIC || w->kind > LBM_WALL_NO_SLIP) {
            return "lbm2d: unknown wall kind";
        }
    }
    if ((c.bottom.kind == LBM_WALL_PERIODIC) != (c.top.kind == LBM_WALL_PERIODIC)) {
        return "lbm2d: y is periodic on both sides or on neither";
    }
    std::string bad = check_mach("lbm2d: ", "bottom wall speed", c.bottom.velocity_x);
    if (bad.empty()) {
        bad = check_mach("lbm2d: ", "top wall speed", c.top.velocity_x);
    }
    if (!bad.empty()) {
        return bad;
    }
    if (c.x_boundary == LBM_X_INFLOW_OUTFLOW) {
        bad = check_mach("lbm2d: ", "inlet speed", std::hypot(c.inlet_ux, c.inlet_uy));
        if (!bad.empty()) {
            return bad;
        }
        for (std::size_t y = 0; y < c.ny; ++y) {
            if (solid[c.nx * y] != 0 || solid[c.nx - 1 + c.nx * y] != 0) {
                return "lbm2d: the inlet and outlet columns must be fluid";
            }
        }
        if (c.outlet == LBM_OUTLET_CONVECTIVE && !(c.inlet_ux > 0.0)) {
            return "lbm2d: the convective outlet needs a positive inlet speed";
        }
    }
    return {};
}

namespace {
// round_to_float32 (src/aero/compute/float32_pack.cpp:17-54): the value a binary32 kernel
// sees, infinite when it overflows.
double as_float32(double v) { return static_cast<double>(static_cast<float>(v)); }

struct Parameter {
    const char* name;
    double value;
};

std::string binary32_check(const char* prefix, std::initializer_list<Parameter> params, double tau) {
    for (const Parameter& p : params) {
        if (!std::isfinite(as_float32(p.value))) {
            std::ostringstream message;
            message << prefix << p.name << " " << p.value << " is outside the binary32 range of the kernels";
            return message.str();
        }
    }
    if (!(as_float32(tau) > 0.5)) {
        std::ostringstream message;
        message << std::setprecision(17) << prefix << "tau " << tau
                << " rounds to 1/2 in binary32 (zero viscosity in the kernels); tau - 1/2 must exceed 2^-25 "
                   "(about 3e-8)";
        return message.str();
    }
    return {};
}
} // namespace

// check_binary32 (src/aero/lbm3d/lbm3d_gpu.cpp:94-127)
std::string binary32_problem3(const lbm3d_config& c) {
    const double k = c.smagorinsky_enabled ? c.smagorinsky_constant : 0.0;
    return binary32_check("lbm3d gpu: ",
                          {{"tau", c.tau},
                           {"y = 0 wall speed", c.y_low.velocity_x},
                           {"y = ny - 1 wall speed", c.y_high.velocity_x},
                           {"z = 0 wall speed", c.z_low.velocity_x},
                           {"z = nz - 1 wall speed", c.z_high.velocity_x},
                           {"inlet x velocity", c.inlet_ux},
                           {"inlet y velocity", c.inlet_uy},
                           {"inlet z velocity", c.inlet_uz},
                           {"squared Smagorinsky constant", k * k},
                           {"body force x", c.body_force_x},
                           {"body force y", c.body_force_y},
                           {"body force z", c.body_force_z}},
                          c.tau);
}

std::string binary32_problem2(const lbm2d_config& c) {
    const double k = c.smagorinsky_enabled ? c.smagorinsky_constant : 0.0;
    return binary32_check("lbm2d gpu: ",
                          {{"tau", c.tau},
                           {"bottom wall speed", c.bottom.velocity_x},
                           {"top wall speed", c.top.velocity_x},
                           {"inlet x velocity", c.inlet_ux},
                           {"inlet y velocity", c.inlet_uy},
                           {"squared Smagorinsky constant", k * k},
                           {"body force x", c.body_force_x},
                           {"body force y", c.body_force_y}},
                          c.tau);
}

// check_inlet_velocity (lbm3d_common.cpp:163-174)
std::string inlet_problem3(const lbm3d_config& c, const double v[3]) {
    if (c.x_boundary != LBM_X_INFLOW_OUTFLOW) {
        return "lbm3d: a periodic lattice has no inlet";
    }
    if (!finite(v[0]) || !finite(v[1]) || !finite(v[2])) {
        return "lbm3d: the inlet velocity must be finite";
    }
    return check_mach("lbm3d: ", "inlet speed", std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
}

namespace {

// link_fraction (src/aero/lbm3d/lbm3d_internal.h:139-179)
double link_fraction(const lbm3d_config& c, const std::uint8_t* solid, const double* distance, std::size_t x,
                     std::size_t y, std::size_t z, int i) {
    if (c.solid_wall != LBM_SOLID_INTERPOLATED) {
        return 0.5;
    }
    const auto inside = [&](std::ptrdiff_t a, std::ptrdiff_t b, std::ptrdiff_t d) {
        return a >= 0 && b >= 0 && d >= 0 && a < static_cast<std::ptrdiff_t>(c.nx) &&
               b < static_cast<std::ptrdiff_t>(c.ny) && d < static_cast<std::ptrdiff_t>(c.nz);
    };
    const auto at = [&](std::ptrdiff_t a, std::ptrdiff_t b, std::ptrdiff_t d) {
        return static_cast<std::size_t>(a) + c.nx * (static_cast<std::size_t>(b) + c.ny * static_cast<std::size_t>(d));
    };
    const std::ptrdiff_t sx = static_cast<std::ptrdiff_t>(x) - D3Q19::cx(i);
    const std::ptrdiff_t sy = static_cast<std::ptrdiff_t>(y) - D3Q19::cy(i);
    const std::ptrdiff_t sz = static_cast<std::ptrdiff_t>(z) - D3Q19::cz(i);
    if (!inside(sx, sy, sz)) {
        return 0.5;
    }
    const double fluid = distance[at(static_cast<std::ptrdiff_t>(x), static_cast<std::ptrdiff_t>(y),
                                     static_cast<std::ptrdiff_t>(z))];
    const double sol = distance[at(sx, sy, sz)];
    if (!std::isfinite(fluid) || !std::isfinite(sol) || fluid < 0.0 || sol > 0.0 || (fluid == 0.0 && sol == 0.0)) {
        return 0.5;
    }
    const double span = fluid - sol;
    const double ratio = std::isfinite(span) ? fluid / span : 0.5 * fluid / (0.5 * fluid - 0.5 * sol);
    const double q = std::min(ratio, 1.0);
    if (q < 0.5) {
        const std::ptrdiff_t fx = static_cast<std::ptrdiff_t>(x) + D3Q19::cx(i);
        const std::ptrdiff_t fy = static_cast<std::ptrdiff_t>(y) + D3Q19::cy(i);
        const std::ptrdiff_t fz = static_cast<std::ptrdiff_t>(z) + D3Q19::cz(i);
        if (!inside(fx, fy, fz) || solid[at(fx, fy, fz)] != 0) {
            return 0.5;
        }
    }
    return q;
}

} // namespace

Tables tables3(const lbm3d_config& c, const std::uint8_t* solid, const double* distance) {
    Tables t;
    const std::size_t nx = c.nx;
    const std::size_t ny = c.ny;
    const std::size_t nz = c.nz;
    const std::size_t plane = nx * ny;
    const std::size_t cells = plane * nz;
    // classify_nodes (lbm3d_internal.h:219-248): faces and neighbours of solids are boundary.
    t.cls.assign(cells, 1);
    std::ptrdiff_t offset[19];
    for (int i = 0; i < 19; ++i) {
        offset[i] = -static_cast<std::ptrdiff_t>(D3Q19::cx(i)) - static_cast<std::ptrdiff_t>(nx) * D3Q19::cy(i) -
                    static_cast<std::ptrdiff_t>(plane) * D3Q19::cz(i);
    }
    const int sides[4] = {c.y_low.kind, c.y_high.kind, c.z_low.kind, c.z_high.kind};
    const int inx = static_cast<int>(nx);
    const int iny = static_cast<int>(ny);
    const int inz = static_cast<int>(nz);
    const auto solid_at = [&](int a, int b, int d) {
        return solid[static_cast<std::size_t>(a) + nx * (static_cast<std::size_t>(b) + ny * static_cast<std::size_t>(d))] != 0;
    };
    const auto cell_of = [&](int a, int b, int d) {
        return static_cast<std::uint32_t>(static_cast<std::size_t>(a) +
                                          nx * (static_cast<std::size_t>(b) + ny * static_cast<std::size_t>(d)));
    };
    std::size_t nb = 0;
    for (std::size_t z = 0; z < nz; ++z) {
        for (std::size_t y = 0; y < ny; ++y) {
            const std::size_t row = nx * (y + ny * z);
            const bool face_row = y == 0 || z == 0 || y + 1 == ny || z + 1 == nz;
            for (std::size_t x = 0; x < nx; ++x) {
                const std::size_t here = row + x;
                if (solid[here] != 0) {
                    t.cls[here] = 0;
                    continue;
                }
                if (face_row || x == 0 || x + 1 == nx) {
                    t.cls[here] = 2;
                    // A side node: on a y or z face, not on an x face, and every source in the
                    // lattice after the sides' mirrors and wraps (no wall, solid, inlet or outlet
                    // source). An inlet node: on the velocity inlet face, off the y and z faces,
                    // every source in the lattice or beyond the inlet. Both stream in place
                    // through the bulk kernel.
                    const bool side = face_row && x > 0 && x + 1 < nx;
                    const bool inlet = !face_row && x == 0 && c.x_boundary == LBM_X_INFLOW_OUTFLOW;
                    bool in_place = side || inlet;
                    for (int i = 1; i < 19 && in_place; ++i) {
                        const Resolved r = resolve_d3q19(static_cast<int>(x), static_cast<int>(y), static_cast<int>(z), i,
                                                         inx, iny, inz, c.x_boundary, sides, solid_at);
                        in_place = r.kind == kNode || (inlet && r.kind == kInlet);
                    }
                    if (in_place) {
                        t.cls[here] = side ? 3 : 4;
                    } else {
                        ++nb;
                    }
                    continue;
                }
                for (int i = 1; i < 19; ++i) {
                    if (solid[static_cast<std::size_t>(static_cast<std::ptrdiff_t>(here) + offset[i])] != 0) {
                        t.cls[here] = 2;
                        ++nb;
                        break;
                    }
                }
            }
        }
    }
    t.bcell.reserve(nb);
    std::vector<std::uint32_t> slot_of;
    slot_of.assign(cells, 0xffffffffU);
    t.tile_slot.assign(cells / 64 + 1, 0U);
    for (std::size_t cell = 0; cell < cells; ++cell) {
        if ((cell & 63U) == 0) {
            t.tile_slot[cell >> 6] = static_cast<std::uint32_t>(t.bcell.size());
        }
        if (t.cls[cell] == 2) {
            slot_of[cell] = static_cast<std::uint32_t>(t.bcell.size());
            t.bcell.push_back(static_cast<std::uint32_t>(cell));
        } else if (t.cls[cell] == 3) {
            t.scell.push_back(static_cast<std::uint32_t>(cell));
        }
    }
    if ((cells & 63U) == 0) {
        t.tile_slot[cells >> 6] = static_cast<std::uint32_t>(t.bcell.size());
    }
    t.bstride = padded(nb, 64);
    t.bcode.assign(static_cast<std::size_t>(19) * t.bstride, 0U);
    if (c.solid_wall == LBM_SOLID_INTERPOLATED) {
        t.bq.assign(static_cast<std::size_t>(19) * t.bstride, 0.5);
    }
    for (std::size_t b = 0; b < nb; ++b) {
        const std::size_t cell = t.bcell[b];
        const int z = static_cast<int>(cell / plane);
        const int y = static_cast<int>((cell % plane) / nx);
        const int x = static_cast<int>(cell % nx);
        for (int i = 0; i < 19; ++i) {
            const Resolved r = resolve_d3q19(x, y, z, i, inx, iny, inz, c.x_boundary, sides, solid_at);
            std::uint32_t code = 0;
            const bool my = (r.mirror & 1) != 0;
            const bool mz = (r.mirror & 2) != 0;
            switch (r.kind) {
                case kNode: {
                    const std::uint32_t src = cell_of(r.sx, r.sy, r.sz);
                    if (t.cls[src] == 2) {
                        // The source's output is also in its stash slot, densely.
                        code = make_code(kStash, my, mz, slot_of[src]);
                    } else {
                        // A bulk, side or inlet node: through the lattice (mirrored only for side
                        // nodes).
                        code = make_code(kNode, my, mz, src);
                    }
                    break;
                }
                case kBounce: {
                    code = make_code(kBounce, false, false, 0);
                    // boundary_links order: z, y, x, direction (i >= 1 always: c_0 = 0)
                    const double q = link_fraction(c, solid, distance, static_cast<std::size_t>(x),
                                                   static_cast<std::size_t>(y), static_cast<std::size_t>(z), i);
                    t.links.push_back(Link{static_cast<std::uint32_t>(cell), static_cast<std::uint32_t>(i), q});
                    t.lslot.push_back(static_cast<std::uint32_t>(b));
                    if (!t.bq.empty()) {
                        t.bq[static_cast<std::size_t>(i) * t.bstride + b] = q;
                    }
                    break;
                }
                case kWall:
                    code = make_code(kWall, false, false, static_cast<std::uint32_t>(r.wall_side));
                    break;
                case kInlet:
                    code = make_code(kInlet, false, false, 0);
                    break;
                default:
                    code = make_code(kOutlet, my, mz, slot_of[cell_of(r.sx, r.sy, r.sz)]);
                    break;
            }
            t.bcode[static_cast<std::size_t>(i) * t.bstride + b] = code;
        }
    }
    // Which outputs of a boundary node need their lattice copy: output i of node w streams to
    // n = w + c_i through the sides (periodic: wrapped; slip: the row kept, the direction
    // mirrored); n reads it from the lattice when n is a bulk or side node, or when n is a
    // boundary node, reached without a mirror, whose link opposite(i) is a bounce-back (the far
    // read of a short link, also made by the force). Every other reader is a boundary node
    // reading w's stash (kStash), and targets through walls, inlet, outlet or into solids have
    // no reader.
    t.bwrite.assign(t.bstride, 0U);
    for (std::size_t b = 0; b < nb; ++b) {
        const std::size_t cell = t.bcell[b];
        const int z = static_cast<int>(cell / plane);
        const int y = static_cast<int>((cell % plane) / nx);
        const int x = static_cast<int>(cell % nx);
        std::uint32_t mask = 0;
        for (int i = 0; i < 19; ++i) {
            int tx = x + D3Q19::cx(i);
            int ty = y + D3Q19::cy(i);
            int tz = z + D3Q19::cz(i);
            bool inside = true;
            bool mirrored = false;
            if (ty < 0 || ty >= iny) {
                if (sides[ty < 0 ? 0 : 1] == LBM_WALL_PERIODIC) {
                    ty = ty < 0 ? ty + iny : ty - iny;
                } else if (sides[ty < 0 ? 0 : 1] == LBM_WALL_SLIP) {
                    ty = y;
                    mirrored = true;
                } else {
                    inside = false;
                }
            }
            if (tz < 0 || tz >= inz) {
                if (sides[tz < 0 ? 2 : 3] == LBM_WALL_PERIODIC) {
                    tz = tz < 0 ? tz + inz : tz - inz;
                } else if (sides[tz < 0 ? 2 : 3] == LBM_WALL_SLIP) {
                    tz = z;
                    mirrored = true;
                } else {
                    inside = false;
                }
            }
            if (tx < 0 || tx >= inx) {
                if (c.x_boundary == LBM_X_PERIODIC) {
                    tx = tx < 0 ? tx + inx : tx - inx;
                } else {
                    inside = false;
                }
            }
            if (!inside) {
                continue;
            }
            const std::uint32_t n = cell_of(tx, ty, tz);
            bool needed = t.cls[n] == 1 || t.cls[n] == 3 || t.cls[n] == 4;
            if (t.cls[n] == 2 && !mirrored) {
                const std::uint32_t far = t.bcode[static_cast<std::size_t>(D3Q19::opp(i)) * t.bstride + slot_of[n]];
                needed = static_cast<int>(far & 7U) == kBounce;
            }
            mask |= needed ? 1U << i : 0U;
        }
        t.bwrite[b] = mask;
    }
    return t;
}

Tables tables2(const lbm2d_config& c, const std::uint8_t* solid) {
    Tables t;
    const std::size_t nx = c.nx;
    const std::size_t ny = c.ny;
    const std::size_t cells = nx * ny;
    t.cls.assign(cells, 1);
    std::ptrdiff_t offset[9];
    for (int i = 0; i < 9; ++i) {
        offset[i] = -static_cast<std::ptrdiff_t>(D2Q9::cx(i)) - static_cast<std::ptrdiff_t>(nx) * D2Q9::cy(i);
    }
    // classify_nodes (lbm2d_cpu.cpp:316-341)
    for (std::size_t y = 0; y < ny; ++y) {
        for (std::size_t x = 0; x < nx; ++x) {
            const std::size_t here = x + nx * y;
            if (solid[here] != 0) {
                t.cls[here] = 0;
                continue;
            }
            if (x == 0 || y == 0 || x + 1 == nx || y + 1 == ny) {
                t.cls[here] = 2;
                continue;
            }
            for (int i = 1; i < 9; ++i) {
                if (solid[static_cast<std::size_t>(static_cast<std::ptrdiff_t>(here) + offset[i])] != 0) {
                    t.cls[here] = 2;
                    break;
                }
            }
        }
    }
    std::vector<std::uint32_t> slot_of(cells, 0xffffffffU);
    for (std::size_t cell = 0; cell < cells; ++cell) {
        if (t.cls[cell] == 2) {
            slot_of[cell] = static_cast<std::uint32_t>(t.bcell.size());
            t.bcell.push_back(static_cast<std::uint32_t>(cell));
        }
    }
    const std::size_t nb = t.bcell.size();
    t.bstride = padded(nb, 64);
    t.bcode.assign(static_cast<std::size_t>(9) * t.bstride, 0U);
    const int sides[2] = {c.bottom.kind, c.top.kind};
    const auto solid_at = [&](int a, int b, int) { return solid[static_cast<std::size_t>(a) + nx * static_cast<std::size_t>(b)] != 0; };
    for (std::size_t b = 0; b < nb; ++b) {
        const std::size_t cell = t.bcell[b];
        const int y = static_cast<int>(cell / nx);
        const int x = static_cast<int>(cell % nx);
        for (int i = 0; i < 9; ++i) {
            const Resolved r = resolve_d2q9(x, y, i, static_cast<int>(nx), static_cast<int>(ny), c.x_boundary, sides, solid_at);
            const std::uint32_t src = static_cast<std::uint32_t>(static_cast<std::size_t>(r.sx) + nx * static_cast<std::size_t>(r.sy));
            const bool my = (r.mirror & 1) != 0;
            std::uint32_t code = 0;
            switch (r.kind) {
                case kNode:
                    code = r.mirror == 0 ? make_code(kNode, false, false, src) : make_code(kNode, my, false, slot_of[src]);
                    break;
                case kBounce:
                    code = make_code(kBounce, false, false, 0);
                    t.links.push_back(Link{static_cast<std::uint32_t>(cell), static_cast<std::uint32_t>(i), 0.5});
                    t.lslot.push_back(static_cast<std::uint32_t>(b));
                    break;
                case kWall:
                    code = make_code(kWall, false, false, static_cast<std::uint32_t>(r.wall_side));
                    break;
                case kInlet:
                    code = make_code(kInlet, false, false, 0);
                    break;
                default:
                    code = make_code(kOutlet, my, false, slot_of[src]);
                    break;
            }
            t.bcode[static_cast<std::size_t>(i) * t.bstride + b] = code;
        }
    }
    return t;
}

void set_memory_limit(std::size_t bytes) { g_memory_limit = bytes; }

int device_memory(std::size_t* available, std::size_t* max_alloc) {
    std::size_t free_bytes = 0;
    std::size_t total = 0;
    const hipError_t e = hipMemGetInfo(&free_bytes, &total);
    if (e != hipSuccess) {
        return fail(LBM_EDEVICE, std::string("hipMemGetInfo: ") + hipGetErrorString(e));
    }
    if (g_memory_limit != 0) {
        free_bytes = std::min(free_bytes, g_memory_limit);
    }
    if (available != nullptr) {
        *available = free_bytes;
    }
    if (max_alloc != nullptr) {
        *max_alloc = free_bytes;
    }
    return ok();
}

// megabytes (lbm3d_gpu.cpp:84-89)
std::string megabytes(std::size_t bytes) {
    std::ostringstream out;
    out << std::fixed << std::setprecision(1) << static_cast<double>(bytes) / (1024.0 * 1024.0) << " MiB";
    return out.str();
}

} // namespace lbm::host

extern "C" {

const char* lbm_last_error(void) { return lbm::host::last_error(); }

const char* lbm_version(void) { return "orrery_lbm 0.1 gfx950"; }

int lbm_device_memory(size_t* available, size_t* max_alloc) { return lbm::host::device_memory(available, max_alloc); }

int lbm_device_name(char* name, size_t capacity) {
    if (name == nullptr || capacity == 0) {
        return lbm::host::fail(LBM_EINVAL, "lbm: null argument");
    }
    int dev = 0;
    hipDeviceProp_t prop{};
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) {
        e = hipGetDeviceProperties(&prop, dev);
    }
    if (e != hipSuccess) {
        return lbm::host::fail(LBM_EDEVICE, std::string("lbm: device properties: ") + hipGetErrorString(e));
    }
    const std::string text = std::string(prop.name) + " " + prop.gcnArchName + " (" +
                             std::to_string(prop.multiProcessorCount) + " CUs)";
    const std::size_t n = std::min(capacity - 1, text.size());
    std::memcpy(name, text.data(), n);
    name[n] = '\0';
    return LBM_OK;
}

int lbm_set_memory_limit(size_t bytes) {
    lbm::host::set_memory_limit(bytes);
    return LBM_OK;
}

int lbm3d_validate(const lbm3d_config* config, const uint8_t* solid, const double* distance) {
    if (config == nullptr) {
        return lbm::host::fail(LBM_EINVAL, "lbm3d: null argument");
    }
    const std::string p = lbm::host::problem3(*config, solid, distance);
    return p.empty() ? LBM_OK : lbm::host::fail(LBM_EINVAL, p);
}

int lbm2d_validate(const lbm2d_config* config, const uint8_t* solid) {
    if (config == nullptr) {
        return lbm::host::fail(LBM_EINVAL, "lbm2d: null argument");
    }
    const std::string p = lbm::host::problem2(*config, solid);
    return p.empty() ? LBM_OK : lbm::host::fail(LBM_EINVAL, p);
}

int lbm3d_classify(const lbm3d_config* config, const uint8_t* solid, uint8_t* classes) {
    if (config == nullptr || classes == nullptr) {
        return lbm::host::fail(LBM_EINVAL, "lbm3d: null argument");
    }
    // Node classes do not depend on the wall model: validate and classify with the halfway wall.
    lbm3d_config c = *config;
    c.solid_wall = LBM_SOLID_HALFWAY;
    if (int rc = lbm3d_validate(&c, solid, nullptr); rc != LBM_OK) {
        return rc;
    }
    const lbm::host::Tables t = lbm::host::tables3(c, solid, nullptr);
    // the reference's classes: side and inlet nodes are boundary nodes
    std::transform(t.cls.begin(), t.cls.end(), classes, [](std::uint8_t k) { return k >= 3 ? std::uint8_t{2} : k; });
    return LBM_OK;
}

int lbm3d_boundary_links(const lbm3d_config* config, const uint8_t* solid, const double* distance, uint32_t* cell,
                         uint32_t* direction, double* q, size_t capacity, size_t* count) {
    if (int rc = lbm3d_validate(config, solid, distance); rc != LBM_OK) {
        return rc;
    }
    const lbm::host::Tables t = lbm::host::tables3(*config, solid, distance);
    for (std::size_t k = 0; k < t.links.size() && k < capacity; ++k) {
        if (cell != nullptr) cell[k] = t.links[k].cell;
        if (direction != nullptr) direction[k] = t.links[k].direction;
        if (q != nullptr) q[k] = t.links[k].q;
    }
    if (count != nullptr) {
        *count = t.links.size();
    }
    return LBM_OK;
}

int lbm2d_boundary_links(const lbm2d_config* config, const uint8_t* solid, uint32_t* cell, uint32_t* direction,
                         size_t capacity, size_t* count) {
    if (int rc = lbm2d_validate(config, solid); rc != LBM_OK) {
        return rc;
    }
    const lbm::host::Tables t = lbm::host::tables2(*config, solid);
    for (std::size_t k = 0; k < t.links.size() && k < capacity; ++k) {
        if (cell != nullptr) cell[k] = t.links[k].cell;
        if (direction != nullptr) direction[k] = t.links[k].direction;
    }
    if (count != nullptr) {
        *count = t.links.size();
    }
    return LBM_OK;
}

// SplitMix64 finaliser (Steele, Lea and Flood 2014) keyed by (seed, cell, component); the
// device copy sits in the init kernels.
double lbm_perturbation(uint64_t seed, uint64_t cell, uint32_t component) {
    uint64_t z = seed + 0x9E3779B97F4A7C15ULL * (cell * 4U + component + 1U);
    z = (z ^ (z >> 30U)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27U)) * 0x94D049BB133111EBULL;
    z ^= z >> 31U;
    const double u = static_cast<double>(z >> 11U) * (1.0 / 9007199254740992.0);
    return 2.0 * u - 1.0;
}

} // extern "C"
