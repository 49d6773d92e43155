// TEST INFRASTRUCTURE ONLY -- the oracle. A CPU restatement, in binary64, of the reference
// solver's algorithm for the D3Q19 and D2Q9 lattice Boltzmann step (pull streaming into a
// second buffer, every boundary rule resolved per link, BGK + Guo forcing + Smagorinsky +
// absorbing layers). Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load
// it, and only to check results; the product library never links it.
//
// Each function names the reference lines it restates (paths relative to the reference
// root). Expression order follows the reference so that rounding matches; the library is
// compiled with -ffp-contract=off -fno-fast-math as the reference is (cmake/FpDeterminism.cmake).
//
// Pinning (tests/test_oracle_*.py): golden values recorded from the reference (Poiseuille
// 4x24x4 velocity, C1/C2 link counts), the reference's own closed-form tests (Poiseuille,
// duct, Couette, co-moving walls, uniform stream, mass, momentum balance, Smagorinsky shear,
// interpolated-wall fractions), and oracle/_ref (reference setup code compiled from its own
// sources) for resolve_source, classify_nodes, boundary_links, link_fraction, sponge_sigma,
// velocity_inlet_west and zou_he_west. collide() itself sits in lbm3d_cpu.cpp, which needs
// oneTBB (absent), so it is pinned through the closed forms and the recorded golden values.
#include <omp.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

#include "../include/orrery_lbm.h"

namespace {

constexpr std::size_t kMaxGridCells = std::size_t{1} << 27U; // src/geom/include/orrery/geom/voxel_grid.h:29
constexpr double kMaxLatticeMach = 0.1;                      // lattice_units.h:475
constexpr double kSmagorinskyFactor = 18.0 * 1.4142135623730951; // lbm3d_cpu.cpp:26

double lattice_mach(double u) { return std::abs(u) / std::sqrt(1.0 / 3.0); } // lattice_units.cpp:18-20

// lbm3d_common.cpp:21-30 / lbm2d_common.cpp:18-27 (check_mach); empty when below the limit.
std::string check_mach(const char* prefix, const char* what, double speed) {
    if (!(lattice_mach(speed) < kMaxLatticeMach)) {
        std::ostringstream message;
        message << prefix << what << " " << speed << " has lattice Mach number " << lattice_mach(speed)
                << ", not below " << kMaxLatticeMach << " (APPROXIMATION A3)";
        return message.str();
    }
    return {};
}

void put(char* err, std::size_t len, const std::string& msg) {
    if (err != nullptr && len > 0) {
        std::snprintf(err, len, "%s", msg.c_str());
    }
}

// ------------------------------------------------------------------------------------------
// D3Q19 tables (src/aero/lbm3d/include/orrery/aero/lbm3d/d3q19.h:15-36)
namespace d3 {
constexpr int Q = 19;
constexpr int cx[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
constexpr int cy[Q] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
constexpr int cz[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
constexpr double w[Q] = {1.0 / 3.0,  1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0,
                         1.0 / 18.0, 1.0 / 18.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
                         1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
                         1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
constexpr int opp[Q] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
constexpr int mirror_y[Q] = {0, 1, 2, 4, 3, 5, 6, 9, 10, 7, 8, 11, 12, 13, 14, 18, 17, 16, 15};
constexpr int mirror_z[Q] = {0, 1, 2, 3, 4, 6, 5, 7, 8, 9, 10, 13, 14, 11, 12, 17, 18, 15, 16};
constexpr int outlet_dirs[5] = {2, 8, 10, 12, 14}; // lbm3d_internal.h:32
constexpr int inlet_dirs[5] = {1, 7, 9, 11, 13};   // lbm3d_internal.h:48

// lbm3d_internal.h:33-46
int outlet_slot(int i) {
    switch (i) {
        case 2: return 0;
        case 8: return 1;
        case 10: return 2;
        case 12: return 3;
        default: return 4;
    }
}

// d3q19.h:49-55
double equilibrium(int i, double rho, double ux, double uy, double uz) {
    const double cu = static_cast<double>(cx[i]) * ux + static_cast<double>(cy[i]) * uy +
                      static_cast<double>(cz[i]) * uz;
    const double uu = ux * ux + uy * uy + uz * uz;
    return w[i] * rho * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * uu);
}
} // namespace d3

enum Kind : int { kNode = 0, kBounce = 1, kWall = 2, kInlet = 3, kOutlet = 4 };

struct Source {
    int kind = kNode;
    std::size_t cell = 0;
    int direction = 0;
    double wall_ux = 0.0;
};

struct Side {
    bool wall = false;
    double wall_ux = 0.0;
};

// lbm3d_internal.h:65-85 (apply_side)
Side apply_side(const lbm_wall& low, const lbm_wall& high, std::ptrdiff_t& s, std::ptrdiff_t n,
                std::ptrdiff_t own, int& direction, const int* mirror) {
    if (s >= 0 && s < n) {
        return {};
    }
    const lbm_wall& wall = s < 0 ? low : high;
    switch (wall.kind) {
        case LBM_WALL_PERIODIC:
            s = (s + n) % n;
            break;
        case LBM_WALL_SLIP:
            s = own;
            direction = mirror[direction];
            break;
        default:
            return Side{true, wall.velocity_x};
    }
    return {};
}

std::size_t index3(const lbm3d_config& c, std::size_t x, std::size_t y, std::size_t z) {
    return x + c.nx * (y + c.ny * z);
}

// lbm3d_internal.h:90-130 (resolve_source): y sides, then z sides, then x.
Source resolve3(const lbm3d_config& c, const std::uint8_t* solid, std::size_t x, std::size_t y,
                std::size_t z, int i) {
    const auto nx = static_cast<std::ptrdiff_t>(c.nx);
    const auto ny = static_cast<std::ptrdiff_t>(c.ny);
    const auto nz = static_cast<std::ptrdiff_t>(c.nz);
    const std::size_t here = index3(c, x, y, z);
    std::ptrdiff_t sx = static_cast<std::ptrdiff_t>(x) - d3::cx[i];
    std::ptrdiff_t sy = static_cast<std::ptrdiff_t>(y) - d3::cy[i];
    std::ptrdiff_t sz = static_cast<std::ptrdiff_t>(z) - d3::cz[i];
    int direction = i;
    Side side = apply_side(c.y_low, c.y_high, sy, ny, static_cast<std::ptrdiff_t>(y), direction,
                           d3::mirror_y);
    if (side.wall) {
        return Source{kWall, here, d3::opp[i], side.wall_ux};
    }
    side = apply_side(c.z_low, c.z_high, sz, nz, static_cast<std::ptrdiff_t>(z), direction,
                      d3::mirror_z);
    if (side.wall) {
        return Source{kWall, here, d3::opp[i], side.wall_ux};
    }
    if (sx < 0 || sx >= nx) {
        if (c.x_boundary == LBM_X_PERIODIC) {
            sx = (sx + nx) % nx;
        } else if (sx < 0) {
            return Source{kInlet, here, i, 0.0};
        } else {
            const std::size_t face = index3(c, c.nx - 1, static_cast<std::size_t>(sy),
                                            static_cast<std::size_t>(sz));
            return Source{kOutlet, face, direction, 0.0};
        }
    }
    const std::size_t cell = index3(c, static_cast<std::size_t>(sx), static_cast<std::size_t>(sy),
                                    static_cast<std::size_t>(sz));
    if (solid[cell] != 0) {
        return Source{kBounce, here, d3::opp[i], 0.0};
    }
    return Source{kNode, cell, direction, 0.0};
}

// lbm3d_internal.h:139-179 (link_fraction)
double link_fraction3(const lbm3d_config& c, const std::uint8_t* solid, const double* distance,
                      std::size_t x, std::size_t y, std::size_t z, int i) {
    if (c.solid_wall != LBM_SOLID_INTERPOLATED) {
        return 0.5;
    }
    const auto sx = static_cast<std::ptrdiff_t>(x) - d3::cx[i];
    const auto sy = static_cast<std::ptrdiff_t>(y) - d3::cy[i];
    const auto sz = static_cast<std::ptrdiff_t>(z) - d3::cz[i];
    const auto inside = [&](std::ptrdiff_t a, std::ptrdiff_t b, std::ptrdiff_t d) {
        return a >= 0 && b >= 0 && d >= 0 && a < static_cast<std::ptrdiff_t>(c.nx) &&
               b < static_cast<std::ptrdiff_t>(c.ny) && d < static_cast<std::ptrdiff_t>(c.nz);
    };
    if (!inside(sx, sy, sz)) {
        return 0.5;
    }
    const double fluid = distance[index3(c, x, y, z)];
    const double sol = distance[index3(c, static_cast<std::size_t>(sx), static_cast<std::size_t>(sy),
                                       static_cast<std::size_t>(sz))];
    if (!std::isfinite(fluid) || !std::isfinite(sol) || fluid < 0.0 || sol > 0.0 ||
        (fluid == 0.0 && sol == 0.0)) {
        return 0.5;
    }
    const double span = fluid - sol;
    const double ratio = std::isfinite(span) ? fluid / span : 0.5 * fluid / (0.5 * fluid - 0.5 * sol);
    const double q = std::min(ratio, 1.0);
    if (q < 0.5) {
        const auto fx = static_cast<std::ptrdiff_t>(x) + d3::cx[i];
        const auto fy = static_cast<std::ptrdiff_t>(y) + d3::cy[i];
        const auto fz = static_cast<std::ptrdiff_t>(z) + d3::cz[i];
        if (!inside(fx, fy, fz) ||
            solid[index3(c, static_cast<std::size_t>(fx), static_cast<std::size_t>(fy),
                         static_cast<std::size_t>(fz))] != 0) {
            return 0.5;
        }
    }
    return q;
}

// lbm3d_internal.h:256-281 (sponge_sigma)
double sponge_sigma3(const lbm3d_config& c, std::size_t x, std::size_t y, std::size_t z) {
    if (!(c.sponge_strength > 0.0) || c.x_boundary != LBM_X_INFLOW_OUTFLOW) {
        return 0.0;
    }
    double best = 0.0;
    const auto layer = [&best](std::size_t distance, std::size_t thickness) {
        if (distance < thickness) {
            const double r = static_cast<double>(thickness - distance) / static_cast<double>(thickness);
            best = std::max(best, r * r);
        }
    };
    layer(x, c.sponge_inlet);
    layer(c.nx - 1 - x, c.sponge_outlet);
    if (c.y_low.kind != LBM_WALL_PERIODIC) {
        layer(y, c.sponge_sides);
        layer(c.ny - 1 - y, c.sponge_sides);
    }
    if (c.z_low.kind != LBM_WALL_PERIODIC) {
        layer(z, c.sponge_sides);
        layer(c.nz - 1 - z, c.sponge_sides);
    }
    return c.sponge_strength * best;
}

// lbm3d_internal.h:324-365 (velocity_inlet_west)
void inlet_west3(double* f, double ux, double uy, double uz) {
    double known = 0.0;
    for (int i = 0; i < d3::Q; ++i) {
        if (d3::cx[i] == 0) {
            known += f[i];
        } else if (d3::cx[i] < 0) {
            known += 2.0 * f[i];
        }
    }
    const double rho = known / (1.0 - ux);
    for (const int i : d3::inlet_dirs) {
        const double cu = static_cast<double>(d3::cx[i]) * ux + static_cast<double>(d3::cy[i]) * uy +
                          static_cast<double>(d3::cz[i]) * uz;
        f[i] = f[d3::opp[i]] + 6.0 * d3::w[i] * rho * cu;
    }
    double feq[d3::Q];
    double pi[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < d3::Q; ++i) {
        feq[i] = d3::equilibrium(i, rho, ux, uy, uz);
        const double neq = f[i] - feq[i];
        const auto ex = static_cast<double>(d3::cx[i]);
        const auto ey = static_cast<double>(d3::cy[i]);
        const auto ez = static_cast<double>(d3::cz[i]);
        pi[0] += ex * ex * neq;
        pi[1] += ey * ey * neq;
        pi[2] += ez * ez * neq;
        pi[3] += ex * ey * neq;
        pi[4] += ex * ez * neq;
        pi[5] += ey * ez * neq;
    }
    const double cs2 = 1.0 / 3.0;
    for (int i = 0; i < d3::Q; ++i) {
        const auto ex = static_cast<double>(d3::cx[i]);
        const auto ey = static_cast<double>(d3::cy[i]);
        const auto ez = static_cast<double>(d3::cz[i]);
        f[i] = feq[i] + 4.5 * d3::w[i] *
                            ((ex * ex - cs2) * pi[0] + (ey * ey - cs2) * pi[1] + (ez * ez - cs2) * pi[2] +
                             2.0 * (ex * ey * pi[3] + ex * ez * pi[4] + ey * ez * pi[5]));
    }
}

// lbm3d_cpu.cpp:32-100 (collide): BGK + Guo forcing + Smagorinsky + absorbing layer.
void collide3(double* f, const lbm3d_config& c, double sponge, const double* stream) {
    double rho = 0.0;
    for (int i = 0; i < d3::Q; ++i) {
        rho += f[i];
    }
    const double jx = f[1] - f[2] + f[7] - f[8] + f[9] - f[10] + f[11] - f[12] + f[13] - f[14];
    const double jy = f[3] - f[4] + f[7] - f[8] - f[9] + f[10] + f[15] - f[16] + f[17] - f[18];
    const double jz = f[5] - f[6] + f[11] - f[12] - f[13] + f[14] + f[15] - f[16] - f[17] + f[18];
    const double fx = c.body_force_x;
    const double fy = c.body_force_y;
    const double fz = c.body_force_z;
    const double ux = (jx + 0.5 * fx) / rho;
    const double uy = (jy + 0.5 * fy) / rho;
    const double uz = (jz + 0.5 * fz) / rho;
    const double base = 1.0 - 1.5 * (ux * ux + uy * uy + uz * uz);
    const double cu[d3::Q] = {0.0,     ux,       -ux,      uy,      -uy,     uz,       -uz,
                              ux + uy, -ux - uy, ux - uy,  uy - ux, ux + uz, -ux - uz, ux - uz,
                              uz - ux, uy + uz,  -uy - uz, uy - uz, uz - uy};
    double feq[d3::Q];
    for (int i = 0; i < d3::Q; ++i) {
        feq[i] = d3::w[i] * rho * (base + cu[i] * (3.0 + 4.5 * cu[i]));
    }
    double tau = c.tau;
    if (c.smagorinsky_enabled != 0) {
        double pi[6] = {0, 0, 0, 0, 0, 0};
        for (int i = 0; i < d3::Q; ++i) {
            const double neq = f[i] - feq[i];
            const auto ex = static_cast<double>(d3::cx[i]);
            const auto ey = static_cast<double>(d3::cy[i]);
            const auto ez = static_cast<double>(d3::cz[i]);
            pi[0] += ex * ex * neq;
            pi[1] += ey * ey * neq;
            pi[2] += ez * ez * neq;
            pi[3] += ex * ey * neq;
            pi[4] += ex * ez * neq;
            pi[5] += ey * ez * neq;
        }
        const double strain = std::sqrt(pi[0] * pi[0] + pi[1] * pi[1] + pi[2] * pi[2] +
                                        2.0 * (pi[3] * pi[3] + pi[4] * pi[4] + pi[5] * pi[5]));
        const double cc = c.smagorinsky_constant;
        tau = 0.5 * (tau + std::sqrt(tau * tau + kSmagorinskyFactor * cc * cc * strain / rho));
    }
    const double omega = 1.0 / tau;
    const bool forced = fx != 0.0 || fy != 0.0 || fz != 0.0;
    const double force_scale = 1.0 - 0.5 * omega;
    for (int i = 0; i < d3::Q; ++i) {
        double value = f[i] - omega * (f[i] - feq[i]);
        if (forced) {
            const auto ex = static_cast<double>(d3::cx[i]);
            const auto ey = static_cast<double>(d3::cy[i]);
            const auto ez = static_cast<double>(d3::cz[i]);
            value += force_scale * d3::w[i] *
                     (3.0 * ((ex - ux) * fx + (ey - uy) * fy + (ez - uz) * fz) +
                      9.0 * cu[i] * (ex * fx + ey * fy + ez * fz));
        }
        f[i] = value;
    }
    if (sponge > 0.0) {
        for (int i = 0; i < d3::Q; ++i) {
            const double target = d3::equilibrium(i, 1.0, stream[0], stream[1], stream[2]);
            f[i] -= sponge * (f[i] - target);
        }
    }
}

struct Link3 {
    std::uint32_t cell;
    std::uint32_t direction;
    double q;
};

// lbm3d_common.cpp:74-155 (problem_of); empty string when the configuration can run.
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
    if (!std::isfinite(c.tau) || !(c.tau > 0.5)) {
        return "lbm3d: tau must be finite and above 1/2 (positive viscosity)";
    }
    for (const double v : {c.inlet_ux, c.inlet_uy, c.inlet_uz, c.body_force_x, c.body_force_y,
                           c.body_force_z, c.y_low.velocity_x, c.y_high.velocity_x,
                           c.z_low.velocity_x, c.z_high.velocity_x}) {
        if (!std::isfinite(v)) {
            return "lbm3d: velocities and body force must be finite";
        }
    }
    if (c.smagorinsky_enabled != 0 &&
        (!std::isfinite(c.smagorinsky_constant) || c.smagorinsky_constant < 0.0)) {
        return "lbm3d: the Smagorinsky constant must be finite and non-negative";
    }
    if ((c.y_low.kind == LBM_WALL_PERIODIC) != (c.y_high.kind == LBM_WALL_PERIODIC) ||
        (c.z_low.kind == LBM_WALL_PERIODIC) != (c.z_high.kind == LBM_WALL_PERIODIC)) {
        return "lbm3d: an axis is periodic on both sides or on neither";
    }
    const char* names[4] = {"y = 0 wall speed", "y = ny - 1 wall speed", "z = 0 wall speed",
                            "z = nz - 1 wall speed"};
    const lbm_wall* walls[4] = {&c.y_low, &c.y_high, &c.z_low, &c.z_high};
    for (int k = 0; k < 4; ++k) {
        std::string bad = check_mach("lbm3d: ", names[k], walls[k]->velocity_x);
        if (!bad.empty()) {
            return bad;
        }
    }
    if (!std::isfinite(c.sponge_strength) || c.sponge_strength < 0.0 || c.sponge_strength > 1.0) {
        return "lbm3d: the absorbing-layer strength must be in [0, 1]";
    }
    if (c.sponge_strength > 0.0 &&
        (c.x_boundary != LBM_X_INFLOW_OUTFLOW || c.sponge_inlet > c.nx || c.sponge_outlet > c.nx ||
         c.sponge_sides > std::min(c.ny, c.nz))) {
        return "lbm3d: absorbing layers need the inflow/outflow x boundary and must fit in the "
               "lattice";
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
        const double speed =
            std::sqrt(c.inlet_ux * c.inlet_ux + c.inlet_uy * c.inlet_uy + c.inlet_uz * c.inlet_uz);
        std::string bad = check_mach("lbm3d: ", "inlet speed", speed);
        if (!bad.empty()) {
            return bad;
        }
        for (std::size_t z = 0; z < c.nz; ++z) {
            for (std::size_t y = 0; y < c.ny; ++y) {
                if (solid[index3(c, 0, y, z)] != 0 || solid[index3(c, c.nx - 1, y, z)] != 0) {
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

class Oracle3D {
public:
    Oracle3D(const lbm3d_config& config, const std::uint8_t* solid, const double* distance)
        : c_(config), cells_(config.nx * config.ny * config.nz), solid_(solid, solid + cells_) {
        if (distance != nullptr) {
            distance_.assign(distance, distance + cells_);
        }
        // lbm3d_common.cpp:176-198 (boundary_links): z slabs in parallel, concatenated in
        // z order, so the list keeps the reference's (z, y, x, direction) order.
        std::vector<std::vector<Link3>> slabs(c_.nz);
#pragma omp parallel for schedule(dynamic)
        for (std::ptrdiff_t zz = 0; zz < static_cast<std::ptrdiff_t>(c_.nz); ++zz) {
            const auto z = static_cast<std::size_t>(zz);
            for (std::size_t y = 0; y < c_.ny; ++y) {
                for (std::size_t x = 0; x < c_.nx; ++x) {
                    const std::size_t here = index3(c_, x, y, z);
                    if (solid_[here] != 0) {
                        continue;
                    }
                    for (int i = 1; i < d3::Q; ++i) {
                        if (resolve3(c_, solid_.data(), x, y, z, i).kind == kBounce) {
                            slabs[z].push_back(Link3{static_cast<std::uint32_t>(here),
                                                     static_cast<std::uint32_t>(i),
                                                     link_fraction3(c_, solid_.data(), distance_.data(),
                                                                    x, y, z, i)});
                        }
                    }
                }
            }
        }
        for (const std::vector<Link3>& slab : slabs) {
            links_.insert(links_.end(), slab.begin(), slab.end());
        }
        // lbm3d_internal.h:201-212 (source_offsets)
        const auto nx = static_cast<std::ptrdiff_t>(c_.nx);
        const auto plane = static_cast<std::ptrdiff_t>(c_.nx * c_.ny);
        for (int i = 0; i < d3::Q; ++i) {
            offset_[i] = -static_cast<std::ptrdiff_t>(d3::cx[i]) - nx * d3::cy[i] - plane * d3::cz[i];
        }
        classify();
        // lbm3d_internal.h:293-310 (wall_table)
        if (c_.solid_wall == LBM_SOLID_INTERPOLATED) {
            slot_.assign(cells_, kNoSlot);
            std::uint32_t next = 0;
            for (const Link3& link : links_) {
                std::uint32_t& s = slot_[link.cell];
                if (s == kNoSlot) {
                    s = next++;
                    wall_q_.insert(wall_q_.end(), d3::Q, 0.5);
                }
                wall_q_[static_cast<std::size_t>(s) * d3::Q + link.direction] = link.q;
            }
        }
        initialize(nullptr);
    }

    void initialize(const double* moments) {
        // lbm3d_common.cpp:243-270 (initial_populations, uniform_inlet_state)
        f_.assign(d3::Q * cells_, 0.0);
        const bool inflow = c_.x_boundary == LBM_X_INFLOW_OUTFLOW;
#pragma omp parallel for schedule(static)
        for (std::size_t cell = 0; cell < cells_; ++cell) {
            double rho = 1.0;
            double ux = 0.0;
            double uy = 0.0;
            double uz = 0.0;
            if (solid_[cell] == 0) {
                if (moments != nullptr) {
                    rho = moments[cell];
                    ux = moments[cells_ + cell];
                    uy = moments[2 * cells_ + cell];
                    uz = moments[3 * cells_ + cell];
                } else if (inflow) {
                    ux = c_.inlet_ux;
                    uy = c_.inlet_uy;
                    uz = c_.inlet_uz;
                }
            }
            for (int i = 0; i < d3::Q; ++i) {
                f_[i * cells_ + cell] = d3::equilibrium(i, rho, ux, uy, uz);
            }
        }
        g_ = f_;
        reset_outlet();
        inlet_ = {c_.inlet_ux, c_.inlet_uy, c_.inlet_uz};
        steps_ = 0;
    }

    void step(std::uint64_t count) {
        const std::size_t rows = c_.ny * c_.nz;
        for (std::uint64_t n = 0; n < count; ++n) {
#pragma omp parallel for schedule(static)
            for (std::ptrdiff_t r = 0; r < static_cast<std::ptrdiff_t>(rows); ++r) {
                update_row(static_cast<std::size_t>(r) % c_.ny, static_cast<std::size_t>(r) / c_.ny);
            }
            std::swap(f_, g_);
            ++steps_;
        }
    }

    // lbm3d_cpu.cpp:154-186
    void macroscopic(double* rho, double* ux, double* uy, double* uz) const {
        for (std::size_t cell = 0; cell < cells_; ++cell) {
            double r = 1.0;
            double vx = 0.0;
            double vy = 0.0;
            double vz = 0.0;
            if (solid_[cell] == 0) {
                double s = 0.0;
                double jx = 0.0;
                double jy = 0.0;
                double jz = 0.0;
                for (int i = 0; i < d3::Q; ++i) {
                    const double v = f_[i * cells_ + cell];
                    s += v;
                    jx += static_cast<double>(d3::cx[i]) * v;
                    jy += static_cast<double>(d3::cy[i]) * v;
                    jz += static_cast<double>(d3::cz[i]) * v;
                }
                r = s;
                vx = (jx - 0.5 * c_.body_force_x) / s;
                vy = (jy - 0.5 * c_.body_force_y) / s;
                vz = (jz - 0.5 * c_.body_force_z) / s;
            }
            if (rho != nullptr) rho[cell] = r;
            if (ux != nullptr) ux[cell] = vx;
            if (uy != nullptr) uy[cell] = vy;
            if (uz != nullptr) uz[cell] = vz;
        }
    }

    // lbm3d_cpu.cpp:192-203 with add_link_force / link_wall_point (lbm3d_common.cpp:213-241)
    void force(double* F, double* M) const {
        double force[3] = {0, 0, 0};
        double moment[3] = {0, 0, 0};
        const std::size_t plane = c_.nx * c_.ny;
        for (const Link3& link : links_) {
            const int out = d3::opp[link.direction];
            const auto farther = static_cast<std::size_t>(static_cast<std::ptrdiff_t>(link.cell) -
                                                          offset_[link.direction]);
            const double back = reflected(link.cell, farther, static_cast<int>(link.direction), link.q);
            const double exchanged = f_[out * cells_ + link.cell] + back;
            const double fv[3] = {static_cast<double>(d3::cx[out]) * exchanged,
                                  static_cast<double>(d3::cy[out]) * exchanged,
                                  static_cast<double>(d3::cz[out]) * exchanged};
            const std::size_t z = link.cell / plane;
            const std::size_t rest = link.cell % plane;
            const std::size_t y = rest / c_.nx;
            const std::size_t x = rest % c_.nx;
            const double r[3] = {static_cast<double>(x) - link.q * d3::cx[link.direction],
                                 static_cast<double>(y) - link.q * d3::cy[link.direction],
                                 static_cast<double>(z) - link.q * d3::cz[link.direction]};
            for (int a = 0; a < 3; ++a) {
                force[a] += fv[a];
            }
            moment[0] += r[1] * fv[2] - r[2] * fv[1];
            moment[1] += r[2] * fv[0] - r[0] * fv[2];
            moment[2] += r[0] * fv[1] - r[1] * fv[0];
        }
        for (int a = 0; a < 3; ++a) {
            F[a] = force[a];
            M[a] = moment[a];
        }
    }

    // lbm3d_cpu.cpp:205-215
    double mass() const {
        double total = 0.0;
        for (std::size_t cell = 0; cell < cells_; ++cell) {
            if (solid_[cell] == 0) {
                for (int i = 0; i < d3::Q; ++i) {
                    total += f_[i * cells_ + cell];
                }
            }
        }
        return total;
    }

    const lbm3d_config& config() const { return c_; }
    const std::vector<double>& populations() const { return f_; }
    const std::vector<Link3>& links() const { return links_; }
    // Test hooks: any population state (the logical layout f[i * cells + cell]), and the
    // convective-outlet history [5][nz][ny] (slot-major).
    void set_populations(const double* f) { std::copy(f, f + f_.size(), f_.begin()); }
    const std::vector<double>& outlet_history() const { return outlet_; }
    std::uint64_t steps() const { return steps_; }
    void set_inlet(const double* v) { inlet_ = {v[0], v[1], v[2]}; }
    const std::array<double, 3>& inlet() const { return inlet_; }
    const std::vector<std::uint8_t>& classes() const { return class_; }

private:
    static constexpr std::uint32_t kNoSlot = 0xffffffffU;

    // lbm3d_internal.h:219-248 (classify_nodes)
    void classify() {
        class_.assign(cells_, 1);
        for (std::size_t z = 0; z < c_.nz; ++z) {
            for (std::size_t y = 0; y < c_.ny; ++y) {
                for (std::size_t x = 0; x < c_.nx; ++x) {
                    const std::size_t here = index3(c_, x, y, z);
                    if (solid_[here] != 0) {
                        class_[here] = 0;
                        continue;
                    }
                    if (x == 0 || y == 0 || z == 0 || x + 1 == c_.nx || y + 1 == c_.ny ||
                        z + 1 == c_.nz) {
                        class_[here] = 2;
                        continue;
                    }
                    for (int i = 1; i < d3::Q; ++i) {
                        const auto source =
                            static_cast<std::size_t>(static_cast<std::ptrdiff_t>(here) + offset_[i]);
                        if (solid_[source] != 0) {
                            class_[here] = 2;
                            break;
                        }
                    }
                }
            }
        }
    }

    // lbm3d_internal.h:185-197 (reflected)
    double reflected(std::size_t here, std::size_t farther, int i, double q) const {
        const int o = d3::opp[i];
        const double out = f_[o * cells_ + here];
        if (q == 0.5) {
            return out;
        }
        if (q < 0.5) {
            return 2.0 * q * out + (1.0 - 2.0 * q) * f_[o * cells_ + farther];
        }
        return (out + (2.0 * q - 1.0) * f_[i * cells_ + here]) / (2.0 * q);
    }

    std::size_t outlet_index(int slot, std::size_t y, std::size_t z) const {
        return static_cast<std::size_t>(slot) * c_.ny * c_.nz + y + c_.ny * z;
    }

    // lbm3d_cpu.cpp:227-237
    void reset_outlet() {
        outlet_.assign(5 * c_.ny * c_.nz, 0.0);
        for (std::size_t z = 0; z < c_.nz; ++z) {
            for (std::size_t y = 0; y < c_.ny; ++y) {
                const std::size_t cell = index3(c_, c_.nx - 1, y, z);
                for (const int i : d3::outlet_dirs) {
                    outlet_[outlet_index(d3::outlet_slot(i), y, z)] = f_[i * cells_ + cell];
                }
            }
        }
    }

    // lbm3d_cpu.cpp:242-250
    double outlet_value(double upstream, int i, std::size_t y, std::size_t z) {
        if (c_.outlet == LBM_OUTLET_EXTRAPOLATION) {
            return upstream;
        }
        const double u = c_.inlet_ux;
        double& previous = outlet_[outlet_index(d3::outlet_slot(i), y, z)];
        previous = (previous + u * upstream) / (1.0 + u);
        return previous;
    }

    double density_at(std::size_t cell) const {
        double rho = 0.0;
        for (int i = 0; i < d3::Q; ++i) {
            rho += f_[i * cells_ + cell];
        }
        return rho;
    }

    // lbm3d_cpu.cpp:261-322 (update_row)
    void update_row(std::size_t y, std::size_t z) {
        const bool inlet = c_.x_boundary == LBM_X_INFLOW_OUTFLOW;
        const std::size_t row = index3(c_, 0, y, z);
        for (std::size_t x = 0; x < c_.nx; ++x) {
            const std::size_t here = row + x;
            const std::uint8_t node = class_[here];
            if (node == 0) {
                continue;
            }
            double f[d3::Q];
            if (node == 1) {
                for (int i = 0; i < d3::Q; ++i) {
                    f[i] = f_[i * cells_ + static_cast<std::size_t>(static_cast<std::ptrdiff_t>(here) + offset_[i])];
                }
            } else {
                double wall_rho = -1.0;
                for (int i = 0; i < d3::Q; ++i) {
                    const Source s = resolve3(c_, solid_.data(), x, y, z, i);
                    switch (s.kind) {
                        case kNode:
                            f[i] = f_[s.direction * cells_ + s.cell];
                            break;
                        case kBounce:
                            f[i] = slot_.empty() || slot_[here] == kNoSlot
                                       ? f_[s.direction * cells_ + s.cell]
                                       : reflected(here,
                                                   static_cast<std::size_t>(static_cast<std::ptrdiff_t>(here) - offset_[i]),
                                                   i, wall_q_[static_cast<std::size_t>(slot_[here]) * d3::Q + i]);
                            break;
                        case kWall:
                            if (wall_rho < 0.0) {
                                wall_rho = density_at(here);
                            }
                            f[i] = f_[s.direction * cells_ + s.cell] +
                                   6.0 * d3::w[i] * wall_rho * static_cast<double>(d3::cx[i]) * s.wall_ux;
                            break;
                        case kInlet:
                            f[i] = 0.0;
                            break;
                        default:
                            f[i] = outlet_value(f_[s.direction * cells_ + s.cell], i, y, z);
                            break;
                    }
                }
                if (inlet && x == 0) {
                    inlet_west3(f, inlet_[0], inlet_[1], inlet_[2]);
                }
            }
            collide3(f, c_, sponge_sigma3(c_, x, y, z), inlet_.data());
            for (int i = 0; i < d3::Q; ++i) {
                g_[i * cells_ + here] = f[i];
            }
        }
    }

    lbm3d_config c_;
    std::size_t cells_;
    std::vector<std::uint8_t> solid_;
    std::vector<double> distance_;
    std::vector<Link3> links_;
    std::array<std::ptrdiff_t, d3::Q> offset_{};
    std::vector<std::uint8_t> class_;
    std::vector<std::uint32_t> slot_;
    std::vector<double> wall_q_;
    std::vector<double> f_;
    std::vector<double> g_;
    std::vector<double> outlet_;
    std::array<double, 3> inlet_{};
    std::uint64_t steps_ = 0;
};

// ------------------------------------------------------------------------------------------
// D2Q9 (src/aero/lbm/include/orrery/aero/lbm/d2q9.h:12-41)
namespace d2 {
constexpr int Q = 9;
constexpr int cx[Q] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
constexpr int cy[Q] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
constexpr double w[Q] = {4.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0, 1.0 / 9.0,
                         1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
constexpr int opp[Q] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
constexpr int mirror_y[Q] = {0, 1, 4, 3, 2, 8, 7, 6, 5};

// d2q9.h:53-58
double equilibrium(int i, double rho, double ux, double uy) {
    const double cu = static_cast<double>(cx[i]) * ux + static_cast<double>(cy[i]) * uy;
    const double uu = ux * ux + uy * uy;
    return w[i] * rho * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * uu);
}

// lbm2d_internal.h:25-27
int outlet_slot(int i) { return i == 3 ? 0 : (i == 6 ? 1 : 2); }
} // namespace d2

// lbm2d_internal.h:37-75 (resolve_source)
Source resolve2(const lbm2d_config& c, const std::uint8_t* solid, std::size_t x, std::size_t y, int i) {
    const auto nx = static_cast<std::ptrdiff_t>(c.nx);
    const auto ny = static_cast<std::ptrdiff_t>(c.ny);
    const std::size_t here = x + c.nx * y;
    std::ptrdiff_t sx = static_cast<std::ptrdiff_t>(x) - d2::cx[i];
    std::ptrdiff_t sy = static_cast<std::ptrdiff_t>(y) - d2::cy[i];
    int direction = i;
    if (sy < 0 || sy >= ny) {
        const lbm_wall& wall = sy < 0 ? c.bottom : c.top;
        switch (wall.kind) {
            case LBM_WALL_PERIODIC:
                sy = (sy + ny) % ny;
                break;
            case LBM_WALL_SLIP:
                sy = static_cast<std::ptrdiff_t>(y);
                direction = d2::mirror_y[i];
                break;
            default:
                return Source{kWall, here, d2::opp[i], wall.velocity_x};
        }
    }
    if (sx < 0 || sx >= nx) {
        if (c.x_boundary == LBM_X_PERIODIC) {
            sx = (sx + nx) % nx;
        } else if (sx < 0) {
            return Source{kInlet, here, i, 0.0};
        } else {
            const std::size_t column = c.nx - 1 + c.nx * static_cast<std::size_t>(sy);
            return Source{kOutlet, column, direction, 0.0};
        }
    }
    const std::size_t cell = static_cast<std::size_t>(sx) + c.nx * static_cast<std::size_t>(sy);
    if (solid[cell] != 0) {
        return Source{kBounce, here, d2::opp[i], 0.0};
    }
    return Source{kNode, cell, direction, 0.0};
}

// lbm2d_internal.h:85-111 (zou_he_west)
void zou_he_west(double* f, double ux, double uy) {
    const double rho = (f[0] + f[2] + f[4] + 2.0 * (f[3] + f[6] + f[7])) / (1.0 - ux);
    f[1] = f[3] + (2.0 / 3.0) * rho * ux;
    f[5] = f[7] - 0.5 * (f[2] - f[4]) + (1.0 / 6.0) * rho * ux + 0.5 * rho * uy;
    f[8] = f[6] + 0.5 * (f[2] - f[4]) + (1.0 / 6.0) * rho * ux - 0.5 * rho * uy;
    double feq[d2::Q];
    double pxx = 0.0;
    double pyy = 0.0;
    double pxy = 0.0;
    for (int i = 0; i < d2::Q; ++i) {
        feq[i] = d2::equilibrium(i, rho, ux, uy);
        const double neq = f[i] - feq[i];
        const auto ex = static_cast<double>(d2::cx[i]);
        const auto ey = static_cast<double>(d2::cy[i]);
        pxx += ex * ex * neq;
        pyy += ey * ey * neq;
        pxy += ex * ey * neq;
    }
    const double cs2 = 1.0 / 3.0;
    for (int i = 0; i < d2::Q; ++i) {
        const auto ex = static_cast<double>(d2::cx[i]);
        const auto ey = static_cast<double>(d2::cy[i]);
        f[i] = feq[i] + 4.5 * d2::w[i] * ((ex * ex - cs2) * pxx + (ey * ey - cs2) * pyy + 2.0 * ex * ey * pxy);
    }
}

// lbm2d_cpu.cpp:141-188 (collide)
void collide2(double* f, const lbm2d_config& c) {
    const double rho = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8];
    const double jx = f[1] - f[3] + f[5] - f[6] - f[7] + f[8];
    const double jy = f[2] - f[4] + f[5] + f[6] - f[7] - f[8];
    const double fx = c.body_force_x;
    const double fy = c.body_force_y;
    const double ux = (jx + 0.5 * fx) / rho;
    const double uy = (jy + 0.5 * fy) / rho;
    const double base = 1.0 - 1.5 * (ux * ux + uy * uy);
    const double cu[d2::Q] = {0.0, ux, uy, -ux, -uy, ux + uy, uy - ux, -ux - uy, ux - uy};
    double feq[d2::Q];
    for (int i = 0; i < d2::Q; ++i) {
        feq[i] = d2::w[i] * rho * (base + cu[i] * (3.0 + 4.5 * cu[i]));
    }
    double tau = c.tau;
    if (c.smagorinsky_enabled != 0) {
        double pxx = 0.0;
        double pyy = 0.0;
        double pxy = 0.0;
        for (int i = 0; i < d2::Q; ++i) {
            const double neq = f[i] - feq[i];
            const auto ex = static_cast<double>(d2::cx[i]);
            const auto ey = static_cast<double>(d2::cy[i]);
            pxx += ex * ex * neq;
            pyy += ey * ey * neq;
            pxy += ex * ey * neq;
        }
        const double strain = std::sqrt(pxx * pxx + pyy * pyy + 2.0 * pxy * pxy);
        const double cc = c.smagorinsky_constant;
        tau = 0.5 * (tau + std::sqrt(tau * tau + kSmagorinskyFactor * cc * cc * strain / rho));
    }
    const double omega = 1.0 / tau;
    const bool forced = fx != 0.0 || fy != 0.0;
    const double force_scale = 1.0 - 0.5 * omega;
    for (int i = 0; i < d2::Q; ++i) {
        double value = f[i] - omega * (f[i] - feq[i]);
        if (forced) {
            const auto ex = static_cast<double>(d2::cx[i]);
            const auto ey = static_cast<double>(d2::cy[i]);
            value += force_scale * d2::w[i] * (3.0 * ((ex - ux) * fx + (ey - uy) * fy) + 9.0 * cu[i] * (ex * fx + ey * fy));
        }
        f[i] = value;
    }
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
    if (!std::isfinite(c.tau) || !(c.tau > 0.5)) {
        return "lbm2d: tau must be finite and above 1/2 (positive viscosity)";
    }
    if (!std::isfinite(c.inlet_ux) || !std::isfinite(c.inlet_uy) || !std::isfinite(c.body_force_x) ||
        !std::isfinite(c.body_force_y) || !std::isfinite(c.bottom.velocity_x) ||
        !std::isfinite(c.top.velocity_x)) {
        return "lbm2d: velocities and body force must be finite";
    }
    if (c.smagorinsky_enabled != 0 &&
        (!std::isfinite(c.smagorinsky_constant) || c.smagorinsky_constant < 0.0)) {
        return "lbm2d: the Smagorinsky constant must be finite and non-negative";
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

struct Link2 {
    std::uint32_t cell;
    std::uint32_t direction;
};

class Oracle2D {
public:
    Oracle2D(const lbm2d_config& config, const std::uint8_t* solid)
        : c_(config), cells_(config.nx * config.ny), solid_(solid, solid + cells_) {
        // lbm2d_common.cpp:79-97 (boundary_links)
        for (std::size_t y = 0; y < c_.ny; ++y) {
            for (std::size_t x = 0; x < c_.nx; ++x) {
                const std::size_t here = x + c_.nx * y;
                if (solid_[here] != 0) {
                    continue;
                }
                for (int i = 1; i < d2::Q; ++i) {
                    if (resolve2(c_, solid_.data(), x, y, i).kind == kBounce) {
                        links_.push_back(Link2{static_cast<std::uint32_t>(here), static_cast<std::uint32_t>(i)});
                    }
                }
            }
        }
        for (int i = 0; i < d2::Q; ++i) {
            offset_[i] = -static_cast<std::ptrdiff_t>(d2::cx[i]) - static_cast<std::ptrdiff_t>(c_.nx) * d2::cy[i];
        }
        // lbm2d_cpu.cpp:316-341 (classify_nodes)
        class_.assign(cells_, 1);
        for (std::size_t y = 0; y < c_.ny; ++y) {
            for (std::size_t x = 0; x < c_.nx; ++x) {
                const std::size_t here = x + c_.nx * y;
                if (solid_[here] != 0) {
                    class_[here] = 0;
                    continue;
                }
                if (x == 0 || y == 0 || x + 1 == c_.nx || y + 1 == c_.ny) {
                    class_[here] = 2;
                    continue;
                }
                for (int i = 1; i < d2::Q; ++i) {
                    if (solid_[static_cast<std::size_t>(static_cast<std::ptrdiff_t>(here) + offset_[i])] != 0) {
                        class_[here] = 2;
                        break;
                    }
                }
            }
        }
        initialize(nullptr);
    }

    void initialize(const double* moments) {
        // lbm2d_common.cpp:178-199
        f_.assign(d2::Q * cells_, 0.0);
        const bool inflow = c_.x_boundary == LBM_X_INFLOW_OUTFLOW;
        for (std::size_t cell = 0; cell < cells_; ++cell) {
            double rho = 1.0;
            double ux = 0.0;
            double uy = 0.0;
            if (solid_[cell] == 0) {
                if (moments != nullptr) {
                    rho = moments[cell];
                    ux = moments[cells_ + cell];
                    uy = moments[2 * cells_ + cell];
                } else if (inflow) {
                    ux = c_.inlet_ux;
                    uy = c_.inlet_uy;
                }
            }
            for (int i = 0; i < d2::Q; ++i) {
                f_[i * cells_ + cell] = d2::equilibrium(i, rho, ux, uy);
            }
        }
        g_ = f_;
        // lbm2d_cpu.cpp:289-297 (reset_outlet)
        outlet_.assign(3 * c_.ny, 0.0);
        for (std::size_t y = 0; y < c_.ny; ++y) {
            const std::size_t cell = c_.nx - 1 + c_.nx * y;
            for (const int i : {3, 6, 7}) {
                outlet_[static_cast<std::size_t>(d2::outlet_slot(i)) * c_.ny + y] = f_[i * cells_ + cell];
            }
        }
        steps_ = 0;
    }

    void step(std::uint64_t count) {
        for (std::uint64_t n = 0; n < count; ++n) {
#pragma omp parallel for schedule(static)
            for (std::ptrdiff_t y = 0; y < static_cast<std::ptrdiff_t>(c_.ny); ++y) {
                update_row(static_cast<std::size_t>(y));
            }
            std::swap(f_, g_);
            ++steps_;
        }
    }

    // lbm2d_cpu.cpp:232-259
    void macroscopic(double* rho, double* ux, double* uy) const {
        for (std::size_t cell = 0; cell < cells_; ++cell) {
            double r = 1.0;
            double vx = 0.0;
            double vy = 0.0;
            if (solid_[cell] == 0) {
                double s = 0.0;
                double jx = 0.0;
                double jy = 0.0;
                for (int i = 0; i < d2::Q; ++i) {
                    const double v = f_[i * cells_ + cell];
                    s += v;
                    jx += static_cast<double>(d2::cx[i]) * v;
                    jy += static_cast<double>(d2::cy[i]) * v;
                }
                r = s;
                vx = (jx - 0.5 * c_.body_force_x) / s;
                vy = (jy - 0.5 * c_.body_force_y) / s;
            }
            if (rho != nullptr) rho[cell] = r;
            if (ux != nullptr) ux[cell] = vx;
            if (uy != nullptr) uy[cell] = vy;
        }
    }

    // lbm2d_cpu.cpp:265-272 with add_link_force (lbm2d.h:145-154)
    void force(double* fx, double* fy, double* moment) const {
        double tx = 0.0;
        double ty = 0.0;
        double tm = 0.0;
        for (const Link2& link : links_) {
            const int out = d2::opp[link.direction];
            const double outgoing = f_[out * cells_ + link.cell];
            const double vx = 2.0 * static_cast<double>(d2::cx[out]) * outgoing;
            const double vy = 2.0 * static_cast<double>(d2::cy[out]) * outgoing;
            const std::size_t x = link.cell % c_.nx;
            const std::size_t y = link.cell / c_.nx;
            const double rx = static_cast<double>(x) - 0.5 * d2::cx[link.direction];
            const double ry = static_cast<double>(y) - 0.5 * d2::cy[link.direction];
            tx += vx;
            ty += vy;
            tm += rx * vy - ry * vx;
        }
        *fx = tx;
        *fy = ty;
        *moment = tm;
    }

    double mass() const {
        double total = 0.0;
        for (std::size_t cell = 0; cell < cells_; ++cell) {
            if (solid_[cell] == 0) {
                for (int i = 0; i < d2::Q; ++i) {
                    total += f_[i * cells_ + cell];
                }
            }
        }
        return total;
    }

    const std::vector<double>& populations() const { return f_; }
    const std::vector<Link2>& links() const { return links_; }
    std::uint64_t steps() const { return steps_; }
    const lbm2d_config& config() const { return c_; }

private:
    double density_at(std::size_t cell) const {
        double rho = 0.0;
        for (int i = 0; i < d2::Q; ++i) {
            rho += f_[i * cells_ + cell];
        }
        return rho;
    }

    // lbm2d_cpu.cpp:302-310
    double outlet_value(double upstream, int i, std::size_t y) {
        if (c_.outlet == LBM_OUTLET_EXTRAPOLATION) {
            return upstream;
        }
        const double u = c_.inlet_ux;
        double& previous = outlet_[static_cast<std::size_t>(d2::outlet_slot(i)) * c_.ny + y];
        previous = (previous + u * upstream) / (1.0 + u);
        return previous;
    }

    // lbm2d_cpu.cpp:352-404 (update_row)
    void update_row(std::size_t y) {
        const bool inlet = c_.x_boundary == LBM_X_INFLOW_OUTFLOW;
        for (std::size_t x = 0; x < c_.nx; ++x) {
            const std::size_t here = x + c_.nx * y;
            const std::uint8_t node = class_[here];
            if (node == 0) {
                continue;
            }
            double f[d2::Q];
            if (node == 1) {
                for (int i = 0; i < d2::Q; ++i) {
                    f[i] = f_[i * cells_ + static_cast<std::size_t>(static_cast<std::ptrdiff_t>(here) + offset_[i])];
                }
            } else {
                double wall_rho = -1.0;
                for (int i = 0; i < d2::Q; ++i) {
                    const Source s = resolve2(c_, solid_.data(), x, y, i);
                    switch (s.kind) {
                        case kNode:
                        case kBounce:
                            f[i] = f_[s.direction * cells_ + s.cell];
                            break;
                        case kWall:
                            if (wall_rho < 0.0) {
                                wall_rho = density_at(here);
                            }
                            f[i] = f_[s.direction * cells_ + s.cell] +
                                   6.0 * d2::w[i] * wall_rho * static_cast<double>(d2::cx[i]) * s.wall_ux;
                            break;
                        case kInlet:
                            f[i] = 0.0;
                            break;
                        default:
                            f[i] = outlet_value(f_[s.direction * cells_ + s.cell], i, y);
                            break;
                    }
                }
                if (inlet && x == 0) {
                    zou_he_west(f, c_.inlet_ux, c_.inlet_uy);
                }
            }
            collide2(f, c_);
            for (int i = 0; i < d2::Q; ++i) {
                g_[i * cells_ + here] = f[i];
            }
        }
    }

    lbm2d_config c_;
    std::size_t cells_;
    std::vector<std::uint8_t> solid_;
    std::vector<Link2> links_;
    std::array<std::ptrdiff_t, d2::Q> offset_{};
    std::vector<std::uint8_t> class_;
    std::vector<double> f_;
    std::vector<double> g_;
    std::vector<double> outlet_;
    std::uint64_t steps_ = 0;
};

// SplitMix64 finaliser (Steele, Lea and Flood 2014) keyed by (seed, cell, component).
double perturbation(std::uint64_t seed, std::uint64_t cell, std::uint32_t component) {
    std::uint64_t z = seed + 0x9E3779B97F4A7C15ULL * (cell * 4U + component + 1U);
    z = (z ^ (z >> 30U)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27U)) * 0x94D049BB133111EBULL;
    z ^= z >> 31U;
    const double u = static_cast<double>(z >> 11U) * (1.0 / 9007199254740992.0);
    return 2.0 * u - 1.0;
}

} // namespace

// ------------------------------------------------------------------------------------------
// C ABI of the oracle (ctypes from the tests).
extern "C" {

int orc_set_threads(int n) {
    if (n > 0) {
        omp_set_num_threads(n);
    }
    return omp_get_max_threads();
}

double orc_perturbation(std::uint64_t seed, std::uint64_t cell, std::uint32_t component) {
    return perturbation(seed, cell, component);
}

int orc3d_validate(const lbm3d_config* c, const std::uint8_t* solid, const double* distance, char* err,
                   std::size_t len) {
    const std::string p = problem3(*c, solid, distance);
    put(err, len, p);
    return p.empty() ? 0 : 1;
}

void* orc3d_create(const lbm3d_config* c, const std::uint8_t* solid, const double* distance, char* err,
                   std::size_t len) {
    const std::string p = problem3(*c, solid, distance);
    if (!p.empty()) {
        put(err, len, p);
        return nullptr;
    }
    return new Oracle3D(*c, solid, distance);
}

void orc3d_destroy(void* h) { delete static_cast<Oracle3D*>(h); }

void orc3d_initialize(void* h, const double* moments) { static_cast<Oracle3D*>(h)->initialize(moments); }

void orc3d_initialize_perturbed(void* h, std::uint64_t seed, double amplitude) {
    auto* o = static_cast<Oracle3D*>(h);
    const lbm3d_config& c = o->config();
    const std::size_t cells = c.nx * c.ny * c.nz;
    const bool inflow = c.x_boundary == LBM_X_INFLOW_OUTFLOW;
    std::vector<double> m(4 * cells);
#pragma omp parallel for schedule(static)
    for (std::size_t cell = 0; cell < cells; ++cell) {
        m[cell] = 1.0;
        m[cells + cell] = (inflow ? c.inlet_ux : 0.0) + amplitude * perturbation(seed, cell, 0);
        m[2 * cells + cell] = (inflow ? c.inlet_uy : 0.0) + amplitude * perturbation(seed, cell, 1);
        m[3 * cells + cell] = (inflow ? c.inlet_uz : 0.0) + amplitude * perturbation(seed, cell, 2);
    }
    o->initialize(m.data());
}

void orc3d_step(void* h, std::uint64_t count) { static_cast<Oracle3D*>(h)->step(count); }
std::uint64_t orc3d_steps(void* h) { return static_cast<Oracle3D*>(h)->steps(); }
void orc3d_set_inlet(void* h, const double* v) { static_cast<Oracle3D*>(h)->set_inlet(v); }
void orc3d_force(void* h, double* F, double* M) { static_cast<Oracle3D*>(h)->force(F, M); }
double orc3d_mass(void* h) { return static_cast<Oracle3D*>(h)->mass(); }
void orc3d_macroscopic(void* h, double* rho, double* ux, double* uy, double* uz) {
    static_cast<Oracle3D*>(h)->macroscopic(rho, ux, uy, uz);
}
void orc3d_populations(void* h, double* f) {
    const std::vector<double>& v = static_cast<Oracle3D*>(h)->populations();
    std::memcpy(f, v.data(), v.size() * sizeof(double));
}
void orc3d_set_populations(void* h, const double* f) { static_cast<Oracle3D*>(h)->set_populations(f); }
std::size_t orc3d_outlet_history(void* h, double* out) {
    const std::vector<double>& v = static_cast<Oracle3D*>(h)->outlet_history();
    if (out != nullptr) {
        std::memcpy(out, v.data(), v.size() * sizeof(double));
    }
    return v.size();
}
std::size_t orc3d_links(void* h, std::uint32_t* cell, std::uint32_t* dir, double* q, std::size_t cap) {
    const std::vector<Link3>& links = static_cast<Oracle3D*>(h)->links();
    for (std::size_t k = 0; k < links.size() && k < cap; ++k) {
        if (cell != nullptr) cell[k] = links[k].cell;
        if (dir != nullptr) dir[k] = links[k].direction;
        if (q != nullptr) q[k] = links[k].q;
    }
    return links.size();
}
void orc3d_classes(void* h, std::uint8_t* out) {
    const std::vector<std::uint8_t>& v = static_cast<Oracle3D*>(h)->classes();
    std::memcpy(out, v.data(), v.size());
}

// Building blocks, exposed so that tests compare them with oracle/_ref one by one.
int orc3d_resolve(const lbm3d_config* c, const std::uint8_t* solid, std::uint64_t x, std::uint64_t y,
                  std::uint64_t z, int i, std::uint64_t* cell, int* direction, double* wall_ux) {
    const Source s = resolve3(*c, solid, x, y, z, i);
    *cell = s.cell;
    *direction = s.direction;
    *wall_ux = s.wall_ux;
    return s.kind;
}
double orc3d_link_fraction(const lbm3d_config* c, const std::uint8_t* solid, const double* distance,
                           std::uint64_t x, std::uint64_t y, std::uint64_t z, int i) {
    return link_fraction3(*c, solid, distance, x, y, z, i);
}
double orc3d_sponge(const lbm3d_config* c, std::uint64_t x, std::uint64_t y, std::uint64_t z) {
    return sponge_sigma3(*c, x, y, z);
}
void orc3d_inlet_west(double* f, double ux, double uy, double uz) { inlet_west3(f, ux, uy, uz); }
void orc3d_collide(double* f, const lbm3d_config* c, double sponge, const double* stream) {
    collide3(f, *c, sponge, stream);
}
double orc3d_equilibrium(int i, double rho, double ux, double uy, double uz) {
    return d3::equilibrium(i, rho, ux, uy, uz);
}

int orc2d_validate(const lbm2d_config* c, const std::uint8_t* solid, char* err, std::size_t len) {
    const std::string p = problem2(*c, solid);
    put(err, len, p);
    return p.empty() ? 0 : 1;
}
void* orc2d_create(const lbm2d_config* c, const std::uint8_t* solid, char* err, std::size_t len) {
    const std::string p = problem2(*c, solid);
    if (!p.empty()) {
        put(err, len, p);
        return nullptr;
    }
    return new Oracle2D(*c, solid);
}
void orc2d_destroy(void* h) { delete static_cast<Oracle2D*>(h); }
void orc2d_initialize(void* h, const double* moments) { static_cast<Oracle2D*>(h)->initialize(moments); }
void orc2d_initialize_perturbed(void* h, std::uint64_t seed, double amplitude) {
    auto* o = static_cast<Oracle2D*>(h);
    const lbm2d_config& c = o->config();
    const std::size_t cells = c.nx * c.ny;
    const bool inflow = c.x_boundary == LBM_X_INFLOW_OUTFLOW;
    std::vector<double> m(3 * cells);
    for (std::size_t cell = 0; cell < cells; ++cell) {
        m[cell] = 1.0;
        m[cells + cell] = (inflow ? c.inlet_ux : 0.0) + amplitude * perturbation(seed, cell, 0);
        m[2 * cells + cell] = (inflow ? c.inlet_uy : 0.0) + amplitude * perturbation(seed, cell, 1);
    }
    o->initialize(m.data());
}
void orc2d_step(void* h, std::uint64_t count) { static_cast<Oracle2D*>(h)->step(count); }
std::uint64_t orc2d_steps(void* h) { return static_cast<Oracle2D*>(h)->steps(); }
void orc2d_force(void* h, double* fx, double* fy, double* m) { static_cast<Oracle2D*>(h)->force(fx, fy, m); }
double orc2d_mass(void* h) { return static_cast<Oracle2D*>(h)->mass(); }
void orc2d_macroscopic(void* h, double* rho, double* ux, double* uy) {
    static_cast<Oracle2D*>(h)->macroscopic(rho, ux, uy);
}
void orc2d_populations(void* h, double* f) {
    const std::vector<double>& v = static_cast<Oracle2D*>(h)->populations();
    std::memcpy(f, v.data(), v.size() * sizeof(double));
}
std::size_t orc2d_links(void* h, std::uint32_t* cell, std::uint32_t* dir, std::size_t cap) {
    const std::vector<Link2>& links = static_cast<Oracle2D*>(h)->links();
    for (std::size_t k = 0; k < links.size() && k < cap; ++k) {
        if (cell != nullptr) cell[k] = links[k].cell;
        if (dir != nullptr) dir[k] = links[k].direction;
    }
    return links.size();
}
int orc2d_resolve(const lbm2d_config* c, const std::uint8_t* solid, std::uint64_t x, std::uint64_t y, int i,
                  std::uint64_t* cell, int* direction, double* wall_ux) {
    const Source s = resolve2(*c, solid, x, y, i);
    *cell = s.cell;
    *direction = s.direction;
    *wall_ux = s.wall_ux;
    return s.kind;
}
void orc2d_zou_he_west(double* f, double ux, double uy) { zou_he_west(f, ux, uy); }

} // extern "C"
