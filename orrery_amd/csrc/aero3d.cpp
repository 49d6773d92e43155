// The D3Q19 lattice behind Orrery's aerodynamic solver interface: LbmAeroSolver3D
// (src/aero/lbm3d/lbm_aero_solver3d.cpp:85-590, lbm_aero_solver3d.h:45-166) restated over this
// library's C ABI. A closed body mesh in SI units goes in; coefficients, SI force and moment,
// the force history, wall pressure samples and validity warnings come out.
//
// The setup (wind frame, domain, cell budget, body placement, voxelisation, link fractions,
// step counts, lattice configuration and warnings) follows the reference operation for
// operation, so the lattice and mask are the reference's bit for bit. Per call, the reference
// reads the whole macroscopic field back to the host; here the wall-pressure samples, the
// inlet-plane density and the largest speed are computed on the device and only those values
// are read back (lbm3d_sample, lbm3d_max_speed). The velocity field is produced on request.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <new>
#include <numbers>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/orrery_lbm.h"

#include "aero_common.h"

namespace {

using namespace lbm::aero;

constexpr double kMinCellsPerLength = 16.0;
constexpr double kMinBodyCells = 4.0;
constexpr double kBlockageWarning = 0.05;
constexpr double kMaxExtent = 1e7;
constexpr double kPlacementShift = 0.25;
constexpr std::size_t kDefaultMaxCells = std::size_t{1} << 24U;

struct Force {
    double f[3] = {0, 0, 0};
    double m[3] = {0, 0, 0};
};

const int kCx[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
const int kCy[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
const int kCz[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};

} // namespace

struct lbm_aero3d {
    lbm_aero3d_options options{};
    lbm_flow_conditions conditions{};
    lbm_reference_values reference{};
    Units units{};
    bool planned = false;
    lbm3d_config config{};
    std::vector<std::uint8_t> solid;
    std::vector<double> distance;
    lbm3d* solver = nullptr;
    std::vector<Warning> setup_warnings;
    std::vector<Warning> warnings;
    V3 axes[3];
    double centre[3] = {0, 0, 0};
    double origin[3] = {0, 0, 0};
    std::uint64_t first[3] = {0, 0, 0};
    std::uint64_t count[3] = {0, 0, 0};
    double surface_area = 0.0;
    double blockage = 0.0;
    double thinnest = 0.0;
    std::uint64_t total_steps = 0;
    double ramp_steps = 0.0;
    std::size_t window_samples = 1;
    double previous_cd = 0.0;
    // history (CoefficientHistory + side force) and the lattice force samples
    std::vector<double> time, drag, lift, side;
    std::vector<Force> samples;
    double instantaneous[3] = {0, 0, 0};
    // result
    lbm_aero3d_result result{};
    std::vector<double> residuals;
    std::vector<std::uint32_t> wall_cells;
    std::vector<double> wall_normals; // lattice frame, unit, [3 n]
    std::vector<double> points, normals, areas, cp;
};

namespace {

V3 to_body(const lbm_aero3d& a, const double v[3]) { return a.axes[0] * v[0] + a.axes[1] * v[1] + a.axes[2] * v[2]; }

void release_solver(lbm_aero3d& a) {
    if (a.solver != nullptr) {
        lbm3d_destroy(a.solver);
        a.solver = nullptr;
    }
}

// apply_force (lbm_aero_solver3d.cpp:413-426) and set_coefficients (src/aero/aero_result.cpp:22-35).
void apply_force(const lbm_aero3d& a, const Force& lattice, lbm_aero3d_result& r) {
    const double dx = a.units.dx;
    const double fu = a.units.force_unit();
    const double f[3] = {lattice.f[0] * fu, lattice.f[1] * fu, lattice.f[2] * fu};
    const double m[3] = {lattice.m[0] * fu * dx, lattice.m[1] * fu * dx, lattice.m[2] * fu * dx};
    set_coefficients(to_body(a, f), to_body(a, m), a.conditions, a.reference, r);
}

// forces_settled_3d (lbm_aero_solver3d.cpp:60-83), with the reference's SettledCriteria.
bool forces_settled(const lbm_aero3d& a, double from) {
    return lbm_forces_settled(a.time.data(), a.drag.data(), a.lift.data(), a.side.data(), a.time.size(), from,
                              kSteadyResidual, kMaxPeriodSpread, kStationarity) == 1;
}

// The setup of lbm_aero_solver3d.cpp:85-339 up to (not including) the creation of the lattice.
int plan(lbm_aero3d& a, const lbm_mesh* mesh, const lbm_reference_values& reference,
         const lbm_flow_conditions& conditions) {
    release_solver(a);
    a.planned = false;
    a.result = lbm_aero3d_result{};
    a.setup_warnings.clear();
    a.warnings.clear();
    a.residuals.clear();
    const lbm_aero3d_options& o = a.options;
    // validate_mesh (src/geom/mesh_validation.cpp:58-80) with MeshLimits{}.
    std::size_t nv = 0;
    std::size_t nt = 0;
    if (mesh == nullptr || lbm_mesh_size(mesh, &nv, &nt) != LBM_OK) {
        return aero_fail(LBM_EINVAL, "lbm3d: invalid mesh: mesh has no vertices or no triangles");
    }
    std::vector<double> vtx(3 * nv);
    std::vector<std::uint32_t> tri(3 * nt);
    lbm_mesh_data(mesh, vtx.data(), tri.data());
    if (nv == 0 || nt == 0) {
        return aero_fail(LBM_EINVAL, "lbm3d: invalid mesh: mesh has no vertices or no triangles");
    }
    if (nv > 10000000U || nt > 10000000U) {
        return aero_fail(LBM_EINVAL, "lbm3d: invalid mesh: the vertex or triangle count exceeds the limit");
    }
    for (std::size_t k = 0; k < 3 * nv; ++k) {
        if (!std::isfinite(vtx[k])) {
            return aero_fail(LBM_EINVAL, "lbm3d: invalid mesh: vertex " + std::to_string(k / 3) +
                                             " has a NaN or infinite coordinate");
        }
        if (std::abs(vtx[k]) > 1e12) {
            return aero_fail(LBM_EINVAL, "lbm3d: invalid mesh: vertex " + std::to_string(k / 3) +
                                             " coordinate exceeds the limit");
        }
    }
    if (!positive_finite(conditions.speed) || !positive_finite(conditions.density) ||
        !positive_finite(conditions.viscosity) || !std::isfinite(conditions.angle_of_attack) ||
        !std::isfinite(conditions.sideslip)) {
        return aero_fail(LBM_EINVAL, "lbm3d: speed, density and viscosity must be positive and finite, angles finite");
    }
    if (!positive_finite(reference.length) || !positive_finite(reference.area) ||
        !std::isfinite(reference.moment_centre[0]) || !std::isfinite(reference.moment_centre[1]) ||
        !std::isfinite(reference.moment_centre[2])) {
        return aero_fail(LBM_EINVAL,
                         "lbm3d: reference length and area must be positive and finite, the moment centre finite");
    }
    if (!positive_finite(o.upstream) || !positive_finite(o.downstream) || !positive_finite(o.width) ||
        !positive_finite(o.duration) || !positive_finite(o.average_window) || o.steps_per_call == 0 ||
        o.sample_every == 0) {
        return aero_fail(LBM_EINVAL, "lbm3d: domain, duration and step options must be positive");
    }
    if (!std::isfinite(o.sponge_strength) || o.sponge_strength < 0.0 || o.sponge_strength > 1.0 ||
        !(o.sponge_inlet >= 0.0) || !(o.sponge_outlet >= 0.0) || !(o.sponge_sides >= 0.0) || !(o.start_ramp >= 0.0) ||
        !std::isfinite(o.start_ramp) || !std::isfinite(o.sponge_inlet + o.sponge_outlet + o.sponge_sides)) {
        return aero_fail(LBM_EINVAL, "lbm3d: the absorbing-layer strength must be in [0, 1], the thicknesses finite "
                                     "and non-negative");
    }
    if (o.sides == LBM_WALL_NO_SLIP || (o.sides != LBM_WALL_SLIP && o.sides != LBM_WALL_PERIODIC)) {
        return aero_fail(LBM_EINVAL, "lbm3d: the channel sides are free slip or periodic");
    }
    if (o.max_cells > kMaxGridCells) {
        return aero_fail(LBM_EINVAL, "lbm3d: max_cells above geom::max_grid_cells");
    }
    const double length = reference.length;
    Units units;
    if (int rc = make_units(length, conditions.speed, conditions.viscosity / conditions.density, conditions.density,
                            o.cells_per_length, o.lattice_velocity, units);
        rc != LBM_OK) {
        return rc;
    }
    const double dx = units.dx;

    // Wind frame: lattice x along the free stream, y along the side force, z along the lift.
    a.axes[0] = drag_direction(conditions);
    a.axes[1] = side_direction(conditions);
    a.axes[2] = lift_direction(conditions);
    std::vector<double> wind(3 * nv);
    double lo[3] = {0, 0, 0};
    double hi[3] = {0, 0, 0};
    for (std::size_t k = 0; k < nv; ++k) {
        const V3 p{vtx[3 * k], vtx[3 * k + 1], vtx[3 * k + 2]};
        const double w[3] = {dot(p, a.axes[0]), dot(p, a.axes[1]), dot(p, a.axes[2])};
        for (int c = 0; c < 3; ++c) {
            lo[c] = k == 0 ? w[c] : std::min(lo[c], w[c]);
            hi[c] = k == 0 ? w[c] : std::max(hi[c], w[c]);
            wind[3 * k + c] = w[c];
        }
    }
    const double extent[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
    const double span_real[3] = {(o.upstream * length + extent[0] + o.downstream * length) / dx,
                                 std::max(o.width * length, extent[1] + 4.0 * dx) / dx,
                                 std::max(o.width * length, extent[2] + 4.0 * dx) / dx};
    for (const double s : span_real) {
        if (!(s < kMaxExtent)) {
            return aero_fail(LBM_EINVAL, "lbm3d: the domain is too large for the resolution");
        }
    }
    const auto nx = static_cast<std::size_t>(std::ceil(span_real[0])) + 1;
    const auto ny = static_cast<std::size_t>(std::ceil(span_real[1])) + 1;
    const auto nz = static_cast<std::size_t>(std::ceil(span_real[2])) + 1;
    if (nx < 3 || ny < 3 || nz < 3) {
        return aero_fail(LBM_EINVAL, "lbm3d: the lattice " + std::to_string(nx) + " x " + std::to_string(ny) + " x " +
                                         std::to_string(nz) +
                                         " needs at least 3 nodes along each axis; raise the resolution or the domain");
    }

    // Cell budget: the caller's, else the device memory, else the conservative cap.
    std::size_t budget = o.max_cells != 0 ? o.max_cells : kDefaultMaxCells;
    std::string reason = o.max_cells != 0 ? "LbmSolver3DOptions::max_cells"
                                          : "the default cap of 256^3 nodes (device memory unknown)";
    if (o.max_cells == 0) {
        std::size_t available = 0;
        if (lbm_device_memory(&available, nullptr) == LBM_OK) {
            // Bytes per node of this library's lattice: the populations once (in place), the
            // node class, and the four moment planes a full readback stages.
            const std::size_t per_cell = 19U * (o.storage == LBM_STORE_F64 ? 8U : o.storage == LBM_STORE_F32 ? 4U : 2U) +
                                         1U + 4U * 8U;
            budget = std::min(kMaxGridCells, available / per_cell);
            reason = "the free device memory (" + format(static_cast<double>(available) / 1048576.0) + " MiB)";
        }
    }
    const double needed = static_cast<double>(nx) * static_cast<double>(ny) * static_cast<double>(nz);
    const bool fits = nx <= kMaxGridCells && ny <= kMaxGridCells / nx && nz <= kMaxGridCells / (nx * ny);
    if (!fits || nx * ny * nz > budget) {
        const double fitting = std::floor(10.0 * units.cells * std::cbrt(static_cast<double>(budget) / needed)) / 10.0;
        return aero_fail(LBM_EINVAL, "lbm3d: the lattice " + std::to_string(nx) + " x " + std::to_string(ny) + " x " +
                                         std::to_string(nz) + " (" + format(needed) + " nodes) exceeds " + reason +
                                         ", a budget of " + std::to_string(budget) + " nodes; use at most " +
                                         format(fitting) + " cells per reference length or a smaller domain");
    }

    // Lattice node (i, j, k) sits at origin + dx (i, j, k) in the wind frame, the body a
    // quarter node off the lattice planes.
    a.origin[0] = lo[0] - o.upstream * length - kPlacementShift * dx;
    a.origin[1] = 0.5 * (lo[1] + hi[1]) - (0.5 * static_cast<double>(ny - 1) + kPlacementShift) * dx;
    a.origin[2] = 0.5 * (lo[2] + hi[2]) - (0.5 * static_cast<double>(nz - 1) + kPlacementShift) * dx;
    const std::size_t dims[3] = {nx, ny, nz};
    for (int c = 0; c < 3; ++c) {
        const double from = std::floor((lo[c] - a.origin[c]) / dx) - 1.0;
        const double to = std::ceil((hi[c] - a.origin[c]) / dx) + 1.0;
        const auto last = static_cast<double>(dims[c] - 2);
        const double begin = std::clamp(from, 1.0, last);
        const double end = std::clamp(to, begin, last);
        a.first[c] = static_cast<std::uint64_t>(begin);
        a.count[c] = static_cast<std::uint64_t>(end) - a.first[c] + 1;
    }
    const double spec_origin[3] = {a.origin[0] + (static_cast<double>(a.first[0]) - 0.5) * dx,
                                   a.origin[1] + (static_cast<double>(a.first[1]) - 0.5) * dx,
                                   a.origin[2] + (static_cast<double>(a.first[2]) - 0.5) * dx};
    lbm_mesh* wm = nullptr;
    if (int rc = lbm_mesh_create(wind.data(), nv, tri.data(), nt, &wm); rc != LBM_OK) {
        return aero_fail(rc, std::string("lbm3d: ") + lbm_geom_last_error());
    }
    const std::size_t box = a.count[0] * a.count[1] * a.count[2];
    std::vector<std::uint8_t> voxels(box);
    int rc = lbm_voxelize(wm, spec_origin, dx, a.count[0], a.count[1], a.count[2], 0.5, voxels.data(), nullptr);
    std::vector<double> sdf;
    if (rc == LBM_OK && o.wall == LBM_SOLID_INTERPOLATED) {
        sdf.resize(box);
        rc = lbm_signed_distance_band(wm, spec_origin, dx, a.count[0], a.count[1], a.count[2], voxels.data(), 3.0,
                                      sdf.data(), nullptr);
    }
    lbm_mesh_destroy(wm);
    if (rc != LBM_OK) {
        return aero_fail(rc, std::string("lbm3d: ") + lbm_geom_last_error());
    }
    const std::size_t cells = nx * ny * nz;
    a.solid.assign(cells, 0);
    a.distance.clear();
    // place_voxels / place_distance (src/aero/lbm3d/solid_mask3d.cpp:130-163).
    std::size_t solid_count = 0;
    if (!sdf.empty()) {
        a.distance.assign(cells, 1e9);
    }
    for (std::size_t k = 0; k < a.count[2]; ++k) {
        for (std::size_t j = 0; j < a.count[1]; ++j) {
            for (std::size_t i = 0; i < a.count[0]; ++i) {
                const std::size_t v = i + a.count[0] * (j + a.count[1] * k);
                const std::size_t c = (i + a.first[0]) + nx * ((j + a.first[1]) + ny * (k + a.first[2]));
                if (voxels[v] != 0) {
                    a.solid[c] = 1;
                }
                if (!sdf.empty()) {
                    a.distance[c] = sdf[v] / dx;
                }
            }
        }
    }
    for (std::uint8_t s : a.solid) {
        solid_count += s;
    }
    if (solid_count == 0) {
        return aero_fail(LBM_EINVAL, "lbm3d: the mesh covers no lattice node (resolution too low)");
    }

    // Step counts, checked before they are converted to integers.
    const double cells_per_length = units.cells;
    const double steps_per_time = cells_per_length / units.u_lat;
    const double total_real = std::ceil(o.duration * steps_per_time);
    const double window_real = o.average_window * steps_per_time / static_cast<double>(o.sample_every);
    const double ramp_real = o.start_ramp * steps_per_time;
    if (!(total_real < kMaxSteps) || !(window_real < kMaxSteps) || !(ramp_real < kMaxSteps)) {
        return aero_fail(LBM_EINVAL, "lbm3d: duration and averaging window must stay below " + format(kMaxSteps) +
                                         " lattice steps");
    }

    lbm3d_config& config = a.config;
    config = lbm3d_config{};
    config.nx = nx;
    config.ny = ny;
    config.nz = nz;
    config.tau = units.tau();
    config.x_boundary = LBM_X_INFLOW_OUTFLOW;
    config.outlet = LBM_OUTLET_CONVECTIVE;
    config.y_low = lbm_wall{o.sides, 0.0};
    config.y_high = lbm_wall{o.sides, 0.0};
    config.z_low = lbm_wall{o.sides, 0.0};
    config.z_high = lbm_wall{o.sides, 0.0};
    config.inlet_ux = units.u_lat;
    const double reynolds = units.reynolds;
    const bool les = o.les == LBM_LES_ON || (o.les == LBM_LES_AUTOMATIC && reynolds > kLesReynolds);
    config.smagorinsky_enabled = les ? 1 : 0;
    config.smagorinsky_constant = o.smagorinsky_constant;
    config.solid_wall = o.wall;
    const auto layer = [&](double lengths, std::size_t axis) {
        const double nodes = std::round(lengths * cells_per_length);
        return static_cast<std::uint64_t>(std::min(nodes, std::floor(0.25 * static_cast<double>(axis))));
    };
    config.sponge_strength = o.sponge_strength;
    config.sponge_inlet = layer(o.sponge_inlet, nx);
    config.sponge_outlet = layer(o.sponge_outlet, nx);
    config.sponge_sides = layer(o.sponge_sides, std::min(ny, nz));

    // Frontal area of the voxelised body: (y, z) columns with a solid node.
    std::size_t frontal = 0;
    for (std::size_t z = 0; z < nz; ++z) {
        for (std::size_t y = 0; y < ny; ++y) {
            for (std::size_t x = a.first[0]; x < a.first[0] + a.count[0]; ++x) {
                if (a.solid[x + nx * (y + ny * z)] != 0) {
                    ++frontal;
                    break;
                }
            }
        }
    }
    a.blockage = static_cast<double>(frontal) / (static_cast<double>(ny) * static_cast<double>(nz));
    a.units = units;
    a.ramp_steps = ramp_real;
    const V3 mc{reference.moment_centre[0], reference.moment_centre[1], reference.moment_centre[2]};
    for (int c = 0; c < 3; ++c) {
        a.centre[c] = (dot(mc, a.axes[c]) - a.origin[c]) / dx;
    }
    // surface_area (src/geom/triangle_mesh.cpp:32-38) of the body-frame mesh.
    double area = 0.0;
    for (std::size_t t = 0; t < nt; ++t) {
        const V3 p0{vtx[3 * tri[3 * t]], vtx[3 * tri[3 * t] + 1], vtx[3 * tri[3 * t] + 2]};
        const V3 p1{vtx[3 * tri[3 * t + 1]], vtx[3 * tri[3 * t + 1] + 1], vtx[3 * tri[3 * t + 1] + 2]};
        const V3 p2{vtx[3 * tri[3 * t + 2]], vtx[3 * tri[3 * t + 2] + 1], vtx[3 * tri[3 * t + 2] + 2]};
        const V3 e1{p1.x - p0.x, p1.y - p0.y, p1.z - p0.z};
        const V3 e2{p2.x - p0.x, p2.y - p0.y, p2.z - p0.z};
        const V3 cr = cross(e1, e2);
        area += 0.5 * std::hypot(cr.x, cr.y, cr.z);
    }
    a.surface_area = area;
    a.conditions = conditions;
    a.reference = reference;
    a.total_steps = static_cast<std::uint64_t>(total_real);
    a.window_samples = std::max<std::size_t>(1, static_cast<std::size_t>(window_real));
    a.time.clear();
    a.drag.clear();
    a.lift.clear();
    a.side.clear();
    a.samples.clear();
    a.instantaneous[0] = a.instantaneous[1] = a.instantaneous[2] = 0.0;
    a.previous_cd = 0.0;

    // Warnings: the method's envelope plus the lattice's own conditions.
    a.setup_warnings = check_validity(conditions, length);
    a.thinnest = std::min({extent[0], extent[1], extent[2]}) / dx;
    if (cells_per_length < kMinCellsPerLength || a.thinnest < kMinBodyCells) {
        a.setup_warnings.push_back(
            {"lbm.resolution", format(cells_per_length) + " cells per reference length and " + format(a.thinnest) +
                                   " cells across the thinnest body extent; below " + format(kMinCellsPerLength) +
                                   " and " + format(kMinBodyCells) + " the staircase wall dominates the force error (A4)"});
    }
    if (config.tau < kMarginalTau) {
        a.setup_warnings.push_back({"lbm.tau", "relaxation time " + format(config.tau) + " is within " +
                                                   format(kMarginalTau - 0.5) +
                                                   " of 1/2: BGK is close to instability; raise the resolution or "
                                                   "lower the Reynolds number"});
    }
    if (les) {
        a.setup_warnings.push_back({"lbm.les", "Smagorinsky subgrid closure active at Re " + format(reynolds) +
                                                   " (APPROXIMATION A2: eddy viscosity, not resolved turbulence)"});
    } else if (reynolds > kLesReynolds) {
        a.setup_warnings.push_back({"lbm.les_off", "Re " + format(reynolds) + " above " + format(kLesReynolds) +
                                                       " without the Smagorinsky closure: the flow is under-resolved"});
    }
    if (a.blockage > kBlockageWarning) {
        a.setup_warnings.push_back({"lbm.blockage", "the body blocks " + format(100.0 * a.blockage) +
                                                        " percent of the channel cross-section (above " +
                                                        format(100.0 * kBlockageWarning) +
                                                        "); confinement raises the drag"});
    }
    a.warnings = a.setup_warnings;
    a.planned = true;
    return LBM_OK;
}

// The wall samples: one per fluid node with a link to the body, in cell order, with the
// normalised sum of its link directions (lbm_aero_solver3d.cpp:497-530).
int wall_samples(lbm_aero3d& a) {
    std::size_t n = 0;
    if (int rc = lbm3d_links(a.solver, nullptr, nullptr, nullptr, 0, &n); rc != LBM_OK) {
        return rc;
    }
    std::vector<std::uint32_t> cell(n);
    std::vector<std::uint32_t> dir(n);
    if (int rc = lbm3d_links(a.solver, cell.data(), dir.data(), nullptr, n, &n); rc != LBM_OK) {
        return rc;
    }
    std::map<std::uint32_t, std::array<double, 3>> wall;
    for (std::size_t k = 0; k < n; ++k) {
        std::array<double, 3>& s = wall[cell[k]];
        s[0] += kCx[dir[k]];
        s[1] += kCy[dir[k]];
        s[2] += kCz[dir[k]];
    }
    a.wall_cells.clear();
    a.wall_normals.clear();
    for (const auto& [c, s] : wall) {
        const double norm = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
        a.wall_cells.push_back(c);
        for (int k = 0; k < 3; ++k) {
            a.wall_normals.push_back(norm > 0.0 ? s[k] / norm : 0.0);
        }
    }
    return LBM_OK;
}

int update_result(lbm_aero3d& a) {
    const Units& units = a.units;
    const double dx = units.dx;
    lbm_aero3d_result next{};
    next.iterations = lbm3d_steps(a.solver);
    // Moving average of force and moment over the window.
    const std::size_t n = std::min(a.window_samples, a.samples.size());
    Force mean;
    for (std::size_t k = a.samples.size() - n; k < a.samples.size(); ++k) {
        for (int c = 0; c < 3; ++c) {
            mean.f[c] += a.samples[k].f[c];
            mean.m[c] += a.samples[k].m[c];
        }
    }
    if (n > 0) {
        for (int c = 0; c < 3; ++c) {
            mean.f[c] /= static_cast<double>(n);
            mean.m[c] /= static_cast<double>(n);
        }
    }
    apply_force(a, mean, next);
    const double residual = next.cd != 0.0 ? std::abs(next.cd - a.previous_cd) / std::abs(next.cd) : 0.0;
    a.residuals.push_back(residual);
    a.previous_cd = next.cd;

    // Wall pressure: the densities of the wall nodes and of the plane next to the inlet,
    // sampled on the device.
    const lbm3d_config& c = a.config;
    std::vector<std::uint32_t> nodes(a.wall_cells);
    for (std::size_t z = 0; z < c.nz; ++z) {
        for (std::size_t y = 0; y < c.ny; ++y) {
            nodes.push_back(static_cast<std::uint32_t>(1 + c.nx * (y + c.ny * z)));
        }
    }
    std::vector<double> rho(nodes.size());
    if (int rc = lbm3d_sample(a.solver, nodes.data(), nodes.size(), rho.data(), nullptr, nullptr, nullptr);
        rc != LBM_OK) {
        return aero_fail(rc, std::string("lbm3d: ") + lbm_last_error());
    }
    const std::size_t walls = a.wall_cells.size();
    double rho_inf = 0.0;
    for (std::size_t k = walls; k < nodes.size(); ++k) {
        rho_inf += rho[k];
    }
    rho_inf /= static_cast<double>(c.ny * c.nz);
    const double q_lattice = 0.5 * units.u_lat * units.u_lat;
    const double share = walls == 0 ? 0.0 : a.surface_area / static_cast<double>(walls);
    a.points.assign(3 * walls, 0.0);
    a.normals.assign(3 * walls, 0.0);
    a.areas.assign(walls, share);
    a.cp.assign(walls, 0.0);
    const std::size_t plane = c.nx * c.ny;
    for (std::size_t k = 0; k < walls; ++k) {
        const std::uint32_t cell = a.wall_cells[k];
        // The node position: the wall point of a link with q = 0 (link_wall_point,
        // lbm3d_common.cpp:213-227).
        const double node[3] = {static_cast<double>(cell % plane % c.nx), static_cast<double>(cell % plane / c.nx),
                                static_cast<double>(cell / plane)};
        const double at[3] = {a.origin[0] + node[0] * dx, a.origin[1] + node[1] * dx, a.origin[2] + node[2] * dx};
        const V3 p = to_body(a, at);
        const V3 nrm = to_body(a, &a.wall_normals[3 * k]);
        a.points[3 * k] = p.x;
        a.points[3 * k + 1] = p.y;
        a.points[3 * k + 2] = p.z;
        a.normals[3 * k] = nrm.x;
        a.normals[3 * k + 1] = nrm.y;
        a.normals[3 * k + 2] = nrm.z;
        a.cp[k] = kCs2 * (rho[k] - rho_inf) / q_lattice;
    }

    a.warnings = a.setup_warnings;
    double speed = 0.0;
    if (int rc = lbm3d_max_speed(a.solver, &speed); rc != LBM_OK) {
        return aero_fail(rc, std::string("lbm3d: ") + lbm_last_error());
    }
    const double local_mach = lattice_mach(speed);
    if (local_mach > kLocalMachWarning) {
        a.warnings.push_back({"lbm.local_mach", "largest lattice Mach number in the flow " + format(local_mach) +
                                                    " (compressibility error of order Ma^2, A3)"});
    }
    const bool finished = next.iterations >= a.total_steps;
    if (finished) {
        double from = 0.5 * static_cast<double>(next.iterations);
        if (a.time.size() > a.window_samples) {
            from = std::min(from, a.time[a.time.size() - a.window_samples]);
        }
        next.converged = forces_settled(a, from) ? 1 : 0;
        if (!next.converged) {
            a.warnings.push_back(
                {"lbm.unconverged",
                 "the force history of the last half of the run is neither steady (drag, lift and side-force ranges "
                 "within " +
                     format(kSteadyResidual) + " of the mean drag) nor settled periodic shedding (period spread within " +
                     format(kMaxPeriodSpread) + ", amplitude and mean drag within " + format(kStationarity) +
                     " between the first and the last period)"});
        }
    }
    if (!std::isfinite(next.cd) || !std::isfinite(next.cl) || !std::isfinite(next.cy)) {
        a.warnings.push_back({"lbm.diverged", "the lattice state is no longer finite (numerical instability)"});
    }
    next.residual = residual;
    next.samples = a.time.size();
    next.wall_samples = walls;
    next.warnings = a.warnings.size();
    a.result = next;
    return LBM_OK;
}

} // namespace

extern "C" {

const char* lbm_aero3d_last_error(void) { return g_aero_error.c_str(); }

void lbm_aero3d_default_options(lbm_aero3d_options* o) {
    if (o == nullptr) {
        return;
    }
    *o = lbm_aero3d_options{};
    o->cells_per_length = 16.0;
    o->lattice_velocity = 0.05;
    o->upstream = 5.0;
    o->downstream = 10.0;
    o->width = 6.0;
    o->sides = LBM_WALL_SLIP;
    o->wall = LBM_SOLID_INTERPOLATED;
    o->les = LBM_LES_AUTOMATIC;
    o->smagorinsky_constant = 0.1;
    o->duration = 30.0;
    o->average_window = 8.0;
    o->start_ramp = 3.0;
    o->sponge_strength = 0.1;
    o->sponge_inlet = 0.0;
    o->sponge_outlet = 2.0;
    o->sponge_sides = 0.0;
    o->steps_per_call = 200;
    o->sample_every = 10;
    o->max_cells = 0;
    o->storage = LBM_STORE_F32;
}

void lbm_default_flow_conditions(lbm_flow_conditions* c) {
    if (c == nullptr) {
        return;
    }
    *c = lbm_flow_conditions{0.0, 1.225, 1.789e-5, 288.15, 0.0, 0.0, 1.4, 287.05};
}

int lbm_aero3d_create(const lbm_aero3d_options* options, lbm_aero3d** out) {
    if (options == nullptr || out == nullptr) {
        return aero_fail(LBM_EINVAL, "lbm3d: null argument");
    }
    auto* a = new (std::nothrow) lbm_aero3d();
    if (a == nullptr) {
        return aero_fail(LBM_ENOMEM, "lbm3d: host allocation failed");
    }
    a->options = *options;
    *out = a;
    return LBM_OK;
}

int lbm_aero3d_destroy(lbm_aero3d* a) {
    if (a != nullptr) {
        release_solver(*a);
        delete a;
    }
    return LBM_OK;
}

int lbm_aero3d_plan(lbm_aero3d* a, const lbm_mesh* mesh, const lbm_reference_values* reference,
                    const lbm_flow_conditions* conditions) {
    if (a == nullptr || reference == nullptr || conditions == nullptr) {
        return aero_fail(LBM_EINVAL, "lbm3d: null argument");
    }
    return plan(*a, mesh, *reference, *conditions);
}

int lbm_aero3d_setup(lbm_aero3d* a, const lbm_mesh* mesh, const lbm_reference_values* reference,
                     const lbm_flow_conditions* conditions) {
    if (a == nullptr || reference == nullptr || conditions == nullptr) {
        return aero_fail(LBM_EINVAL, "lbm3d: null argument");
    }
    if (int rc = plan(*a, mesh, *reference, *conditions); rc != LBM_OK) {
        return rc;
    }
    a->planned = false;
    lbm3d* h = nullptr;
    const double* distance = a->distance.empty() ? nullptr : a->distance.data();
    if (int rc = lbm3d_create(&a->config, a->solid.data(), distance, a->options.storage, &h); rc != LBM_OK) {
        return aero_fail(rc, lbm_last_error());
    }
    a->solver = h;
    if (a->ramp_steps > 0.0) {
        // Gradual start: the fluid at rest, the inlet still.
        const std::size_t cells = a->config.nx * a->config.ny * a->config.nz;
        std::vector<double> planes(4 * cells, 0.0);
        std::fill(planes.begin(), planes.begin() + static_cast<std::ptrdiff_t>(cells), 1.0);
        const double still[3] = {0.0, 0.0, 0.0};
        if (int rc = lbm3d_initialize(h, planes.data()); rc != LBM_OK) {
            release_solver(*a);
            return aero_fail(rc, lbm_last_error());
        }
        if (int rc = lbm3d_set_inlet(h, still); rc != LBM_OK) {
            release_solver(*a);
            return aero_fail(rc, lbm_last_error());
        }
    }
    if (int rc = wall_samples(*a); rc != LBM_OK) {
        release_solver(*a);
        return aero_fail(rc, lbm_last_error());
    }
    a->planned = true;
    return LBM_OK;
}

// step (lbm_aero_solver3d.cpp:342-411): steps_per_call lattice steps (fewer at the end), force
// samples at the steps that are multiples of sample_every, the ramped inlet constant over each
// sampling interval.
int lbm_aero3d_step(lbm_aero3d* a, int32_t* finished) {
    if (a == nullptr) {
        return aero_fail(LBM_EINVAL, "lbm3d: null argument");
    }
    if (a->solver == nullptr) {
        return aero_fail(LBM_ESTATE, "lbm3d: step() called before setup()");
    }
    lbm3d* s = a->solver;
    const std::uint64_t now0 = lbm3d_steps(s);
    const std::uint64_t target = now0 + std::min<std::uint64_t>(a->total_steps - now0, a->options.steps_per_call);
    const std::uint64_t every = a->options.sample_every;
    while (lbm3d_steps(s) < target) {
        const std::uint64_t now = lbm3d_steps(s);
        const std::uint64_t next_sample = now - now % every + every;
        const std::uint64_t until = std::min(target, next_sample);
        const std::uint64_t chunk = until - now;
        if (a->ramp_steps > 0.0) {
            const auto start = static_cast<double>(now - now % every);
            const double fraction = start >= a->ramp_steps
                                        ? 1.0
                                        : 0.5 * (1.0 - std::cos(std::numbers::pi * start / a->ramp_steps));
            const double v[3] = {fraction * a->config.inlet_ux, 0.0, 0.0};
            if (int rc = lbm3d_set_inlet(s, v); rc != LBM_OK) {
                return aero_fail(rc, lbm_last_error());
            }
        }
        if (int rc = lbm3d_step(s, chunk); rc != LBM_OK) {
            return aero_fail(rc, lbm_last_error());
        }
        if (lbm3d_steps(s) % every != 0) {
            continue;
        }
        Force f;
        if (int rc = lbm3d_force(s, f.f, f.m); rc != LBM_OK) {
            return aero_fail(rc, lbm_last_error());
        }
        // Moment about the moment centre: M_c = M_o - r_c x F.
        const double* r = a->centre;
        f.m[0] -= r[1] * f.f[2] - r[2] * f.f[1];
        f.m[1] -= r[2] * f.f[0] - r[0] * f.f[2];
        f.m[2] -= r[0] * f.f[1] - r[1] * f.f[0];
        a->samples.push_back(f);
        lbm_aero3d_result sample{};
        apply_force(*a, f, sample);
        a->instantaneous[0] = sample.cd;
        a->instantaneous[1] = sample.cl;
        a->instantaneous[2] = sample.cy;
        a->time.push_back(static_cast<double>(lbm3d_steps(s)));
        a->drag.push_back(sample.cd);
        a->lift.push_back(sample.cl);
        a->side.push_back(sample.cy);
    }
    if (int rc = update_result(*a); rc != LBM_OK) {
        return rc;
    }
    if (finished != nullptr) {
        *finished = a->result.iterations >= a->total_steps ? 1 : 0;
    }
    return LBM_OK;
}

int lbm_aero3d_get_result(const lbm_aero3d* a, lbm_aero3d_result* out) {
    if (a == nullptr || out == nullptr) {
        return aero_fail(LBM_EINVAL, "lbm3d: null argument");
    }
    *out = a->result;
    return LBM_OK;
}

int lbm_aero3d_history(const lbm_aero3d* a, double* time, double* drag, double* lift, double* side, size_t capacity,
                       size_t* count) {
    if (a == nullptr || count == nullptr) {
        return aero_fail(LBM_EINVAL, "lbm3d: null argument");
    }
    const std::size_t n = std::min(capacity, a->time.size());
    for (std::size_t k = 0; k < n; ++k) {
        if (time != nullptr) time[k] = a->time[k];
        if (drag != nullptr) drag[k] = a->drag[k];
        if (lift != nullptr) lift[k] = a->lift[k];
        if (side != nullptr) side[k] = a->side[k];
    }
    *count = a->time.size();
    return LBM_OK;
}

int lbm_aero3d_residuals(const lbm_aero3d* a, double* residuals, size_t capacity, size_t* count) {
    if (a == nullptr || count == nullptr) {
        return aero_fail(LBM_EINVAL, "lbm3d: null argument");
    }
    for (std::size_t k = 0; k < std::min(capacity, a->residuals.size()); ++k) {
        residuals[k] = a->residuals[k];
    }
    *count = a->residuals.size();
    return LBM_OK;
}

int lbm_aero3d_instantaneous(const lbm_aero3d* a, double coefficients[3]) {
    if (a == nullptr || coefficients == nullptr) {
        return aero_fail(LBM_EINVAL, "lbm3d: null argument");
    }
    for (int k = 0; k < 3; ++k) {
        coefficients[k] = a->instantaneous[k];
    }
    return LBM_OK;
}

int lbm_aero3d_surface(const lbm_aero3d* a, double* points, double* normals, double* areas, double* cp,
                       size_t capacity, size_t* count) {
    if (a == nullptr || count == nullptr) {
        return aero_fail(LBM_EINVAL, "lbm3d: null argument");
    }
    const std::size_t n = std::min(capacity, a->cp.size());
    for (std::size_t k = 0; k < n; ++k) {
        for (int c = 0; c < 3; ++c) {
            if (points != nullptr) points[3 * k + c] = a->points[3 * k + c];
            if (normals != nullptr) normals[3 * k + c] = a->normals[3 * k + c];
        }
        if (areas != nullptr) areas[k] = a->areas[k];
        if (cp != nullptr) cp[k] = a->cp[k];
    }
    *count = a->cp.size();
    return LBM_OK;
}

int lbm_aero3d_warning(const lbm_aero3d* a, size_t index, char* code, size_t code_capacity, char* message,
                       size_t message_capacity) {
    if (a == nullptr || index >= a->warnings.size()) {
        return aero_fail(LBM_EINVAL, "lbm3d: no such warning");
    }
    const Warning& w = a->warnings[index];
    if (code != nullptr && code_capacity > 0) {
        std::snprintf(code, code_capacity, "%s", w.code.c_str());
    }
    if (message != nullptr && message_capacity > 0) {
        std::snprintf(message, message_capacity, "%s", w.message.c_str());
    }
    return LBM_OK;
}

int lbm_aero3d_warning_count(const lbm_aero3d* a, size_t* count) {
    if (a == nullptr || count == nullptr) {
        return aero_fail(LBM_EINVAL, "lbm3d: null argument");
    }
    *count = a->warnings.size();
    return LBM_OK;
}

int lbm_aero3d_get_info(const lbm_aero3d* a, lbm_aero3d_info* out) {
    if (a == nullptr || out == nullptr) {
        return aero_fail(LBM_EINVAL, "lbm3d: null argument");
    }
    if (!a->planned) {
        return aero_fail(LBM_ESTATE, "lbm3d: no lattice set up");
    }
    lbm_aero3d_info i{};
    i.config = a->config;
    for (int c = 0; c < 3; ++c) {
        i.origin[c] = a->origin[c];
        i.centre[c] = a->centre[c];
        i.first[c] = a->first[c];
        i.count[c] = a->count[c];
        i.axes[3 * c] = a->axes[c].x;
        i.axes[3 * c + 1] = a->axes[c].y;
        i.axes[3 * c + 2] = a->axes[c].z;
    }
    i.dx = a->units.dx;
    i.dt = a->units.dt;
    i.reynolds = a->units.reynolds;
    i.lattice_velocity = a->units.u_lat;
    i.cells_per_length = a->units.cells;
    i.total_steps = a->total_steps;
    i.window_samples = a->window_samples;
    i.ramp_steps = a->ramp_steps;
    i.blockage = a->blockage;
    i.surface_area = a->surface_area;
    std::size_t solid = 0;
    for (std::uint8_t s : a->solid) {
        solid += s;
    }
    i.solid_nodes = solid;
    *out = i;
    return LBM_OK;
}

int lbm_aero3d_mask(const lbm_aero3d* a, uint8_t* solid, double* distance) {
    if (a == nullptr || !a->planned) {
        return aero_fail(LBM_ESTATE, "lbm3d: no lattice set up");
    }
    if (solid != nullptr) {
        std::memcpy(solid, a->solid.data(), a->solid.size());
    }
    if (distance != nullptr && !a->distance.empty()) {
        std::memcpy(distance, a->distance.data(), a->distance.size() * sizeof(double));
    }
    return LBM_OK;
}

int lbm_aero3d_lattice(const lbm_aero3d* a, lbm3d** out) {
    if (a == nullptr || out == nullptr) {
        return aero_fail(LBM_EINVAL, "lbm3d: null argument");
    }
    if (a->solver == nullptr) {
        return aero_fail(LBM_ESTATE, "lbm3d: no lattice set up");
    }
    *out = a->solver;
    return LBM_OK;
}

// The velocity field on the lattice in SI units, [cells][3] (lbm_aero_solver3d.cpp:535-549): only
// at zero angle of attack and sideslip, where body and lattice axes coincide.
int lbm_aero3d_velocity(lbm_aero3d* a, double* values) {
    if (a == nullptr || values == nullptr) {
        return aero_fail(LBM_EINVAL, "lbm3d: null argument");
    }
    if (a->solver == nullptr) {
        return aero_fail(LBM_ESTATE, "lbm3d: no lattice set up");
    }
    if (a->conditions.angle_of_attack != 0.0 || a->conditions.sideslip != 0.0) {
        return aero_fail(LBM_EINVAL, "lbm3d: the lattice is aligned with the wind, not the body: no velocity field at "
                                     "non-zero angle of attack or sideslip");
    }
    const std::size_t cells = a->config.nx * a->config.ny * a->config.nz;
    std::vector<double> u(cells), v(cells), w(cells);
    if (int rc = lbm3d_macroscopic(a->solver, nullptr, u.data(), v.data(), w.data()); rc != LBM_OK) {
        return aero_fail(rc, lbm_last_error());
    }
    for (std::size_t k = 0; k < cells; ++k) {
        values[3 * k] = a->units.velocity_to_si(u[k]);
        values[3 * k + 1] = a->units.velocity_to_si(v[k]);
        values[3 * k + 2] = a->units.velocity_to_si(w[k]);
    }
    return LBM_OK;
}

} // extern "C"
