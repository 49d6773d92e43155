// The D2Q9 lattice behind Orrery's aerodynamic solver interface: LbmAeroSolver2D
// (src/aero/lbm/lbm_aero_solver.cpp:39-392, lbm_aero_solver.h:38-106) restated over this
// library's C ABI. A closed section (x-z plane, metres, per unit span) goes in; coefficients, SI
// force and pitching moment, the force history, the Strouhal number, wall pressure samples and
// validity warnings come out.
//
// The setup (wind frame, domain, polygon mask, step counts, lattice configuration, warnings)
// follows the reference operation for operation, so the lattice and mask are the reference's
// bit for bit. Per call, the reference reads the whole macroscopic field back to the host; here
// the wall-pressure samples, the density of the column next to the inlet and the largest speed
// are computed on the device and only those values cross the bus (lbm2d_sample,
// lbm2d_max_speed). The velocity field is produced on request.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "../../include/orrery_lbm.h"
#include "aero_common.h"

namespace {

using namespace lbm::aero;

constexpr double kMinCellsPerLength = 20.0; // below: resolution warning
constexpr double kBlockageWarning = 0.1;    // body height / channel height
constexpr double kMaxExtent = 1e7;

const int kCx[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
const int kCy[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};

struct Force2 {
    double x = 0.0, y = 0.0, moment = 0.0;
};

} // namespace

struct lbm_aero2d {
    lbm_aero2d_options options{};
    lbm_flow_conditions conditions{};
    lbm_reference_values reference{};
    Units units{};
    bool planned = false;
    lbm2d_config config{};
    std::vector<std::uint8_t> solid;
    lbm2d* solver = nullptr;
    std::vector<Warning> setup_warnings;
    std::vector<Warning> warnings;
    double origin[2] = {0, 0}; // wind-frame position (m) of lattice node (0, 0)
    double centre[2] = {0, 0}; // moment centre, lattice coordinates
    double blockage = 0.0;
    std::uint64_t total_steps = 0;
    std::size_t window_samples = 1;
    double previous_cd = 0.0;
    // CoefficientHistory and the lattice force samples
    std::vector<double> time, drag, lift;
    std::vector<Force2> samples;
    double instantaneous[2] = {0, 0};
    // result
    lbm_aero3d_result result{};
    std::vector<double> residuals;
    std::vector<std::uint32_t> wall_cells;
    std::vector<double> wall_sums; // summed link directions [2 n]
    std::vector<double> points, normals, areas, cp;
};

namespace {

void release_solver(lbm_aero2d& a) {
    if (a.solver != nullptr) {
        lbm2d_destroy(a.solver);
        a.solver = nullptr;
    }
}

// apply_force (lbm_aero_solver.cpp:237-247): per-span lattice force to SI, along the wind's drag
// and lift directions; the lattice z moment (x right, y up) is minus the pitching moment about
// body +y, with one more length unit than the force.
void apply_force(const lbm_aero2d& a, const Force2& lattice, lbm_aero3d_result& r) {
    const double drag = a.units.force_per_span_to_si(lattice.x);
    const double lift = a.units.force_per_span_to_si(lattice.y);
    const V3 dd = drag_direction(a.conditions);
    const V3 ld = lift_direction(a.conditions);
    const V3 force{drag * dd.x + lift * ld.x, drag * dd.y + lift * ld.y, drag * dd.z + lift * ld.z};
    const V3 moment{0.0, -a.units.force_per_span_to_si(lattice.moment) * a.units.dx, 0.0};
    set_coefficients(force, moment, a.conditions, a.reference, r);
}

// setup (lbm_aero_solver.cpp:39-198) up to (not including) the creation of the lattice.
int plan(lbm_aero2d& a, const double* section, std::size_t n, const lbm_reference_values& reference,
         const lbm_flow_conditions& conditions) {
    release_solver(a);
    a.planned = false;
    a.result = lbm_aero3d_result{};
    a.setup_warnings.clear();
    a.warnings.clear();
    a.residuals.clear();
    const lbm_aero2d_options& o = a.options;
    if (n < 3 || section == nullptr) {
        return aero_fail(LBM_EINVAL, "lbm: the 2D solver needs a section of at least 3 points");
    }
    for (std::size_t k = 0; k < 2 * n; ++k) {
        if (!std::isfinite(section[k])) {
            return aero_fail(LBM_EINVAL, "lbm: section points must be finite");
        }
    }
    if (!positive_finite(conditions.speed) || !positive_finite(conditions.density) ||
        !positive_finite(conditions.viscosity) || !std::isfinite(conditions.angle_of_attack) ||
        !std::isfinite(conditions.sideslip)) {
        return aero_fail(LBM_EINVAL, "lbm: speed, density and viscosity must be positive and finite, angles finite");
    }
    if (!positive_finite(reference.length) || !positive_finite(reference.area)) {
        return aero_fail(LBM_EINVAL, "lbm: reference length and area must be positive and finite");
    }
    if (!positive_finite(o.upstream) || !positive_finite(o.downstream) || !positive_finite(o.height) ||
        !positive_finite(o.duration) || !positive_finite(o.average_window) || o.steps_per_call == 0 ||
        o.sample_every == 0) {
        return aero_fail(LBM_EINVAL, "lbm: domain, duration and step options must be positive");
    }
    const double length = reference.length;
    Units units;
    if (int rc = make_units(length, conditions.speed, conditions.viscosity / conditions.density, conditions.density,
                            o.cells_per_length, o.lattice_velocity, units);
        rc != LBM_OK) {
        return rc;
    }
    const double dx = units.dx;

    // Wind frame: the free stream along +x, lift along +z.
    const double ca = std::cos(conditions.angle_of_attack);
    const double sa = std::sin(conditions.angle_of_attack);
    std::vector<double> wind(2 * n);
    double xmin = 0.0;
    double xmax = 0.0;
    double zmin = 0.0;
    double zmax = 0.0;
    for (std::size_t k = 0; k < n; ++k) {
        const double px = section[2 * k];
        const double pz = section[2 * k + 1];
        const double wx = px * ca + pz * sa;
        const double wz = -px * sa + pz * ca;
        xmin = k == 0 ? wx : std::min(xmin, wx);
        xmax = k == 0 ? wx : std::max(xmax, wx);
        zmin = k == 0 ? wz : std::min(zmin, wz);
        zmax = k == 0 ? wz : std::max(zmax, wz);
        wind[2 * k] = wx;
        wind[2 * k + 1] = wz;
    }
    const double nx_real = (o.upstream * length + (xmax - xmin) + o.downstream * length) / dx;
    const double ny_real = std::max(o.height * length, (zmax - zmin) + 4.0 * dx) / dx;
    if (!(nx_real < kMaxExtent) || !(ny_real < kMaxExtent)) {
        return aero_fail(LBM_EINVAL, "lbm: the domain is too large for the resolution");
    }
    const auto nx = static_cast<std::size_t>(std::ceil(nx_real)) + 1;
    const auto ny = static_cast<std::size_t>(std::ceil(ny_real)) + 1;
    if (nx > kMaxGridCells / ny) {
        return aero_fail(LBM_EINVAL, "lbm: the lattice exceeds geom::max_grid_cells nodes");
    }
    a.origin[0] = xmin - o.upstream * length;
    a.origin[1] = 0.5 * (zmin + zmax) - 0.5 * static_cast<double>(ny - 1) * dx;
    for (std::size_t k = 0; k < n; ++k) {
        wind[2 * k] = (wind[2 * k] - a.origin[0]) / dx;
        wind[2 * k + 1] = (wind[2 * k + 1] - a.origin[1]) / dx;
    }
    a.solid.assign(nx * ny, 0);
    if (int rc = lbm_polygon_mask(nx, ny, wind.data(), n, a.solid.data()); rc != LBM_OK) {
        return aero_fail(rc, std::string("lbm: ") + lbm_geom_last_error());
    }
    std::size_t solid_count = 0;
    for (std::uint8_t s : a.solid) {
        solid_count += s;
    }
    if (solid_count == 0) {
        return aero_fail(LBM_EINVAL, "lbm: the section covers no lattice node (resolution too low)");
    }

    // Step counts, checked before they are converted to integers.
    const double cells = units.cells;
    const double steps_per_time = cells / units.u_lat;
    const double total_real = std::ceil(o.duration * steps_per_time);
    const double window_real = o.average_window * steps_per_time / static_cast<double>(o.sample_every);
    if (!(total_real < kMaxSteps) || !(window_real < kMaxSteps)) {
        return aero_fail(LBM_EINVAL, "lbm: duration and averaging window must stay below " + format(kMaxSteps) +
                                         " lattice steps");
    }

    lbm2d_config& config = a.config;
    config = lbm2d_config{};
    config.nx = nx;
    config.ny = ny;
    config.tau = units.tau();
    config.x_boundary = LBM_X_INFLOW_OUTFLOW;
    config.outlet = LBM_OUTLET_CONVECTIVE;
    config.bottom = lbm_wall{LBM_WALL_SLIP, 0.0};
    config.top = lbm_wall{LBM_WALL_SLIP, 0.0};
    config.inlet_ux = units.u_lat;
    const double reynolds = units.reynolds;
    const bool les = o.les == LBM_LES_ON || (o.les == LBM_LES_AUTOMATIC && reynolds > kLesReynolds);
    config.smagorinsky_enabled = les ? 1 : 0;
    config.smagorinsky_constant = o.smagorinsky_constant;

    a.centre[0] = (reference.moment_centre[0] * ca + reference.moment_centre[2] * sa - a.origin[0]) / dx;
    a.centre[1] = (-reference.moment_centre[0] * sa + reference.moment_centre[2] * ca - a.origin[1]) / dx;
    a.units = units;
    a.conditions = conditions;
    a.reference = reference;
    a.total_steps = static_cast<std::uint64_t>(total_real);
    a.window_samples = std::max<std::size_t>(1, static_cast<std::size_t>(window_real));
    a.time.clear();
    a.drag.clear();
    a.lift.clear();
    a.samples.clear();
    a.instantaneous[0] = a.instantaneous[1] = 0.0;
    a.previous_cd = 0.0;

    // Envelope of the method plus the lattice-specific conditions.
    a.setup_warnings = check_validity(conditions, length);
    if (conditions.sideslip != 0.0) {
        a.setup_warnings.push_back(
            {"lbm.sideslip", "the 2D solver ignores the sideslip angle " + format(conditions.sideslip) + " rad"});
    }
    if (cells < kMinCellsPerLength) {
        a.setup_warnings.push_back({"lbm.resolution", format(cells) + " cells per reference length; below " +
                                                          format(kMinCellsPerLength) +
                                                          " the staircase wall dominates the force error (A4)"});
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
    a.blockage = (zmax - zmin) / (static_cast<double>(ny - 1) * dx);
    if (a.blockage > kBlockageWarning) {
        a.setup_warnings.push_back({"lbm.blockage", "the body fills " + format(100.0 * a.blockage) +
                                                        " percent of the channel height; confinement raises the drag"});
    }
    a.warnings = a.setup_warnings;
    a.planned = true;
    return LBM_OK;
}

// The wall samples: one per fluid node with a link to the body, in cell order, with the sum of
// its link directions (lbm_aero_solver.cpp:308-313).
int wall_samples(lbm_aero2d& a) {
    std::size_t n = 0;
    if (int rc = lbm2d_links(a.solver, nullptr, nullptr, 0, &n); rc != LBM_OK) {
        return rc;
    }
    std::vector<std::uint32_t> cell(n);
    std::vector<std::uint32_t> dir(n);
    if (int rc = lbm2d_links(a.solver, cell.data(), dir.data(), n, &n); rc != LBM_OK) {
        return rc;
    }
    std::map<std::uint32_t, std::array<double, 2>> wall;
    for (std::size_t k = 0; k < n; ++k) {
        std::array<double, 2>& s = wall[cell[k]];
        s[0] += kCx[dir[k]];
        s[1] += kCy[dir[k]];
    }
    a.wall_cells.clear();
    a.wall_sums.clear();
    for (const auto& [c, s] : wall) {
        a.wall_cells.push_back(c);
        a.wall_sums.push_back(s[0]);
        a.wall_sums.push_back(s[1]);
    }
    return LBM_OK;
}

// update_result (lbm_aero_solver.cpp:265-389).
int update_result(lbm_aero2d& a) {
    const Units& units = a.units;
    const double dx = units.dx;
    lbm_aero3d_result next{};
    next.iterations = lbm2d_steps(a.solver);

    // Moving average of force and moment over the window.
    const std::size_t n = std::min(a.window_samples, a.samples.size());
    Force2 mean;
    for (std::size_t k = a.samples.size() - n; k < a.samples.size(); ++k) {
        mean.x += a.samples[k].x;
        mean.y += a.samples[k].y;
        mean.moment += a.samples[k].moment;
    }
    if (n > 0) {
        mean.x /= static_cast<double>(n);
        mean.y /= static_cast<double>(n);
        mean.moment /= static_cast<double>(n);
    }
    apply_force(a, mean, next);
    const double residual = next.cd != 0.0 ? std::abs(next.cd - a.previous_cd) / std::abs(next.cd) : 0.0;
    a.residuals.push_back(residual);
    a.previous_cd = next.cd;

    // Wall pressure: the densities of the wall nodes and of the column next to the inlet,
    // sampled on the device.
    const lbm2d_config& c = a.config;
    std::vector<std::uint32_t> nodes(a.wall_cells);
    for (std::size_t y = 0; y < c.ny; ++y) {
        nodes.push_back(static_cast<std::uint32_t>(1 + c.nx * y));
    }
    std::vector<double> rho(nodes.size());
    if (int rc = lbm2d_sample(a.solver, nodes.data(), nodes.size(), rho.data(), nullptr, nullptr); rc != LBM_OK) {
        return aero_fail(rc, std::string("lbm: ") + lbm_last_error());
    }
    const std::size_t walls = a.wall_cells.size();
    double rho_inf = 0.0;
    for (std::size_t k = walls; k < nodes.size(); ++k) {
        rho_inf += rho[k];
    }
    rho_inf /= static_cast<double>(c.ny);
    const double ca = std::cos(a.conditions.angle_of_attack);
    const double sa = std::sin(a.conditions.angle_of_attack);
    const double q_lattice = 0.5 * units.u_lat * units.u_lat;
    a.points.assign(3 * walls, 0.0);
    a.normals.assign(3 * walls, 0.0);
    a.areas.assign(walls, dx); // per metre of span
    a.cp.assign(walls, 0.0);
    for (std::size_t k = 0; k < walls; ++k) {
        const std::uint32_t cell = a.wall_cells[k];
        const std::size_t column = cell % c.nx;
        const std::size_t row = cell / c.nx;
        const double xw = a.origin[0] + static_cast<double>(column) * dx;
        const double zw = a.origin[1] + static_cast<double>(row) * dx;
        const double sx = a.wall_sums[2 * k];
        const double sz = a.wall_sums[2 * k + 1];
        const double norm = std::hypot(sx, sz);
        const double nxw = norm > 0.0 ? sx / norm : 0.0;
        const double nzw = norm > 0.0 ? sz / norm : 0.0;
        a.points[3 * k] = xw * ca - zw * sa;
        a.points[3 * k + 2] = xw * sa + zw * ca;
        a.normals[3 * k] = nxw * ca - nzw * sa;
        a.normals[3 * k + 2] = nxw * sa + nzw * ca;
        a.cp[k] = kCs2 * (rho[k] - rho_inf) / q_lattice;
    }

    a.warnings = a.setup_warnings;
    double speed = 0.0;
    if (int rc = lbm2d_max_speed(a.solver, &speed); rc != LBM_OK) {
        return aero_fail(rc, std::string("lbm: ") + lbm_last_error());
    }
    const double local_mach = lattice_mach(speed);
    if (local_mach > kLocalMachWarning) {
        a.warnings.push_back({"lbm.local_mach", "largest lattice Mach number in the flow " + format(local_mach) +
                                                    " (compressibility error of order Ma^2, A3)"});
    }
    // Converged: the force history over the last half of the run (or the averaging window, if
    // longer) is steady, or settled periodic shedding.
    const bool finished = next.iterations >= a.total_steps;
    if (finished) {
        const std::size_t size = a.time.size();
        double from = 0.5 * static_cast<double>(next.iterations);
        if (size > a.window_samples) {
            from = std::min(from, a.time[size - a.window_samples]);
        }
        const bool steady =
            lbm_steady_since(a.time.data(), a.drag.data(), a.lift.data(), size, from, kSteadyResidual) == 1;
        lbm_oscillation o{};
        const bool periodic = lbm_settled_oscillation(a.time.data(), a.drag.data(), a.lift.data(), size, from,
                                                      kMaxPeriodSpread, kStationarity, &o) == 1;
        next.converged = steady || periodic ? 1 : 0;
        if (!next.converged) {
            a.warnings.push_back(
                {"lbm.unconverged",
                 "the force history of the last half of the run is neither steady (drag and lift ranges within " +
                     format(kSteadyResidual) + " of the mean drag) nor settled periodic shedding (period spread within " +
                     format(kMaxPeriodSpread) + ", lift amplitude and mean drag within " + format(kStationarity) +
                     " between the first and the last period)"});
        }
    }
    if (!std::isfinite(next.cd) || !std::isfinite(next.cl)) {
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

const char* lbm_aero2d_last_error(void) { return g_aero_error.c_str(); }

void lbm_aero2d_default_options(lbm_aero2d_options* o) {
    if (o == nullptr) {
        return;
    }
    *o = lbm_aero2d_options{};
    o->cells_per_length = 40.0;
    o->lattice_velocity = 0.05;
    o->upstream = 8.0;
    o->downstream = 16.0;
    o->height = 16.0;
    o->les = LBM_LES_AUTOMATIC;
    o->smagorinsky_constant = 0.1;
    o->duration = 60.0;
    o->average_window = 6.0;
    o->steps_per_call = 500;
    o->sample_every = 10;
    o->storage = LBM_STORE_F32;
}

int lbm_aero2d_create(const lbm_aero2d_options* options, lbm_aero2d** out) {
    if (options == nullptr || out == nullptr) {
        return aero_fail(LBM_EINVAL, "lbm: null argument");
    }
    auto* a = new (std::nothrow) lbm_aero2d();
    if (a == nullptr) {
        return aero_fail(LBM_ENOMEM, "lbm: host allocation failed");
    }
    a->options = *options;
    *out = a;
    return LBM_OK;
}

int lbm_aero2d_destroy(lbm_aero2d* a) {
    if (a != nullptr) {
        release_solver(*a);
        delete a;
    }
    return LBM_OK;
}

int lbm_aero2d_plan(lbm_aero2d* a, const double* section, size_t n, const lbm_reference_values* reference,
                    const lbm_flow_conditions* conditions) {
    if (a == nullptr || reference == nullptr || conditions == nullptr) {
        return aero_fail(LBM_EINVAL, "lbm: null argument");
    }
    return plan(*a, section, n, *reference, *conditions);
}

int lbm_aero2d_setup(lbm_aero2d* a, const double* section, size_t n, const lbm_reference_values* reference,
                     const lbm_flow_conditions* conditions) {
    if (a == nullptr || reference == nullptr || conditions == nullptr) {
        return aero_fail(LBM_EINVAL, "lbm: null argument");
    }
    if (int rc = plan(*a, section, n, *reference, *conditions); rc != LBM_OK) {
        return rc;
    }
    a->planned = false;
    lbm2d* h = nullptr;
    if (int rc = lbm2d_create(&a->config, a->solid.data(), a->options.storage, &h); rc != LBM_OK) {
        return aero_fail(rc, lbm_last_error());
    }
    a->solver = h;
    if (int rc = wall_samples(*a); rc != LBM_OK) {
        release_solver(*a);
        return aero_fail(rc, lbm_last_error());
    }
    a->planned = true;
    return LBM_OK;
}

// step (lbm_aero_solver.cpp:200-235): steps_per_call lattice steps (fewer at the end), force
// samples at the steps that are multiples of sample_every, whatever the call size.
int lbm_aero2d_step(lbm_aero2d* a, int32_t* finished) {
    if (a == nullptr) {
        return aero_fail(LBM_EINVAL, "lbm: null argument");
    }
    if (a->solver == nullptr) {
        return aero_fail(LBM_ESTATE, "lbm: step() called before setup()");
    }
    lbm2d* s = a->solver;
    const std::uint64_t now0 = lbm2d_steps(s);
    const std::uint64_t target = now0 + std::min<std::uint64_t>(a->total_steps - now0, a->options.steps_per_call);
    const std::uint64_t every = a->options.sample_every;
    while (lbm2d_steps(s) < target) {
        const std::uint64_t now = lbm2d_steps(s);
        const std::uint64_t next_sample = now - now % every + every;
        const std::uint64_t until = std::min(target, next_sample);
        if (int rc = lbm2d_step(s, until - now); rc != LBM_OK) {
            return aero_fail(rc, lbm_last_error());
        }
        if (lbm2d_steps(s) % every != 0) {
            continue; // the call ends between two samples
        }
        Force2 f;
        if (int rc = lbm2d_force(s, &f.x, &f.y, &f.moment); rc != LBM_OK) {
            return aero_fail(rc, lbm_last_error());
        }
        // Moment about the moment centre instead of the lattice origin.
        f.moment -= a->centre[0] * f.y - a->centre[1] * f.x;
        a->samples.push_back(f);
        lbm_aero3d_result sample{};
        apply_force(*a, f, sample);
        a->instantaneous[0] = sample.cd;
        a->instantaneous[1] = sample.cl;
        a->time.push_back(static_cast<double>(lbm2d_steps(s)));
        a->drag.push_back(sample.cd);
        a->lift.push_back(sample.cl);
    }
    if (int rc = update_result(*a); rc != LBM_OK) {
        return rc;
    }
    if (finished != nullptr) {
        *finished = a->result.iterations >= a->total_steps ? 1 : 0;
    }
    return LBM_OK;
}

int lbm_aero2d_get_result(const lbm_aero2d* a, lbm_aero3d_result* out) {
    if (a == nullptr || out == nullptr) {
        return aero_fail(LBM_EINVAL, "lbm: null argument");
    }
    *out = a->result;
    return LBM_OK;
}

int lbm_aero2d_history(const lbm_aero2d* a, double* time, double* drag, double* lift, size_t capacity,
                       size_t* count) {
    if (a == nullptr || count == nullptr) {
        return aero_fail(LBM_EINVAL, "lbm: null argument");
    }
    const std::size_t n = std::min(capacity, a->time.size());
    for (std::size_t k = 0; k < n; ++k) {
        if (time != nullptr) time[k] = a->time[k];
        if (drag != nullptr) drag[k] = a->drag[k];
        if (lift != nullptr) lift[k] = a->lift[k];
    }
    *count = a->time.size();
    return LBM_OK;
}

int lbm_aero2d_residuals(const lbm_aero2d* a, double* residuals, size_t capacity, size_t* count) {
    if (a == nullptr || count == nullptr) {
        return aero_fail(LBM_EINVAL, "lbm: null argument");
    }
    for (std::size_t k = 0; k < std::min(capacity, a->residuals.size()); ++k) {
        residuals[k] = a->residuals[k];
    }
    *count = a->residuals.size();
    return LBM_OK;
}

int lbm_aero2d_instantaneous(const lbm_aero2d* a, double coefficients[2]) {
    if (a == nullptr || coefficients == nullptr) {
        return aero_fail(LBM_EINVAL, "lbm: null argument");
    }
    coefficients[0] = a->instantaneous[0];
    coefficients[1] = a->instantaneous[1];
    return LBM_OK;
}

// strouhal (lbm_aero_solver.cpp:249-263): the dominant lift oscillation over the averaging
// window, f L / U from the frequency in lattice steps.
int lbm_aero2d_strouhal(const lbm_aero2d* a, double* strouhal, int32_t* found) {
    if (a == nullptr || strouhal == nullptr || found == nullptr) {
        return aero_fail(LBM_EINVAL, "lbm: null argument");
    }
    *found = 0;
    *strouhal = 0.0;
    const std::size_t size = a->time.size();
    if (a->solver == nullptr || size < 2) {
        return LBM_OK;
    }
    const std::size_t first = size > a->window_samples ? size - a->window_samples : 0;
    lbm_oscillation o{};
    if (lbm_dominant_oscillation(a->time.data(), a->lift.data(), size, a->time[first], &o) != 1) {
        return LBM_OK;
    }
    *strouhal = o.frequency * a->units.cells / a->units.u_lat;
    *found = 1;
    return LBM_OK;
}

int lbm_aero2d_surface(const lbm_aero2d* a, double* points, double* normals, double* areas, double* cp,
                       size_t capacity, size_t* count) {
    if (a == nullptr || count == nullptr) {
        return aero_fail(LBM_EINVAL, "lbm: null argument");
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

int lbm_aero2d_warning_count(const lbm_aero2d* a, size_t* count) {
    if (a == nullptr || count == nullptr) {
        return aero_fail(LBM_EINVAL, "lbm: null argument");
    }
    *count = a->warnings.size();
    return LBM_OK;
}

int lbm_aero2d_warning(const lbm_aero2d* a, size_t index, char* code, size_t code_capacity, char* message,
                       size_t message_capacity) {
    if (a == nullptr || index >= a->warnings.size()) {
        return aero_fail(LBM_EINVAL, "lbm: no such warning");
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

int lbm_aero2d_get_info(const lbm_aero2d* a, lbm_aero2d_info* out) {
    if (a == nullptr || out == nullptr) {
        return aero_fail(LBM_EINVAL, "lbm: null argument");
    }
    if (!a->planned) {
        return aero_fail(LBM_ESTATE, "lbm: no lattice set up");
    }
    lbm_aero2d_info i{};
    i.config = a->config;
    i.origin[0] = a->origin[0];
    i.origin[1] = a->origin[1];
    i.centre[0] = a->centre[0];
    i.centre[1] = a->centre[1];
    i.dx = a->units.dx;
    i.dt = a->units.dt;
    i.reynolds = a->units.reynolds;
    i.lattice_velocity = a->units.u_lat;
    i.cells_per_length = a->units.cells;
    i.total_steps = a->total_steps;
    i.window_samples = a->window_samples;
    i.blockage = a->blockage;
    std::size_t solid = 0;
    for (std::uint8_t s : a->solid) {
        solid += s;
    }
    i.solid_nodes = solid;
    *out = i;
    return LBM_OK;
}

int lbm_aero2d_mask(const lbm_aero2d* a, uint8_t* solid) {
    if (a == nullptr || !a->planned) {
        return aero_fail(LBM_ESTATE, "lbm: no lattice set up");
    }
    if (solid != nullptr) {
        std::memcpy(solid, a->solid.data(), a->solid.size());
    }
    return LBM_OK;
}

int lbm_aero2d_lattice(const lbm_aero2d* a, lbm2d** out) {
    if (a == nullptr || out == nullptr) {
        return aero_fail(LBM_EINVAL, "lbm: null argument");
    }
    if (a->solver == nullptr) {
        return aero_fail(LBM_ESTATE, "lbm: no lattice set up");
    }
    *out = a->solver;
    return LBM_OK;
}

// The velocity field on the lattice in SI units, [ny][nx][3] with components (x, 0, z)
// (lbm_aero_solver.cpp:324-337): only at zero angle of attack, where body and lattice axes
// coincide.
int lbm_aero2d_velocity(lbm_aero2d* a, double* values) {
    if (a == nullptr || values == nullptr) {
        return aero_fail(LBM_EINVAL, "lbm: null argument");
    }
    if (a->solver == nullptr) {
        return aero_fail(LBM_ESTATE, "lbm: no lattice set up");
    }
    if (a->conditions.angle_of_attack != 0.0) {
        return aero_fail(LBM_EINVAL,
                         "lbm: the lattice is aligned with the wind, not the body: no velocity field at non-zero "
                         "angle of attack");
    }
    const std::size_t cells = a->config.nx * a->config.ny;
    std::vector<double> u(cells), v(cells);
    if (int rc = lbm2d_macroscopic(a->solver, nullptr, u.data(), v.data()); rc != LBM_OK) {
        return aero_fail(rc, lbm_last_error());
    }
    for (std::size_t k = 0; k < cells; ++k) {
        values[3 * k] = a->units.velocity_to_si(u[k]);
        values[3 * k + 1] = 0.0;
        values[3 * k + 2] = a->units.velocity_to_si(v[k]);
    }
    return LBM_OK;
}

} // extern "C"
