// Pieces shared by the 2D and 3D aerodynamic adapters (aero2d.cpp, aero3d.cpp): the error text,
// FlowConditions directions (src/aero/flow_conditions.cpp:7-39), LatticeUnits
// (src/aero/lbm/lattice_units.cpp:22-52), the method's validity envelope
// (src/aero/regime.cpp:34-107) and set_coefficients (src/aero/aero_result.cpp:22-35).
#pragma once

#include <cmath>
#include <cstddef>
#include <numbers>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/orrery_lbm.h"

namespace lbm::aero {

// Last adapter error of this thread (both adapters report through it).
inline thread_local std::string g_aero_error;

inline int aero_fail(int code, const std::string& message) {
    g_aero_error = message;
    return code;
}

struct V3 {
    double x = 0.0, y = 0.0, z = 0.0;
};
inline V3 operator+(const V3& a, const V3& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator*(const V3& a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(const V3& a, const V3& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline V3 normalized(const V3& a) {
    const double n = std::hypot(a.x, a.y, a.z);
    return n > 0.0 ? V3{a.x / n, a.y / n, a.z / n} : V3{};
}

inline V3 drag_direction(const lbm_flow_conditions& c) {
    const double cb = std::cos(c.sideslip);
    return {std::cos(c.angle_of_attack) * cb, std::sin(c.sideslip), std::sin(c.angle_of_attack) * cb};
}
inline V3 lift_direction(const lbm_flow_conditions& c) {
    return normalized(cross(drag_direction(c), V3{0.0, 1.0, 0.0}));
}
inline V3 side_direction(const lbm_flow_conditions& c) { return cross(lift_direction(c), drag_direction(c)); }
inline double mach_of(const lbm_flow_conditions& c) {
    return c.speed / std::sqrt(c.gamma * c.gas_constant * c.temperature);
}
inline double reynolds_of(const lbm_flow_conditions& c, double length) {
    return c.density * c.speed * length / c.viscosity;
}
inline double dynamic_pressure(const lbm_flow_conditions& c) { return 0.5 * c.density * c.speed * c.speed; }

inline bool positive_finite(double v) { return std::isfinite(v) && v > 0.0; }

// operator<< of a double at the stream's default precision (6 significant digits).
inline std::string format(double value) {
    std::ostringstream out;
    out << value;
    return out.str();
}

inline std::string format4(double value) {
    std::ostringstream out;
    out.precision(4);
    out << value;
    return out.str();
}

constexpr double kLesReynolds = 2000.0;      // automatic Smagorinsky above this Re
constexpr double kMarginalTau = 0.51;        // below: BGK stability warning
constexpr double kLocalMachWarning = 0.2;    // largest lattice Mach in the flow
constexpr double kSteadyResidual = 1e-3;     // steady: force ranges / mean drag
constexpr double kMaxPeriodSpread = 0.05;    // periodic: period spread
constexpr double kStationarity = 0.02;       // periodic: amplitude and mean drag drift
constexpr double kMaxSteps = 9007199254740992.0; // 2^53 lattice steps
constexpr double kMaxLatticeMach = 0.1;
constexpr double kCs2 = 1.0 / 3.0;
constexpr std::size_t kMaxGridCells = std::size_t{1} << 27U;
constexpr double kIncompressibleMach = 0.3;
constexpr double kPanelMinReynolds = 1e5;
constexpr double kDegree = std::numbers::pi / 180.0;

struct Units {
    double dx = 0.0, dt = 0.0, rho = 0.0, u_lat = 0.0, nu_lat = 0.0, reynolds = 0.0, cells = 0.0;
    double tau() const { return 3.0 * nu_lat + 0.5; }
    // SI force of a lattice force in 3D.
    double force_unit() const { return 1.0 * rho * dx * dx * dx / (dt * dt) * dx; }
    // SI force per metre of span of a lattice force in 2D (lattice_units.h:79-81).
    double force_per_span_to_si(double lattice) const { return lattice * rho * dx * dx * dx / (dt * dt); }
    double velocity_to_si(double v) const { return v * dx / dt; }
};

inline double lattice_mach(double u) { return std::abs(u) / std::sqrt(kCs2); }

inline int make_units(double length, double velocity, double nu, double density, double cells_per_length,
                      double u_lat, Units& out) {
    if (!positive_finite(length) || !positive_finite(velocity) || !positive_finite(nu) || !positive_finite(density)) {
        return aero_fail(LBM_EINVAL,
                         "lattice units: length, velocity, viscosity and density must be positive and finite");
    }
    if (!positive_finite(cells_per_length) || !positive_finite(u_lat)) {
        return aero_fail(LBM_EINVAL,
                         "lattice units: cells per length and lattice velocity must be positive and finite");
    }
    const double mach = lattice_mach(u_lat);
    if (!(mach < kMaxLatticeMach)) {
        std::ostringstream message;
        message << "lattice units: lattice Mach number " << mach << " (u_lat " << u_lat << ") is not below "
                << kMaxLatticeMach << "; lower the lattice velocity (weak compressibility, APPROXIMATION A3)";
        return aero_fail(LBM_EINVAL, message.str());
    }
    const double dx = length / cells_per_length;
    const double dt = u_lat * dx / velocity;
    const double nu_lat = nu * dt / (dx * dx);
    if (!positive_finite(dx) || !positive_finite(dt) || !positive_finite(nu_lat)) {
        return aero_fail(LBM_EINVAL,
                         "lattice units: spacing, time step or viscosity underflows or overflows double precision");
    }
    out = Units{dx, dt, density, u_lat, nu_lat, velocity * length / nu, cells_per_length};
    return LBM_OK;
}

struct Warning {
    std::string code;
    std::string message;
};

inline void range_warning(std::vector<Warning>& out, const char* what, const char* label, double value, double lo,
                          double hi) {
    if (value >= lo && value <= hi) {
        return;
    }
    out.push_back({std::string("lbm.") + what, std::string("lbm: ") + label + " " + format4(value) +
                                                   " is outside the validity envelope [" + format4(lo) + ", " +
                                                   format4(hi) + "]; the result is computed but not reliable"});
}

// check_validity for the lattice Boltzmann envelope.
inline std::vector<Warning> check_validity(const lbm_flow_conditions& c, double length) {
    std::vector<Warning> out;
    const double max_angle = std::numbers::pi;
    range_warning(out, "mach", "Mach number", mach_of(c), 0.0, kIncompressibleMach);
    range_warning(out, "reynolds", "Reynolds number", reynolds_of(c, length), 0.0, kPanelMinReynolds);
    range_warning(out, "angle_of_attack", "angle of attack (deg)", c.angle_of_attack / kDegree, -max_angle / kDegree,
                  max_angle / kDegree);
    range_warning(out, "sideslip", "sideslip (deg)", c.sideslip / kDegree, -max_angle / kDegree, max_angle / kDegree);
    return out;
}

// set_coefficients: Cd, Cl, Cy along the wind axes and Cm about body y, over q S (and L).
inline void set_coefficients(const V3& force, const V3& moment, const lbm_flow_conditions& c,
                             const lbm_reference_values& reference, lbm_aero3d_result& r) {
    r.force[0] = force.x;
    r.force[1] = force.y;
    r.force[2] = force.z;
    r.moment[0] = moment.x;
    r.moment[1] = moment.y;
    r.moment[2] = moment.z;
    const double qs = dynamic_pressure(c) * reference.area;
    r.cd = dot(force, drag_direction(c)) / qs;
    r.cl = dot(force, lift_direction(c)) / qs;
    r.cy = dot(force, side_direction(c)) / qs;
    r.cm = moment.y / (qs * reference.length);
}

} // namespace lbm::aero
