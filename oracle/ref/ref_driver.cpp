// TEST INFRASTRUCTURE ONLY -- C-ABI driver over the reference's own setup code, compiled from
// the reference sources where they lie (oracle/ref/Makefile) into oracle/_ref/liborrery_ref.so.
// Tests call it to check the oracle's restatement and the device library's host setup one
// function at a time; nothing else loads it.
//
// The reference's CPU step (lbm3d_cpu.cpp, lbm2d_cpu.cpp) includes oneTBB, which this image
// lacks, so it is not built; the functions bound here compile from their own sources.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <optional>
#include <string>
#include <vector>

#include "../../include/orrery_lbm.h"
#include "lbm2d_internal.h"
#include "lbm3d_internal.h"
#include "orrery/aero/lbm/lattice_units.h"
#include "orrery/aero/lbm/lbm2d.h"
#include "orrery/aero/lbm/solid_mask.h"
#include "orrery/aero/lbm3d/lbm3d.h"
#include "orrery/aero/lbm3d/solid_mask3d.h"
#include "orrery/geom/reference_quantities.h"
#include "orrery/geom/naca.h"

using namespace orrery::aero::lbm;

namespace {

Wall wall_of(const lbm_wall& w) {
    Wall out;
    out.kind = w.kind == LBM_WALL_PERIODIC ? WallKind::periodic : (w.kind == LBM_WALL_SLIP ? WallKind::slip : WallKind::no_slip);
    out.velocity_x = w.velocity_x;
    return out;
}

Lbm3DConfig config3(const lbm3d_config& c) {
    Lbm3DConfig r;
    r.nx = c.nx;
    r.ny = c.ny;
    r.nz = c.nz;
    r.tau = c.tau;
    r.x_boundary = c.x_boundary == LBM_X_PERIODIC ? XBoundary::periodic : XBoundary::inflow_outflow;
    r.outlet = c.outlet == LBM_OUTLET_CONVECTIVE ? Outlet::convective : Outlet::extrapolation;
    r.y_low = wall_of(c.y_low);
    r.y_high = wall_of(c.y_high);
    r.z_low = wall_of(c.z_low);
    r.z_high = wall_of(c.z_high);
    r.inlet_ux = c.inlet_ux;
    r.inlet_uy = c.inlet_uy;
    r.inlet_uz = c.inlet_uz;
    r.body_force_x = c.body_force_x;
    r.body_force_y = c.body_force_y;
    r.body_force_z = c.body_force_z;
    r.smagorinsky = Smagorinsky{c.smagorinsky_enabled != 0, c.smagorinsky_constant};
    r.sponge = Sponge{c.sponge_strength, c.sponge_inlet, c.sponge_outlet, c.sponge_sides};
    r.solid_wall = c.solid_wall == LBM_SOLID_INTERPOLATED ? SolidWall::interpolated : SolidWall::halfway;
    return r;
}

Lbm2DConfig config2(const lbm2d_config& c) {
    Lbm2DConfig r;
    r.nx = c.nx;
    r.ny = c.ny;
    r.tau = c.tau;
    r.x_boundary = c.x_boundary == LBM_X_PERIODIC ? XBoundary::periodic : XBoundary::inflow_outflow;
    r.outlet = c.outlet == LBM_OUTLET_CONVECTIVE ? Outlet::convective : Outlet::extrapolation;
    r.bottom = wall_of(c.bottom);
    r.top = wall_of(c.top);
    r.inlet_ux = c.inlet_ux;
    r.inlet_uy = c.inlet_uy;
    r.body_force_x = c.body_force_x;
    r.body_force_y = c.body_force_y;
    r.smagorinsky = Smagorinsky{c.smagorinsky_enabled != 0, c.smagorinsky_constant};
    return r;
}

SolidMask3D mask3(const lbm3d_config& c, const std::uint8_t* solid, const double* distance) {
    SolidMask3D m;
    m.nx = c.nx;
    m.ny = c.ny;
    m.nz = c.nz;
    const std::size_t cells = c.nx * c.ny * c.nz;
    m.solid.assign(solid, solid + cells);
    if (distance != nullptr) {
        m.distance.assign(distance, distance + cells);
    }
    return m;
}

SolidMask2D mask2(const lbm2d_config& c, const std::uint8_t* solid) {
    SolidMask2D m;
    m.nx = c.nx;
    m.ny = c.ny;
    m.solid.assign(solid, solid + c.nx * c.ny);
    return m;
}

void put(char* err, std::size_t len, const std::optional<std::string>& msg) {
    if (err != nullptr && len > 0) {
        std::snprintf(err, len, "%s", msg ? msg->c_str() : "");
    }
}

void copy_mask(const SolidMask3D& m, std::uint8_t* solid, double* distance) {
    std::memcpy(solid, m.solid.data(), m.solid.size());
    if (distance != nullptr && !m.distance.empty()) {
        std::memcpy(distance, m.distance.data(), m.distance.size() * sizeof(double));
    }
}

} // namespace

extern "C" {

int ref3d_validate(const lbm3d_config* c, const std::uint8_t* solid, const double* distance, char* err, std::size_t len) {
    const auto p = validate(config3(*c), mask3(*c, solid, distance));
    put(err, len, p);
    return p ? 1 : 0;
}

std::size_t ref3d_boundary_links(const lbm3d_config* c, const std::uint8_t* solid, const double* distance,
                                 std::uint32_t* cell, std::uint32_t* dir, double* q, std::size_t cap) {
    const auto links = boundary_links(config3(*c), mask3(*c, solid, distance));
    for (std::size_t k = 0; k < links.size() && k < cap; ++k) {
        cell[k] = links[k].cell;
        dir[k] = links[k].direction;
        q[k] = links[k].q;
    }
    return links.size();
}

// The reference's Bouzidi reflection (lbm3d_internal.h:185-197) on a caller's state.
double ref3d_reflected(const double* f, std::size_t cells, std::size_t here, std::size_t farther, std::size_t i,
                       double q) {
    const std::vector<double> v(f, f + 19 * cells);
    return detail3d::reflected(v, cells, here, farther, i, q);
}

// Lbm3DSolver::force of the CPU solver (lbm3d_cpu.cpp:192-203: for each boundary link, the
// reflected population and add_link_force / link_wall_point, lbm3d_common.cpp:213-241) over a
// caller's population state f[i * cells + cell] and the reference's own boundary links.
void ref3d_force_of(const lbm3d_config* c, const std::uint8_t* solid, const double* distance, const double* f,
                    double* force, double* moment) {
    const Lbm3DConfig config = config3(*c);
    const SolidMask3D mask = mask3(*c, solid, distance);
    const std::size_t cells = c->nx * c->ny * c->nz;
    const std::vector<double> v(f, f + 19 * cells);
    const auto offset = detail3d::source_offsets(config);
    Force3D total;
    for (const BoundaryLink3D& link : boundary_links(config, mask)) {
        const std::size_t out = d3q19::opposite[link.direction];
        const auto farther =
            static_cast<std::size_t>(static_cast<std::ptrdiff_t>(link.cell) - offset[link.direction]);
        const double back = detail3d::reflected(v, cells, link.cell, farther, link.direction, link.q);
        add_link_force(total, config, link, v[out * cells + link.cell] + back);
    }
    for (int a = 0; a < 3; ++a) {
        force[a] = total.force[a];
        moment[a] = total.moment[a];
    }
}

void ref3d_classify(const lbm3d_config* c, const std::uint8_t* solid, std::uint8_t* out) {
    const auto classes = detail3d::classify_nodes(config3(*c), mask3(*c, solid, nullptr));
    std::memcpy(out, classes.data(), classes.size());
}

int ref3d_resolve(const lbm3d_config* c, const std::uint8_t* solid, std::uint64_t x, std::uint64_t y, std::uint64_t z,
                  int i, std::uint64_t* cell, int* direction, double* wall_ux) {
    const detail3d::Source s = detail3d::resolve_source(config3(*c), mask3(*c, solid, nullptr), x, y, z, static_cast<std::size_t>(i));
    *cell = s.cell;
    *direction = static_cast<int>(s.direction);
    *wall_ux = s.wall_ux;
    return static_cast<int>(s.kind);
}

double ref3d_link_fraction(const lbm3d_config* c, const std::uint8_t* solid, const double* distance, std::uint64_t x,
                           std::uint64_t y, std::uint64_t z, int i) {
    return detail3d::link_fraction(config3(*c), mask3(*c, solid, distance), x, y, z, static_cast<std::size_t>(i));
}

double ref3d_sponge(const lbm3d_config* c, std::uint64_t x, std::uint64_t y, std::uint64_t z) {
    return detail3d::sponge_sigma(config3(*c), x, y, z);
}

void ref3d_inlet_west(double* f, double ux, double uy, double uz) {
    d3q19::Populations p{};
    for (std::size_t i = 0; i < d3q19::q; ++i) {
        p[i] = f[i];
    }
    detail3d::velocity_inlet_west(p, ux, uy, uz);
    for (std::size_t i = 0; i < d3q19::q; ++i) {
        f[i] = p[i];
    }
}

double ref3d_equilibrium(int i, double rho, double ux, double uy, double uz) {
    return d3q19::equilibrium(static_cast<std::size_t>(i), rho, ux, uy, uz);
}

void ref3d_initial_populations(const lbm3d_config* c, const std::uint8_t* solid, const double* planes, double* out) {
    const std::size_t nx = c->nx;
    const std::size_t ny = c->ny;
    const std::size_t cells = nx * ny * c->nz;
    const auto f = initial_populations(config3(*c), mask3(*c, solid, nullptr),
                                       [&](std::size_t x, std::size_t y, std::size_t z) {
                                           const std::size_t k = x + nx * (y + ny * z);
                                           return d3q19::Moments{planes[k], planes[cells + k], planes[2 * cells + k],
                                                                 planes[3 * cells + k]};
                                       });
    std::memcpy(out, f.data(), f.size() * sizeof(double));
}

std::uint64_t ref3d_gpu_memory_bytes(const lbm3d_config* c, std::uint64_t links, std::uint64_t slots) {
    const auto b = gpu_memory_bytes(config3(*c), links, slots);
    return b ? *b : 0;
}

void ref3d_sphere_mask(std::uint64_t nx, std::uint64_t ny, std::uint64_t nz, double c0, double c1, double c2, double r,
                       std::uint8_t* solid, double* distance) {
    copy_mask(sphere_mask(nx, ny, nz, {c0, c1, c2}, r), solid, distance);
}

void ref3d_box_mask(std::uint64_t nx, std::uint64_t ny, std::uint64_t nz, const double* lo, const double* hi,
                    std::uint8_t* solid, double* distance) {
    copy_mask(box_mask(nx, ny, nz, {lo[0], lo[1], lo[2]}, {hi[0], hi[1], hi[2]}), solid, distance);
}

int ref2d_validate(const lbm2d_config* c, const std::uint8_t* solid, char* err, std::size_t len) {
    const auto p = validate(config2(*c), mask2(*c, solid));
    put(err, len, p);
    return p ? 1 : 0;
}

std::size_t ref2d_boundary_links(const lbm2d_config* c, const std::uint8_t* solid, std::uint32_t* cell,
                                 std::uint32_t* dir, std::size_t cap) {
    const auto links = boundary_links(config2(*c), mask2(*c, solid));
    for (std::size_t k = 0; k < links.size() && k < cap; ++k) {
        cell[k] = links[k].cell;
        dir[k] = links[k].direction;
    }
    return links.size();
}

int ref2d_resolve(const lbm2d_config* c, const std::uint8_t* solid, std::uint64_t x, std::uint64_t y, int i,
                  std::uint64_t* cell, int* direction, double* wall_ux) {
    const detail::Source s = detail::resolve_source(config2(*c), mask2(*c, solid), x, y, static_cast<std::size_t>(i));
    *cell = s.cell;
    *direction = static_cast<int>(s.direction);
    *wall_ux = s.wall_ux;
    return static_cast<int>(s.kind);
}

void ref2d_zou_he_west(double* f, double ux, double uy) {
    d2q9::Populations p{};
    for (std::size_t i = 0; i < d2q9::q; ++i) {
        p[i] = f[i];
    }
    detail::zou_he_west(p, ux, uy);
    for (std::size_t i = 0; i < d2q9::q; ++i) {
        f[i] = p[i];
    }
}

void ref2d_circle_mask(std::uint64_t nx, std::uint64_t ny, double cx, double cy, double r, std::uint8_t* solid) {
    const SolidMask2D m = circle_mask(nx, ny, cx, cy, r);
    std::memcpy(solid, m.solid.data(), m.solid.size());
}

void ref2d_polygon_mask(std::uint64_t nx, std::uint64_t ny, const double* points, std::size_t n, std::uint8_t* solid) {
    std::vector<orrery::geom::SectionPoint> polygon(n);
    for (std::size_t k = 0; k < n; ++k) {
        polygon[k] = orrery::geom::SectionPoint{points[2 * k], points[2 * k + 1]};
    }
    const SolidMask2D m = polygon_mask(nx, ny, polygon);
    std::memcpy(solid, m.solid.data(), m.solid.size());
}

void ref_naca_section(double max_camber, double camber_position, double thickness, int n, double* out) {
    orrery::geom::Naca4 a;
    a.max_camber = max_camber;
    a.camber_position = camber_position;
    a.thickness = thickness;
    const std::vector<orrery::geom::SectionPoint> points = a.section(n);
    for (std::size_t k = 0; k < points.size(); ++k) {
        out[2 * k] = points[k].x;
        out[2 * k + 1] = points[k].z;
    }
}

double ref_naca_section_area(double max_camber, double camber_position, double thickness) {
    orrery::geom::Naca4 a;
    a.max_camber = max_camber;
    a.camber_position = camber_position;
    a.thickness = thickness;
    return a.section_area();
}

double ref_lattice_tau(double length, double velocity, double viscosity, double density, double cells_per_length,
                       double lattice_velocity) {
    const auto u = LatticeUnits::create(SiFlow{length, velocity, viscosity, density},
                                        LatticeResolution{cells_per_length, lattice_velocity});
    return u ? u.value().tau() : std::nan("");
}

} // extern "C"

#include "orrery/aero/lbm/force_history.h"

extern "C" {

int ref_dominant_oscillation(const double* time, const double* values, std::size_t n, double from, double* out5) {
    const auto o = dominant_oscillation(std::span<const double>(time, n), std::span<const double>(values, n), from);
    if (!o) {
        return 0;
    }
    out5[0] = o->period;
    out5[1] = o->frequency;
    out5[2] = static_cast<double>(o->cycles);
    out5[3] = o->amplitude;
    out5[4] = o->period_spread;
    return 1;
}

int ref_settled_oscillation(const double* time, const double* drag, const double* lift, std::size_t n, double from,
                            double* out5) {
    const auto o = settled_oscillation(std::span<const double>(time, n), std::span<const double>(drag, n),
                                       std::span<const double>(lift, n), from);
    if (!o) {
        return 0;
    }
    out5[0] = o->period;
    out5[1] = o->frequency;
    out5[2] = static_cast<double>(o->cycles);
    out5[3] = o->amplitude;
    out5[4] = o->period_spread;
    return 1;
}

int ref_steady_since(const double* time, const double* drag, const double* lift, std::size_t n, double from) {
    return steady_since(std::span<const double>(time, n), std::span<const double>(drag, n),
                        std::span<const double>(lift, n), from)
               ? 1
               : 0;
}

double ref_mean_since(const double* time, const double* values, std::size_t n, double from) {
    return mean_since(std::span<const double>(time, n), std::span<const double>(values, n), from);
}

} // extern "C"

#include "orrery/geom/naca.h"
#include "orrery/geom/primitives.h"
#include "orrery/geom/voxel_grid.h"
#include "orrery/geom/winding_number.h"

namespace {

std::size_t copy_mesh(const orrery::geom::TriangleMesh& m, double* v, std::size_t vcap, std::uint32_t* t,
                      std::size_t tcap, std::size_t* nt) {
    for (std::size_t k = 0; k < m.vertices.size() && k < vcap; ++k) {
        v[3 * k] = m.vertices[k].x;
        v[3 * k + 1] = m.vertices[k].y;
        v[3 * k + 2] = m.vertices[k].z;
    }
    for (std::size_t k = 0; k < m.triangles.size() && k < tcap; ++k) {
        for (int a = 0; a < 3; ++a) {
            t[3 * k + a] = m.triangles[k][a];
        }
    }
    *nt = m.triangles.size();
    return m.vertices.size();
}

orrery::geom::TriangleMesh mesh_of(const double* v, std::size_t nv, const std::uint32_t* t, std::size_t nt) {
    orrery::geom::TriangleMesh m;
    for (std::size_t k = 0; k < nv; ++k) {
        m.vertices.push_back({v[3 * k], v[3 * k + 1], v[3 * k + 2]});
    }
    for (std::size_t k = 0; k < nt; ++k) {
        m.triangles.push_back({t[3 * k], t[3 * k + 1], t[3 * k + 2]});
    }
    return m;
}

} // namespace

extern "C" {

std::size_t ref_mesh_revolution(int kind, double a, double b, double c, double d, double e, int around,
                                double arc_step, double* v, std::size_t vcap, std::uint32_t* t, std::size_t tcap,
                                std::size_t* nt) {
    using namespace orrery::geom;
    RevolutionProfile p;
    switch (kind) {
        case 0: p = sphere_profile(a); break;
        case 1: p = cylinder_profile(a, b); break;
        case 2: p = cone_profile(a, b); break;
        default: {
            CapsuleShape s;
            s.max_diameter = a;
            s.nose_radius = b;
            s.shoulder_radius = c;
            s.afterbody_half_angle = d;
            s.top_diameter = e;
            p = capsule_profile(s);
        }
    }
    return copy_mesh(p.mesh(RevolutionResolution{around, arc_step}), v, vcap, t, tcap, nt);
}

std::size_t ref_mesh_wedge(double length, double half_angle, double span, double* v, std::size_t vcap,
                           std::uint32_t* t, std::size_t tcap, std::size_t* nt) {
    orrery::geom::Wedge w;
    w.length = length;
    w.half_angle = half_angle;
    w.span = span;
    return copy_mesh(w.mesh(), v, vcap, t, tcap, nt);
}

std::size_t ref_mesh_naca_wing(double m, double p, double th, double chord, double span, int n, int div, double* v,
                               std::size_t vcap, std::uint32_t* t, std::size_t tcap, std::size_t* nt) {
    orrery::geom::Naca4 a;
    a.max_camber = m;
    a.camber_position = p;
    a.thickness = th;
    return copy_mesh(orrery::geom::naca_wing(a, chord, span, n, div).mesh, v, vcap, t, tcap, nt);
}

void ref_voxelize(const double* v, std::size_t nv, const std::uint32_t* t, std::size_t nt, const double* origin,
                  double spacing, std::size_t nx, std::size_t ny, std::size_t nz, double threshold, std::uint8_t* solid,
                  double* winding) {
    using namespace orrery::geom;
    const TriangleMesh m = mesh_of(v, nv, t, nt);
    GridSpec g;
    g.origin = {origin[0], origin[1], origin[2]};
    g.spacing = spacing;
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    const VoxelGrid vg = voxelize(m, g, threshold);
    std::memcpy(solid, vg.solid.data(), vg.solid.size());
    std::memcpy(winding, vg.winding.data(), vg.winding.size() * sizeof(double));
}

// reference_quantities (src/geom/reference_quantities.cpp) with flow +x and lift +z: out = wetted,
// frontal, planform, frontal bound, planform bound, reference length, span, volume, centre x/y/z.
void ref_reference_quantities(const double* v, std::size_t nv, const std::uint32_t* t, std::size_t nt,
                              std::size_t resolution, double* out) {
    using namespace orrery::geom;
    const TriangleMesh m = mesh_of(v, nv, t, nt);
    ReferenceOptions o;
    o.raster_resolution = resolution;
    const ReferenceQuantities q = reference_quantities(m, o);
    const double values[11] = {q.wetted_area, q.frontal_area, q.planform_area, q.frontal_area_bound,
                               q.planform_area_bound, q.reference_length, q.span, q.volume,
                               q.centre_of_mass.x, q.centre_of_mass.y, q.centre_of_mass.z};
    std::memcpy(out, values, sizeof(values));
}

void ref_signed_distance_band(const double* v, std::size_t nv, const std::uint32_t* t, std::size_t nt,
                              const double* origin, double spacing, std::size_t nx, std::size_t ny, std::size_t nz,
                              const std::uint8_t* solid, double band_cells, double* distance) {
    using namespace orrery::geom;
    const TriangleMesh m = mesh_of(v, nv, t, nt);
    VoxelGrid vg;
    vg.spec.origin = {origin[0], origin[1], origin[2]};
    vg.spec.spacing = spacing;
    vg.spec.nx = nx;
    vg.spec.ny = ny;
    vg.spec.nz = nz;
    vg.solid.assign(solid, solid + nx * ny * nz);
    const SignedDistanceField f = signed_distance_band(m, vg, band_cells);
    std::memcpy(distance, f.distance.data(), f.distance.size() * sizeof(double));
}

} // extern "C"
