// Drives orrery_amd::make_hip_solver (include/orrery_lbm_solver.hpp) through the reference's own
// Lbm3DSolver types: prints the refusal message of each configuration named on the command line,
// or "created" when the solver was made. Built by tests/test_adapter.py (and by the package
// build, for the device runs of tests/test_gpu_adapter.py).
//
// Device runs through the reference interface:
//   run3d FILE   a 40 x 20 x 18 sphere channel (Smagorinsky, convective outlet), a seeded
//                initial state, step(7), set_inlet_velocity, step(33); prints the step count,
//                force, moment, mass and inlet velocity, and writes populations() and the
//                density and velocity planes of macroscopic() to FILE (float64).
//   run2d FILE   a 48 x 24 D2Q9 channel around a square: step(9), step(31), the same readbacks.
//   solve3d      the IAeroSolver wrapper on a unit cube: step() until it reports the end;
//                prints the result's coefficients, force and iteration count.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "orrery/geom/primitives.h"
#include "orrery_lbm_aero.hpp"
#include "orrery_lbm_solver.hpp"

namespace ref = orrery::aero::lbm;

namespace {

// A small integer pattern for seeded states: the same doubles in any language.
double pattern(std::size_t k, std::size_t period) {
    return static_cast<double>(k % period) / static_cast<double>(period);
}

bool write_all(const char* path, const std::vector<const std::vector<double>*>& arrays) {
    std::FILE* out = std::fopen(path, "wb");
    if (out == nullptr) {
        return false;
    }
    bool ok = true;
    for (const auto* a : arrays) {
        ok = ok && std::fwrite(a->data(), sizeof(double), a->size(), out) == a->size();
    }
    return std::fclose(out) == 0 && ok;
}

int run3d(const char* path) {
    ref::Lbm3DConfig c;
    c.nx = 40;
    c.ny = 20;
    c.nz = 18;
    c.tau = 0.6;
    c.inlet_ux = 0.05;
    c.smagorinsky.enabled = true;
    c.smagorinsky.constant = 0.17;
    ref::SolidMask3D m;
    m.nx = c.nx;
    m.ny = c.ny;
    m.nz = c.nz;
    m.solid.assign(c.nx * c.ny * c.nz, 0);
    for (std::size_t z = 0; z < c.nz; ++z) {
        for (std::size_t y = 0; y < c.ny; ++y) {
            for (std::size_t x = 0; x < c.nx; ++x) {
                const long dx = static_cast<long>(x) - 12;
                const long dy = static_cast<long>(y) - 10;
                const long dz = static_cast<long>(z) - 9;
                m.solid[x + c.nx * (y + c.ny * z)] = dx * dx + dy * dy + dz * dz <= 16 ? 1 : 0;
            }
        }
    }
    auto made = orrery_amd::make_hip_solver(c, m);
    if (!made) {
        std::printf("run3d\t%s\n", made.error().c_str());
        return 1;
    }
    ref::Lbm3DSolver& s = *made.value();
    const ref::InitialState3D initial = [](std::size_t x, std::size_t y, std::size_t z) {
        ref::d3q19::Moments q;
        q.rho = 1.0 + 1e-3 * pattern(7 * x + 3 * y + 5 * z, 11);
        q.ux = 0.05 + 1e-3 * pattern(x + 2 * y, 5);
        q.uy = 1e-3 * pattern(3 * x + z, 7);
        q.uz = -5e-4 * pattern(y + z, 3);
        return q;
    };
    if (!s.initialize(initial) || !s.step(7) || !s.set_inlet_velocity({0.045, 0.002, -0.001}) || !s.step(33)) {
        std::printf("run3d\tfailed\n");
        return 1;
    }
    auto f = s.force();
    auto mass = s.mass();
    auto pops = s.populations();
    auto macro = s.macroscopic();
    if (!f || !mass || !pops || !macro) {
        std::printf("run3d\treadback failed\n");
        return 1;
    }
    const auto v = s.inlet_velocity();
    std::printf("run3d\t%llu", static_cast<unsigned long long>(s.steps()));
    for (double d : {f.value().force[0], f.value().force[1], f.value().force[2], f.value().moment[0],
                     f.value().moment[1], f.value().moment[2], mass.value(), v[0], v[1], v[2]}) {
        std::printf("\t%.17g", d);
    }
    std::printf("\t%zu\n", s.links().size());
    const auto& mf = macro.value();
    return write_all(path, {&pops.value(), &mf.density, &mf.velocity_x, &mf.velocity_y, &mf.velocity_z}) ? 0 : 1;
}

int run2d(const char* path) {
    ref::Lbm2DConfig c;
    c.nx = 48;
    c.ny = 24;
    c.tau = 0.58;
    c.inlet_ux = 0.04;
    ref::SolidMask2D m;
    m.nx = c.nx;
    m.ny = c.ny;
    m.solid.assign(c.nx * c.ny, 0);
    for (std::size_t y = 9; y < 14; ++y) {
        for (std::size_t x = 10; x < 15; ++x) {
            m.solid[x + c.nx * y] = 1;
        }
    }
    auto made = orrery_amd::make_hip_solver(c, m);
    if (!made) {
        std::printf("run2d\t%s\n", made.error().c_str());
        return 1;
    }
    ref::Lbm2DSolver& s = *made.value();
    const ref::InitialState initial = [](std::size_t x, std::size_t y) {
        ref::d2q9::Moments q;
        q.rho = 1.0 + 1e-3 * pattern(5 * x + 3 * y, 13);
        q.ux = 0.04 + 1e-3 * pattern(x + y, 7);
        q.uy = 1e-3 * pattern(2 * x + y, 5);
        return q;
    };
    if (!s.initialize(initial) || !s.step(9) || !s.step(31)) {
        std::printf("run2d\tfailed\n");
        return 1;
    }
    auto f = s.force();
    auto mass = s.mass();
    auto pops = s.populations();
    auto macro = s.macroscopic();
    if (!f || !mass || !pops || !macro) {
        std::printf("run2d\treadback failed\n");
        return 1;
    }
    std::printf("run2d\t%llu\t%.17g\t%.17g\t%.17g\t%.17g\n", static_cast<unsigned long long>(s.steps()),
                f.value().x, f.value().y, f.value().moment, mass.value());
    const auto& mf = macro.value();
    return write_all(path, {&pops.value(), &mf.density, &mf.velocity_x, &mf.velocity_y}) ? 0 : 1;
}

int solve3d() {
    orrery::aero::AeroGeometry g;
    g.mesh.vertices = {{-0.01, -0.01, -0.01}, {0.01, -0.01, -0.01}, {-0.01, 0.01, -0.01}, {0.01, 0.01, -0.01},
                       {-0.01, -0.01, 0.01},  {0.01, -0.01, 0.01},  {-0.01, 0.01, 0.01},  {0.01, 0.01, 0.01}};
    g.mesh.triangles = {{0, 4, 6}, {0, 6, 2}, {1, 3, 7}, {1, 7, 5}, {0, 1, 5}, {0, 5, 4},
                        {2, 6, 7}, {2, 7, 3}, {0, 2, 3}, {0, 3, 1}, {4, 5, 7}, {4, 7, 6}};
    g.reference.length = 0.02;
    g.reference.area = 0.0004;
    orrery::aero::FlowConditions f;
    f.speed = 0.1;
    f.density = 1000.0;
    f.viscosity = 0.1;
    ref::LbmSolver3DOptions o;
    o.resolution.cells_per_length = 8.0;
    o.resolution.lattice_velocity = 0.05;
    o.upstream = 2.0;
    o.downstream = 3.0;
    o.width = 3.0;
    o.duration = 2.0;
    o.average_window = 1.0;
    o.steps_per_call = 40;
    orrery_amd::HipLbmAeroSolver3D solver(o);
    solver.setup(g, f);
    int calls = 0;
    while (!solver.step()) {
        ++calls;
    }
    const auto& r = solver.result();
    std::printf("solve3d\t%d\t%zu\t%.17g\t%.17g\t%.17g\t%.17g\t%.17g\t%zu\n", calls + 1,
                static_cast<std::size_t>(r.iterations), r.cd, r.cl, r.cy, r.cm, r.force.x, r.surface.cp.size());
    return 0;
}

} // namespace

int main(int argc, char** argv) {
    for (int a = 1; a < argc; ++a) {
        if (std::strcmp(argv[a], "run3d") == 0 && a + 1 < argc) {
            return run3d(argv[a + 1]);
        }
        if (std::strcmp(argv[a], "run2d") == 0 && a + 1 < argc) {
            return run2d(argv[a + 1]);
        }
        if (std::strcmp(argv[a], "solve3d") == 0) {
            return solve3d();
        }
        if (std::strncmp(argv[a], "aero-", 5) == 0) {
            // The IAeroSolver wrapper: a unit cube, a refused flow condition or option.
            orrery::aero::AeroGeometry g;
            g.mesh.vertices = {{-0.5, -0.5, -0.5}, {0.5, -0.5, -0.5}, {-0.5, 0.5, -0.5}, {0.5, 0.5, -0.5},
                               {-0.5, -0.5, 0.5},  {0.5, -0.5, 0.5},  {-0.5, 0.5, 0.5},  {0.5, 0.5, 0.5}};
            g.mesh.triangles = {{0, 4, 6}, {0, 6, 2}, {1, 3, 7}, {1, 7, 5}, {0, 1, 5}, {0, 5, 4},
                                {2, 6, 7}, {2, 7, 3}, {0, 2, 3}, {0, 3, 1}, {4, 5, 7}, {4, 7, 6}};
            g.reference.length = 1.0;
            g.reference.area = 1.0;
            orrery::aero::FlowConditions f;
            f.speed = 0.1;
            f.density = 1000.0;
            f.viscosity = 5.0;
            ref::LbmSolver3DOptions o;
            const std::string what = argv[a] + 5;
            if (what == "speed") {
                f.speed = 0.0;
            } else if (what == "sides") {
                o.sides = ref::WallKind::no_slip;
            } else if (what == "mach") {
                o.resolution.lattice_velocity = 0.06;
            }
            orrery_amd::HipLbmAeroSolver3D solver(o);
            try {
                solver.setup(g, f);
                std::printf("%s\tset up\n", argv[a]);
            } catch (const std::invalid_argument& e) {
                std::printf("%s\t%s\n", argv[a], e.what());
            }
            continue;
        }
        if (std::strncmp(argv[a], "lattice2d-", 10) == 0) {
            // make_hip_solver for the D2Q9 lattice: a 16 x 8 channel, one parameter refused.
            ref::Lbm2DConfig c;
            c.nx = 16;
            c.ny = 8;
            c.tau = 0.6;
            c.inlet_ux = 0.05;
            ref::SolidMask2D m;
            m.nx = 16;
            m.ny = 8;
            m.solid.assign(16 * 8, 0);
            const std::string what = argv[a] + 10;
            if (what == "tau") {
                c.tau = 0.5;
            } else if (what == "mach") {
                c.inlet_ux = 0.2;
            } else if (what == "size") {
                m.nx = 15;
            } else if (what == "periodic") {
                c.top.kind = ref::WallKind::periodic;
            } else if (what == "inlet") {
                m.solid[16 * 3] = 1;
            }
            auto made = orrery_amd::make_hip_solver(c, m);
            std::printf("%s\t%s\n", argv[a], made ? "created" : made.error().c_str());
            continue;
        }
        if (std::strncmp(argv[a], "aero2d-", 7) == 0) {
            // The 2D IAeroSolver wrapper: a 48-gon of diameter 1, a refused flow or option.
            orrery::aero::AeroGeometry g;
            for (int k = 0; k < 48; ++k) {
                const double t = 2.0 * 3.14159265358979323846 * k / 48.0;
                g.section.push_back({0.5 * std::cos(t), 0.5 * std::sin(t)});
            }
            g.reference.length = 1.0;
            g.reference.area = 1.0;
            orrery::aero::FlowConditions f;
            f.speed = 0.1;
            f.density = 1000.0;
            f.viscosity = 1.0;
            ref::LbmSolverOptions o;
            o.resolution.cells_per_length = 10.0;
            const std::string what = argv[a] + 7;
            if (what == "speed") {
                f.speed = 0.0;
            } else if (what == "section") {
                g.section.resize(2);
            } else if (what == "mach") {
                o.resolution.lattice_velocity = 0.08;
            }
            orrery_amd::HipLbmAeroSolver2D solver(o);
            try {
                solver.setup(g, f);
                std::printf("%s\tset up\n", argv[a]);
            } catch (const std::invalid_argument& e) {
                std::printf("%s\t%s\n", argv[a], e.what());
            }
            continue;
        }
        ref::Lbm3DConfig c;
        c.nx = 16;
        c.ny = 8;
        c.nz = 8;
        c.tau = 0.6;
        c.inlet_ux = 0.05;
        ref::SolidMask3D m;
        m.nx = 16;
        m.ny = 8;
        m.nz = 8;
        m.solid.assign(16 * 8 * 8, 0);
        const std::string what = argv[a];
        if (what == "tau") {
            c.tau = 0.5;
        } else if (what == "mach") {
            c.inlet_ux = 0.2;
        } else if (what == "size") {
            m.nx = 15;
        } else if (what == "small") {
            c.nz = 2;
            m.nz = 2;
            m.solid.assign(16 * 8 * 2, 0);
        } else if (what == "distance") {
            c.solid_wall = ref::SolidWall::interpolated;
        } else if (what == "periodic") {
            c.y_low.kind = ref::WallKind::periodic;
        }
        auto made = orrery_amd::make_hip_solver(c, m);
        std::printf("%s\t%s\n", what.c_str(), made ? "created" : made.error().c_str());
    }
    return 0;
}
