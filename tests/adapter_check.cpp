// Drives orrery_amd::make_hip_solver (include/orrery_lbm_solver.hpp) through the reference's own
// Lbm3DSolver types: prints the refusal message of each configuration named on the command line,
// or "created" when the solver was made. Built by tests/test_adapter.py.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "orrery/geom/primitives.h"
#include "orrery_lbm_aero.hpp"
#include "orrery_lbm_solver.hpp"

namespace ref = orrery::aero::lbm;

int main(int argc, char** argv) {
    for (int a = 1; a < argc; ++a) {
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
