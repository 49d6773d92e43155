// Drives orrery_amd::make_hip_solver (include/orrery_lbm_solver.hpp) through the reference's own
// Lbm3DSolver types: prints the refusal message of each configuration named on the command line,
// or "created" when the solver was made. Built by tests/test_adapter.py.
#include <cstdio>
#include <cstring>
#include <string>

#include "orrery_lbm_solver.hpp"

namespace ref = orrery::aero::lbm;

int main(int argc, char** argv) {
    for (int a = 1; a < argc; ++a) {
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
