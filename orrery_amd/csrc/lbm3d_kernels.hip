// D3Q19 kernels: the fused in-place stream-collide step (bulk and boundary halves), state
// initialisation, and the FP64 readbacks and reductions (populations, moments, mass, force).
//
// Physics follows the reference step (src/aero/lbm3d/lbm3d_cpu.cpp:32-100,261-322 and
// src/render/shaders/lbm/lbm_d3q19_common.glsl:111-368): BGK with Guo forcing, the optional
// Smagorinsky closure, absorbing layers, halfway or interpolated (Bouzidi) bounce-back,
// slip / no-slip / moving walls, a regularised velocity inlet and a convective or
// zero-gradient outlet. The binary32 formats collide on populations shifted by their weights
// (the GLSL form); the binary64 format keeps the CPU solver's unshifted arithmetic so the
// on-device FP64 run reproduces it.
#include <hip/hip_runtime.h>

#include "lbm_device.h"
#include "lbm_kernels.h"

namespace lbm {

namespace {

using L3 = D3Q19;

__device__ __forceinline__ int off3(int i, int nx, int plane) {
    return L3::cx(i) + nx * L3::cy(i) + plane * L3::cz(i);
}

// Relaxation rate of the absorbing layers at (x, y, z) (sponge_sigma,
// src/aero/lbm3d/lbm3d_internal.h:256-281).
template <typename R>
__device__ R sponge_sigma3(const Lattice3<R>& p, std::uint32_t x, std::uint32_t y, std::uint32_t z) {
    R best = 0;
    const auto layer = [&best](std::uint32_t d, std::uint32_t t) {
        if (d < t) {
            const R r = static_cast<R>(t - d) / static_cast<R>(t);
            best = best > r * r ? best : r * r;
        }
    };
    layer(x, p.sponge_in);
    layer(p.nx - 1 - x, p.sponge_out);
    if (p.sponge_y) {
        layer(y, p.sponge_sides);
        layer(p.ny - 1 - y, p.sponge_sides);
    }
    if (p.sponge_z) {
        layer(z, p.sponge_sides);
        layer(p.nz - 1 - z, p.sponge_sides);
    }
    return p.sponge_strength * best;
}

// Collision of one node: f holds the post-streaming populations (shifted for binary32
// formats) and is overwritten with the post-collision ones.
template <class T>
__device__ __forceinline__ void collide3(typename T::R (&f)[19], const Lattice3<typename T::R>& p,
                                         typename T::R sponge) {
    using R = typename T::R;
    if constexpr (T::shifted) {
        // lbm_d3q19_common.glsl:264-326 and :358-368 (shifted equilibrium w (drho + rho (...)))
        R drho = 0;
#pragma unroll
        for (int i = 0; i < 19; ++i) {
            drho += f[i];
        }
        const R jx = f[1] - f[2] + f[7] - f[8] + f[9] - f[10] + f[11] - f[12] + f[13] - f[14];
        const R jy = f[3] - f[4] + f[7] - f[8] - f[9] + f[10] + f[15] - f[16] + f[17] - f[18];
        const R jz = f[5] - f[6] + f[11] - f[12] - f[13] + f[14] + f[15] - f[16] - f[17] + f[18];
        const R rho = R(1) + drho;
        const R inv = R(1) / rho;
        const R ux = (jx + R(0.5) * p.force[0]) * inv;
        const R uy = (jy + R(0.5) * p.force[1]) * inv;
        const R uz = (jz + R(0.5) * p.force[2]) * inv;
        const R uu = ux * ux + uy * uy + uz * uz;
        R feq[19];
#pragma unroll
        for (int i = 0; i < 19; ++i) {
            const R cu = R(L3::cx(i)) * ux + R(L3::cy(i)) * uy + R(L3::cz(i)) * uz;
            feq[i] = R(L3::w(i)) * (drho + rho * (R(3) * cu + R(4.5) * cu * cu - R(1.5) * uu));
        }
        R omega = p.omega;
        if (p.les) {
            R pxx = 0, pyy = 0, pzz = 0, pxy = 0, pxz = 0, pyz = 0;
#pragma unroll
            for (int i = 0; i < 19; ++i) {
                const R ex = R(L3::cx(i));
                const R ey = R(L3::cy(i));
                const R ez = R(L3::cz(i));
                const R neq = f[i] - feq[i];
                pxx += ex * ex * neq;
                pyy += ey * ey * neq;
                pzz += ez * ez * neq;
                pxy += ex * ey * neq;
                pxz += ex * ez * neq;
                pyz += ey * ez * neq;
            }
            const R strain =
                sqrt(pxx * pxx + pyy * pyy + pzz * pzz + R(2) * (pxy * pxy + pxz * pxz + pyz * pyz));
            const R tau = R(0.5) * (p.tau + sqrt(p.tau * p.tau + R(25.455844122715710) * p.smag_c2 * strain / rho));
            omega = R(1) / tau;
        }
        const R force_scale = R(1) - R(0.5) * omega;
#pragma unroll
        for (int i = 0; i < 19; ++i) {
            R value = f[i] - omega * (f[i] - feq[i]);
            if (p.forced) {
                const R ex = R(L3::cx(i));
                const R ey = R(L3::cy(i));
                const R ez = R(L3::cz(i));
                const R cu = ex * ux + ey * uy + ez * uz;
                value += force_scale * R(L3::w(i)) *
                         (R(3) * ((ex - ux) * p.force[0] + (ey - uy) * p.force[1] + (ez - uz) * p.force[2]) +
                          R(9) * cu * (ex * p.force[0] + ey * p.force[1] + ez * p.force[2]));
            }
            f[i] = value;
        }
        if (sponge > R(0)) {
            const R vx = p.inlet[0];
            const R vy = p.inlet[1];
            const R vz = p.inlet[2];
            const R vv = vx * vx + vy * vy + vz * vz;
#pragma unroll
            for (int i = 0; i < 19; ++i) {
                const R cu = R(L3::cx(i)) * vx + R(L3::cy(i)) * vy + R(L3::cz(i)) * vz;
                const R target = R(L3::w(i)) * (R(3) * cu + R(4.5) * cu * cu - R(1.5) * vv);
                f[i] -= sponge * (f[i] - target);
            }
        }
    } else {
        // lbm3d_cpu.cpp:32-100, expression for expression (the library is built without
        // floating-point contraction).
        double rho = 0.0;
#pragma unroll
        for (int i = 0; i < 19; ++i) {
            rho += f[i];
        }
        const double jx = f[1] - f[2] + f[7] - f[8] + f[9] - f[10] + f[11] - f[12] + f[13] - f[14];
        const double jy = f[3] - f[4] + f[7] - f[8] - f[9] + f[10] + f[15] - f[16] + f[17] - f[18];
        const double jz = f[5] - f[6] + f[11] - f[12] - f[13] + f[14] + f[15] - f[16] - f[17] + f[18];
        const double fx = p.force[0];
        const double fy = p.force[1];
        const double fz = p.force[2];
        const double ux = (jx + 0.5 * fx) / rho;
        const double uy = (jy + 0.5 * fy) / rho;
        const double uz = (jz + 0.5 * fz) / rho;
        const double base = 1.0 - 1.5 * (ux * ux + uy * uy + uz * uz);
        const double cu[19] = {0.0,     ux,       -ux,      uy,      -uy,     uz,       -uz,
                               ux + uy, -ux - uy, ux - uy,  uy - ux, ux + uz, -ux - uz, ux - uz,
                               uz - ux, uy + uz,  -uy - uz, uy - uz, uz - uy};
        double feq[19];
#pragma unroll
        for (int i = 0; i < 19; ++i) {
            feq[i] = L3::w(i) * rho * (base + cu[i] * (3.0 + 4.5 * cu[i]));
        }
        double tau = p.tau;
        if (p.les) {
            double pi[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
            for (int i = 0; i < 19; ++i) {
                const double neq = f[i] - feq[i];
                const double ex = L3::cx(i);
                const double ey = L3::cy(i);
                const double ez = L3::cz(i);
                pi[0] += ex * ex * neq;
                pi[1] += ey * ey * neq;
                pi[2] += ez * ez * neq;
                pi[3] += ex * ey * neq;
                pi[4] += ex * ez * neq;
                pi[5] += ey * ez * neq;
            }
            const double strain = sqrt(pi[0] * pi[0] + pi[1] * pi[1] + pi[2] * pi[2] +
                                       2.0 * (pi[3] * pi[3] + pi[4] * pi[4] + pi[5] * pi[5]));
            const double c = p.smag_c;
            tau = 0.5 * (tau + sqrt(tau * tau + 18.0 * 1.4142135623730951 * c * c * strain / rho));
        }
        const double omega = 1.0 / tau;
        const double force_scale = 1.0 - 0.5 * omega;
#pragma unroll
        for (int i = 0; i < 19; ++i) {
            double value = f[i] - omega * (f[i] - feq[i]);
            if (p.forced) {
                const double ex = L3::cx(i);
                const double ey = L3::cy(i);
                const double ez = L3::cz(i);
                value += force_scale * L3::w(i) *
                         (3.0 * ((ex - ux) * fx + (ey - uy) * fy + (ez - uz) * fz) +
                          9.0 * cu[i] * (ex * fx + ey * fy + ez * fz));
            }
            f[i] = value;
        }
        if (sponge > 0.0) {
            const double sx = p.inlet[0];
            const double sy = p.inlet[1];
            const double sz = p.inlet[2];
#pragma unroll
            for (int i = 0; i < 19; ++i) {
                const double c = static_cast<double>(L3::cx(i)) * sx + static_cast<double>(L3::cy(i)) * sy +
                                 static_cast<double>(L3::cz(i)) * sz;
                const double vv = sx * sx + sy * sy + sz * sz;
                const double target = L3::w(i) * 1.0 * (1.0 + 3.0 * c + 4.5 * c * c - 1.5 * vv);
                f[i] -= sponge * (f[i] - target);
            }
        }
    }
}

// Regularised velocity inlet on the west face (velocity_inlet_west,
// src/aero/lbm3d/lbm3d_internal.h:324-365; shifted form lbm_d3q19_common.glsl:111-162).
template <class T>
__device__ __forceinline__ void inlet_west3(typename T::R (&f)[19], const Lattice3<typename T::R>& p) {
    using R = typename T::R;
    const R ux = p.inlet[0];
    const R uy = p.inlet[1];
    const R uz = p.inlet[2];
    R known = 0;
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        if (L3::cx(i) == 0) {
            known += f[i];
        } else if (L3::cx(i) < 0) {
            known += R(2) * f[i];
        }
    }
    R rho;
    R drho;
    if constexpr (T::shifted) {
        drho = (known + ux) / (R(1) - ux);
        rho = R(1) + drho;
    } else {
        rho = known / (1.0 - ux);
        drho = rho;
    }
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        if (L3::cx(i) > 0) {
            const R cu = R(L3::cx(i)) * ux + R(L3::cy(i)) * uy + R(L3::cz(i)) * uz;
            f[i] = f[L3::opp(i)] + R(6) * R(L3::w(i)) * rho * cu;
        }
    }
    const R uu = ux * ux + uy * uy + uz * uz;
    R feq[19];
    R pi[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        const R ex = R(L3::cx(i));
        const R ey = R(L3::cy(i));
        const R ez = R(L3::cz(i));
        const R cu = ex * ux + ey * uy + ez * uz;
        if constexpr (T::shifted) {
            feq[i] = R(L3::w(i)) * (drho + rho * (R(3) * cu + R(4.5) * cu * cu - R(1.5) * uu));
        } else {
            feq[i] = L3::w(i) * rho * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * uu);
        }
        const R neq = f[i] - feq[i];
        pi[0] += ex * ex * neq;
        pi[1] += ey * ey * neq;
        pi[2] += ez * ez * neq;
        pi[3] += ex * ey * neq;
        pi[4] += ex * ez * neq;
        pi[5] += ey * ez * neq;
    }
    const R cs2 = R(1.0 / 3.0);
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        const R ex = R(L3::cx(i));
        const R ey = R(L3::cy(i));
        const R ez = R(L3::cz(i));
        f[i] = feq[i] + R(4.5) * R(L3::w(i)) *
                            ((ex * ex - cs2) * pi[0] + (ey * ey - cs2) * pi[1] + (ez * ez - cs2) * pi[2] +
                             R(2) * (ex * ey * pi[3] + ex * ez * pi[4] + ey * ez * pi[5]));
    }
}

// ---- the step: bulk half ---------------------------------------------------------------------
// Every bulk node (all 18 neighbours inside the domain and fluid) streams plainly: 19 loads,
// the collision, 19 stores. kLocal selects the A step (odd n) or the B step (even n).
template <class T, bool kLocal, int CPT>
__global__ __launch_bounds__(256) void d3q19_bulk(Lattice3<typename T::R> p) {
    using S = typename T::S;
    using R = typename T::R;
    S* __restrict__ f = static_cast<S*>(p.f);
    const std::uint32_t stride = p.stride;
    const int nx = static_cast<int>(p.nx);
    const int plane = static_cast<int>(p.nx * p.ny);
    const std::uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        const std::uint32_t cell = (tile * CPT + k) * blockDim.x + threadIdx.x;
        if (cell >= p.cells) {
            return;
        }
        if (p.cls[cell] != 1) {
            continue;
        }
        R v[19];
#pragma unroll
        for (int i = 0; i < 19; ++i) {
            if constexpr (kLocal) {
                v[i] = T::load(f[i * stride + cell]);
            } else {
                v[i] = T::load(f[L3::opp(i) * stride + static_cast<std::uint32_t>(static_cast<int>(cell) - off3(i, nx, plane))]);
            }
        }
        R sponge = 0;
        if (p.sponge_strength > R(0)) {
            const std::uint32_t z = p.div_plane.div(cell);
            const std::uint32_t rest = cell - z * static_cast<std::uint32_t>(plane);
            const std::uint32_t y = p.div_nx.div(rest);
            sponge = sponge_sigma3(p, rest - y * p.nx, y, z);
        }
        collide3<T>(v, p, sponge);
#pragma unroll
        for (int i = 0; i < 19; ++i) {
            if constexpr (kLocal) {
                f[L3::opp(i) * stride + cell] = T::pack(v[i]);
            } else {
                f[i * stride + static_cast<std::uint32_t>(static_cast<int>(cell) + off3(i, nx, plane))] = T::pack(v[i]);
            }
        }
    }
}

// ---- the step: boundary half -----------------------------------------------------------------
// One thread per boundary node. Sources come from the precomputed codes; values other than
// the node's own natural pull come from the stash of the previous step.
template <class T, bool kLocal>
__global__ __launch_bounds__(256) void d3q19_boundary(Lattice3<typename T::R> p) {
    using S = typename T::S;
    using R = typename T::R;
    const std::uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= p.nb) {
        return;
    }
    S* __restrict__ f = static_cast<S*>(p.f);
    const std::uint32_t stride = p.stride;
    const std::uint32_t bs = p.bstride;
    const int nx = static_cast<int>(p.nx);
    const int plane = static_cast<int>(p.nx * p.ny);
    const std::uint32_t cell = p.bcell[b];
    const std::uint32_t z = p.div_plane.div(cell);
    const std::uint32_t rest = cell - z * static_cast<std::uint32_t>(plane);
    const std::uint32_t y = p.div_nx.div(rest);
    const std::uint32_t x = rest - y * p.nx;
    const R* __restrict__ st = p.stash + (kLocal ? 19U * bs : 0U);
    R* __restrict__ sw = p.stash + (kLocal ? 0U : 19U * bs);

    // Natural pull of direction i at this node (its own address in the A-A scheme).
    const auto natural = [&](int i, std::uint32_t src) -> R {
        if constexpr (kLocal) {
            return T::load(f[i * stride + cell]);
        } else {
            return T::load(f[L3::opp(i) * stride + src]);
        }
    };

    R v[19];
    R wall_rho = R(-1);
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        const std::uint32_t code = p.bcode[i * bs + b];
        const int kind = static_cast<int>(code & 7U);
        const std::uint32_t payload = code >> kCodeShift;
        if (kind == kNode) {
            if ((code & (kCodeMirrorY | kCodeMirrorZ)) == 0U) {
                v[i] = natural(i, payload);
            } else {
                int d = i;
                if (code & kCodeMirrorY) d = L3::mirror_y(d);
                if (code & kCodeMirrorZ) d = L3::mirror_z(d);
                v[i] = st[d * bs + payload];
            }
        } else if (kind == kBounce) {
            // reflected (lbm3d_internal.h:185-197)
            const int o = L3::opp(i);
            const R leaving = st[o * bs + b];
            const R q = p.bq != nullptr ? p.bq[i * bs + b] : R(0.5);
            if (q == R(0.5)) {
                v[i] = leaving;
            } else if (q < R(0.5)) {
                const R farther = natural(o, static_cast<std::uint32_t>(static_cast<int>(cell) + off3(i, nx, plane)));
                v[i] = R(2) * q * leaving + (R(1) - R(2) * q) * farther;
            } else {
                v[i] = (leaving + (R(2) * q - R(1)) * st[i * bs + b]) / (R(2) * q);
            }
        } else if (kind == kWall) {
            if (wall_rho < R(0)) {
                R s = 0;
#pragma unroll
                for (int k = 0; k < 19; ++k) {
                    s += st[k * bs + b];
                }
                wall_rho = T::shifted ? R(1) + s : s;
            }
            v[i] = st[L3::opp(i) * bs + b] + R(6) * R(L3::w(i)) * wall_rho * R(L3::cx(i)) * p.wall_u[payload];
        } else if (kind == kInlet) {
            v[i] = R(0);
        } else {
            int d = i;
            if (code & kCodeMirrorY) d = L3::mirror_y(d);
            if (code & kCodeMirrorZ) d = L3::mirror_z(d);
            R up = st[d * bs + payload];
            if (p.convective) {
                R* h = p.hist + L3::outlet_slot(i) * bs + b;
                up = (*h + p.outlet_speed * up) / (R(1) + p.outlet_speed);
                *h = up;
            }
            v[i] = up;
        }
    }
    if (p.x_mode == 1 && x == 0) {
        inlet_west3<T>(v, p);
    }
    const R sponge = p.sponge_strength > R(0) ? sponge_sigma3(p, x, y, z) : R(0);
    collide3<T>(v, p, sponge);
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        const S packed = T::pack(v[i]);
        sw[i * bs + b] = T::load(packed);
        if constexpr (kLocal) {
            f[L3::opp(i) * stride + cell] = packed;
        } else {
            int tx = static_cast<int>(x) + L3::cx(i);
            int ty = static_cast<int>(y) + L3::cy(i);
            int tz = static_cast<int>(z) + L3::cz(i);
            bool inside = true;
            if (ty < 0 || ty >= static_cast<int>(p.ny)) {
                if (p.side[ty < 0 ? 0 : 1] == kPeriodic) {
                    ty = ty < 0 ? ty + static_cast<int>(p.ny) : ty - static_cast<int>(p.ny);
                } else {
                    inside = false;
                }
            }
            if (tz < 0 || tz >= static_cast<int>(p.nz)) {
                if (p.side[tz < 0 ? 2 : 3] == kPeriodic) {
                    tz = tz < 0 ? tz + static_cast<int>(p.nz) : tz - static_cast<int>(p.nz);
                } else {
                    inside = false;
                }
            }
            if (tx < 0 || tx >= nx) {
                if (p.x_mode == 0) {
                    tx = tx < 0 ? tx + nx : tx - nx;
                } else {
                    inside = false;
                }
            }
            if (inside) {
                f[i * stride + static_cast<std::uint32_t>(tx + nx * ty + plane * tz)] = packed;
            }
        }
    }
}

// ---- state initialisation ------------------------------------------------------------------
// Equilibrium of (rho, u) per node in the post-A layout; solids at rest. mode 0: uniform
// state (density 1, velocity u0); 1: moment planes in `planes`; 2: u0 plus the seeded
// perturbation (lbm_perturbation) on fluid nodes.
__device__ __forceinline__ double perturbation(std::uint64_t seed, std::uint64_t cell, std::uint32_t component) {
    std::uint64_t z = seed + 0x9E3779B97F4A7C15ULL * (cell * 4U + component + 1U);
    z = (z ^ (z >> 30U)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27U)) * 0x94D049BB133111EBULL;
    z ^= z >> 31U;
    const double u = static_cast<double>(z >> 11U) * (1.0 / 9007199254740992.0);
    return 2.0 * u - 1.0;
}

template <class T>
__global__ __launch_bounds__(256) void d3q19_init(Lattice3<typename T::R> p, int mode, const double* planes,
                                                  double u0x, double u0y, double u0z, std::uint64_t seed,
                                                  double amplitude) {
    using S = typename T::S;
    const std::uint32_t cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= p.cells) {
        return;
    }
    S* __restrict__ f = static_cast<S*>(p.f);
    double rho = 1.0;
    double ux = 0.0;
    double uy = 0.0;
    double uz = 0.0;
    if (p.cls[cell] != 0) {
        if (mode == 1) {
            rho = planes[cell];
            ux = planes[p.cells + cell];
            uy = planes[2U * p.cells + cell];
            uz = planes[3U * p.cells + cell];
        } else {
            ux = u0x;
            uy = u0y;
            uz = u0z;
            if (mode == 2) {
                ux = u0x + amplitude * perturbation(seed, cell, 0);
                uy = u0y + amplitude * perturbation(seed, cell, 1);
                uz = u0z + amplitude * perturbation(seed, cell, 2);
            }
        }
    }
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        // d3q19::equilibrium (d3q19.h:49-55)
        const double cu = static_cast<double>(L3::cx(i)) * ux + static_cast<double>(L3::cy(i)) * uy +
                          static_cast<double>(L3::cz(i)) * uz;
        const double uu = ux * ux + uy * uy + uz * uz;
        double value = L3::w(i) * rho * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * uu);
        if constexpr (T::shifted) {
            value -= L3::w(i);
        }
        f[L3::opp(i) * p.stride + cell] = T::from_double(value);
    }
}

// Stash and outlet history of the boundary nodes from the freshly initialised lattice.
template <class T>
__global__ __launch_bounds__(256) void d3q19_init_boundary(Lattice3<typename T::R> p) {
    using S = typename T::S;
    const std::uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= p.nb) {
        return;
    }
    const S* __restrict__ f = static_cast<const S*>(p.f);
    const std::uint32_t cell = p.bcell[b];
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        const typename T::R v = T::load(f[L3::opp(i) * p.stride + cell]);
        p.stash[i * p.bstride + b] = v;
        if (L3::cx(i) < 0) {
            p.hist[L3::outlet_slot(i) * p.bstride + b] = v;
        }
    }
}

// ---- readbacks -------------------------------------------------------------------------------
// Post-collision population i of a bulk node in the layout of `phase` = n & 1.
template <class T>
__device__ __forceinline__ typename T::R bulk_value(const Lattice3<typename T::R>& p, const typename T::S* f,
                                                    int i, std::uint32_t cell, int phase) {
    if (phase == 0) {
        return T::load(f[L3::opp(i) * p.stride + cell]);
    }
    const int off = off3(i, static_cast<int>(p.nx), static_cast<int>(p.nx * p.ny));
    return T::load(f[i * p.stride + static_cast<std::uint32_t>(static_cast<int>(cell) + off)]);
}

// Populations with the weights added back, [19][cells] doubles. Pass 1 over cells covers
// solid and bulk nodes; pass 2 over the boundary list fills boundary nodes from the stash.
template <class T>
__global__ __launch_bounds__(256) void d3q19_populations_cells(Lattice3<typename T::R> p, int phase, double* out) {
    const std::uint32_t cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= p.cells) {
        return;
    }
    const std::uint8_t c = p.cls[cell];
    if (c == 2) {
        return;
    }
    const auto* f = static_cast<const typename T::S*>(p.f);
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        double v;
        if (c == 0) {
            v = L3::w(i);
        } else {
            v = static_cast<double>(bulk_value<T>(p, f, i, cell, phase));
            if constexpr (T::shifted) {
                v += L3::w(i);
            }
        }
        out[static_cast<std::size_t>(i) * p.cells + cell] = v;
    }
}

template <class T>
__global__ __launch_bounds__(256) void d3q19_populations_boundary(Lattice3<typename T::R> p, int phase, double* out) {
    const std::uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= p.nb) {
        return;
    }
    const std::uint32_t cell = p.bcell[b];
    const auto* st = p.stash + (phase ? 19U * p.bstride : 0U);
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        double v = static_cast<double>(st[i * p.bstride + b]);
        if constexpr (T::shifted) {
            v += L3::w(i);
        }
        out[static_cast<std::size_t>(i) * p.cells + cell] = v;
    }
}

// Moments (lbm3d_cpu.cpp:154-186) in FP64 from the stored values: rho, and u with the Guo
// half-force correction; solids report rest.
template <class T>
__device__ __forceinline__ void moments_of(const double (&f)[19], const double* force, double& rho, double& ux,
                                           double& uy, double& uz) {
    double s = 0.0;
    double jx = 0.0;
    double jy = 0.0;
    double jz = 0.0;
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        s += f[i];
        jx += static_cast<double>(L3::cx(i)) * f[i];
        jy += static_cast<double>(L3::cy(i)) * f[i];
        jz += static_cast<double>(L3::cz(i)) * f[i];
    }
    rho = T::shifted ? 1.0 + s : s;
    ux = (jx - 0.5 * force[0]) / rho;
    uy = (jy - 0.5 * force[1]) / rho;
    uz = (jz - 0.5 * force[2]) / rho;
}

template <class T>
__global__ __launch_bounds__(256) void d3q19_moments_cells(Lattice3<typename T::R> p, int phase, double fx,
                                                           double fy, double fz, double* out) {
    const std::uint32_t cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= p.cells) {
        return;
    }
    const std::uint8_t c = p.cls[cell];
    if (c == 2) {
        return;
    }
    double rho = 1.0, ux = 0.0, uy = 0.0, uz = 0.0;
    if (c == 1) {
        const auto* f = static_cast<const typename T::S*>(p.f);
        double v[19];
#pragma unroll
        for (int i = 0; i < 19; ++i) {
            v[i] = static_cast<double>(bulk_value<T>(p, f, i, cell, phase));
        }
        const double force[3] = {fx, fy, fz};
        moments_of<T>(v, force, rho, ux, uy, uz);
    }
    out[cell] = rho;
    out[p.cells + cell] = ux;
    out[2U * static_cast<std::size_t>(p.cells) + cell] = uy;
    out[3U * static_cast<std::size_t>(p.cells) + cell] = uz;
}

template <class T>
__global__ __launch_bounds__(256) void d3q19_moments_boundary(Lattice3<typename T::R> p, int phase, double fx,
                                                              double fy, double fz, double* out) {
    const std::uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= p.nb) {
        return;
    }
    const std::uint32_t cell = p.bcell[b];
    const auto* st = p.stash + (phase ? 19U * p.bstride : 0U);
    double v[19];
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        v[i] = static_cast<double>(st[i * p.bstride + b]);
    }
    const double force[3] = {fx, fy, fz};
    double rho, ux, uy, uz;
    moments_of<T>(v, force, rho, ux, uy, uz);
    out[cell] = rho;
    out[p.cells + cell] = ux;
    out[2U * static_cast<std::size_t>(p.cells) + cell] = uy;
    out[3U * static_cast<std::size_t>(p.cells) + cell] = uz;
}

// ---- FP64 reductions ---------------------------------------------------------------------------
// Block sums of K doubles per thread: wave64 shuffles, then the waves of the block through
// LDS, in a fixed order (deterministic for a fixed launch shape).
template <int K>
__device__ __forceinline__ void block_sum(double (&v)[K], double* partial) {
    __shared__ double lds[16][K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
#pragma unroll
        for (int offset = 32; offset > 0; offset >>= 1) {
            v[k] += __shfl_down(v[k], offset, 64);
        }
    }
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            lds[wave][k] = v[k];
        }
    }
    __syncthreads();
    if (threadIdx.x < K) {
        double s = 0.0;
        const int waves = static_cast<int>(blockDim.x >> 6);
        for (int w = 0; w < waves; ++w) {
            s += lds[w][threadIdx.x];
        }
        partial[static_cast<std::size_t>(blockIdx.x) * K + threadIdx.x] = s;
    }
}

// Momentum exchange per link (lbm3d_cpu.cpp:192-203 with reflected, add_link_force and
// link_wall_point): F = c_o (f*_o(x) + f_j(x, t+1)), M = (x - q c_j) x F, in FP64.
template <class T>
__global__ __launch_bounds__(256) void d3q19_force(Lattice3<typename T::R> p, int phase, const std::uint32_t* lcell,
                                                   const std::uint8_t* ldir, const double* lq,
                                                   const std::uint32_t* lslot, std::uint32_t nlinks,
                                                   double* partial) {
    const std::uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    double v[6] = {0, 0, 0, 0, 0, 0};
    if (k < nlinks) {
        const auto* f = static_cast<const typename T::S*>(p.f);
        const std::uint32_t cell = lcell[k];
        const int j = ldir[k];
        const double q = lq[k];
        // The step kernel reflects with q in the arithmetic type; the force uses the same q for
        // the reflected value and the exact one for the wall point.
        const double qr = static_cast<double>(static_cast<typename T::R>(q));
        const std::uint32_t b = lslot[k];
        const auto* st = p.stash + (phase ? 19U * p.bstride : 0U);
        const int o = L3::opp(j);
        const double out = static_cast<double>(st[o * p.bstride + b]) + (T::shifted ? L3::w(o) : 0.0);
        double back;
        if (qr == 0.5) {
            back = out;
        } else if (qr < 0.5) {
            // f*_o(x + c_j): the natural pull of direction o at x in this phase.
            double farther;
            if (phase == 0) {
                const int off = off3(j, static_cast<int>(p.nx), static_cast<int>(p.nx * p.ny));
                farther = static_cast<double>(T::load(f[j * p.stride + static_cast<std::uint32_t>(static_cast<int>(cell) + off)]));
            } else {
                farther = static_cast<double>(T::load(f[o * p.stride + cell]));
            }
            if constexpr (T::shifted) {
                farther += L3::w(o);
            }
            back = 2.0 * qr * out + (1.0 - 2.0 * qr) * farther;
        } else {
            const double own = static_cast<double>(st[j * p.bstride + b]) + (T::shifted ? L3::w(j) : 0.0);
            back = (out + (2.0 * qr - 1.0) * own) / (2.0 * qr);
        }
        const double exchanged = out + back;
        const double fv[3] = {static_cast<double>(L3::cx(o)) * exchanged, static_cast<double>(L3::cy(o)) * exchanged,
                              static_cast<double>(L3::cz(o)) * exchanged};
        const std::uint32_t plane = p.nx * p.ny;
        const std::uint32_t z = cell / plane;
        const std::uint32_t rest = cell % plane;
        const std::uint32_t y = rest / p.nx;
        const std::uint32_t x = rest % p.nx;
        const double r[3] = {static_cast<double>(x) - q * L3::cx(j), static_cast<double>(y) - q * L3::cy(j),
                             static_cast<double>(z) - q * L3::cz(j)};
        v[0] = fv[0];
        v[1] = fv[1];
        v[2] = fv[2];
        v[3] = r[1] * fv[2] - r[2] * fv[1];
        v[4] = r[2] * fv[0] - r[0] * fv[2];
        v[5] = r[0] * fv[1] - r[1] * fv[0];
    }
    block_sum<6>(v, partial);
}

// Mass: pass over cells (bulk nodes) and pass over the boundary list, each to block partials.
template <class T>
__global__ __launch_bounds__(256) void d3q19_mass_cells(Lattice3<typename T::R> p, int phase, double* partial) {
    const std::uint32_t cell = blockIdx.x * blockDim.x + threadIdx.x;
    double v[1] = {0.0};
    if (cell < p.cells && p.cls[cell] == 1) {
        const auto* f = static_cast<const typename T::S*>(p.f);
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < 19; ++i) {
            s += static_cast<double>(bulk_value<T>(p, f, i, cell, phase)) + (T::shifted ? L3::w(i) : 0.0);
        }
        v[0] = s;
    }
    block_sum<1>(v, partial);
}

template <class T>
__global__ __launch_bounds__(256) void d3q19_mass_boundary(Lattice3<typename T::R> p, int phase, double* partial) {
    const std::uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    double v[1] = {0.0};
    if (b < p.nb) {
        const auto* st = p.stash + (phase ? 19U * p.bstride : 0U);
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < 19; ++i) {
            s += static_cast<double>(st[i * p.bstride + b]) + (T::shifted ? L3::w(i) : 0.0);
        }
        v[0] = s;
    }
    block_sum<1>(v, partial);
}

} // namespace

// Final stage of the FP64 reductions: one block sums `n` partials of K values in a fixed
// order (each thread a strided run, then a tree), so the result is reproducible.
template <int K>
__global__ __launch_bounds__(256) void reduce_partials(const double* partial, std::uint32_t n, double* out) {
    __shared__ double lds[256][K];
    double s[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        s[k] = 0.0;
    }
    for (std::uint32_t r = threadIdx.x; r < n; r += blockDim.x) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            s[k] += partial[static_cast<std::size_t>(r) * K + k];
        }
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        lds[threadIdx.x][k] = s[k];
    }
    __syncthreads();
    for (unsigned width = blockDim.x / 2; width > 0; width >>= 1) {
        if (threadIdx.x < width) {
#pragma unroll
            for (int k = 0; k < K; ++k) {
                lds[threadIdx.x][k] += lds[threadIdx.x + width][k];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < K) {
        out[threadIdx.x] = lds[0][threadIdx.x];
    }
}

template __global__ void reduce_partials<1>(const double*, std::uint32_t, double*);
template __global__ void reduce_partials<6>(const double*, std::uint32_t, double*);
template __global__ void reduce_partials<3>(const double*, std::uint32_t, double*);

// ---- launchers ---------------------------------------------------------------------------------
namespace {
inline std::uint32_t blocks_for(std::uint64_t n, std::uint32_t block) {
    return static_cast<std::uint32_t>((n + block - 1) / block);
}

template <class T, bool kLocal>
hipError_t step_bulk(const Lattice3<typename T::R>& p, int block, int cpt, hipStream_t s) {
    const std::uint32_t per = static_cast<std::uint32_t>(block) * static_cast<std::uint32_t>(cpt);
    const std::uint32_t grid = blocks_for(p.cells, per);
    switch (cpt) {
        case 2:
            hipLaunchKernelGGL((d3q19_bulk<T, kLocal, 2>), dim3(grid), dim3(block), 0, s, p);
            break;
        case 4:
            hipLaunchKernelGGL((d3q19_bulk<T, kLocal, 4>), dim3(grid), dim3(block), 0, s, p);
            break;
        default:
            hipLaunchKernelGGL((d3q19_bulk<T, kLocal, 1>), dim3(grid), dim3(block), 0, s, p);
            break;
    }
    return hipGetLastError();
}
} // namespace

template <class T>
hipError_t d3q19_step_bulk(const Lattice3<typename T::R>& p, bool local, int block, int cpt, hipStream_t s) {
    return local ? step_bulk<T, true>(p, block, cpt, s) : step_bulk<T, false>(p, block, cpt, s);
}

template <class T>
hipError_t d3q19_step_boundary(const Lattice3<typename T::R>& p, bool local, hipStream_t s) {
    if (p.nb == 0) {
        return hipSuccess;
    }
    const std::uint32_t grid = blocks_for(p.nb, 256);
    if (local) {
        hipLaunchKernelGGL((d3q19_boundary<T, true>), dim3(grid), dim3(256), 0, s, p);
    } else {
        hipLaunchKernelGGL((d3q19_boundary<T, false>), dim3(grid), dim3(256), 0, s, p);
    }
    return hipGetLastError();
}

template <class T>
hipError_t d3q19_step(const Lattice3<typename T::R>& p, bool local, int block, int cpt, hipStream_t s) {
    const hipError_t e = d3q19_step_bulk<T>(p, local, block, cpt, s);
    if (e != hipSuccess) {
        return e;
    }
    return d3q19_step_boundary<T>(p, local, s);
}

template <class T>
hipError_t d3q19_initialize(const Lattice3<typename T::R>& p, int mode, const double* planes, const double u0[3],
                            std::uint64_t seed, double amplitude, hipStream_t s) {
    hipLaunchKernelGGL((d3q19_init<T>), dim3(blocks_for(p.cells, 256)), dim3(256), 0, s, p, mode, planes, u0[0],
                       u0[1], u0[2], seed, amplitude);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || p.nb == 0) {
        return e;
    }
    hipLaunchKernelGGL((d3q19_init_boundary<T>), dim3(blocks_for(p.nb, 256)), dim3(256), 0, s, p);
    return hipGetLastError();
}

template <class T>
hipError_t d3q19_populations(const Lattice3<typename T::R>& p, int phase, double* out, hipStream_t s) {
    hipLaunchKernelGGL((d3q19_populations_cells<T>), dim3(blocks_for(p.cells, 256)), dim3(256), 0, s, p, phase, out);
    if (p.nb > 0) {
        hipLaunchKernelGGL((d3q19_populations_boundary<T>), dim3(blocks_for(p.nb, 256)), dim3(256), 0, s, p, phase,
                           out);
    }
    return hipGetLastError();
}

template <class T>
hipError_t d3q19_moments(const Lattice3<typename T::R>& p, int phase, const double force[3], double* out,
                         hipStream_t s) {
    hipLaunchKernelGGL((d3q19_moments_cells<T>), dim3(blocks_for(p.cells, 256)), dim3(256), 0, s, p, phase,
                       force[0], force[1], force[2], out);
    if (p.nb > 0) {
        hipLaunchKernelGGL((d3q19_moments_boundary<T>), dim3(blocks_for(p.nb, 256)), dim3(256), 0, s, p, phase,
                           force[0], force[1], force[2], out);
    }
    return hipGetLastError();
}

std::uint32_t reduction_blocks(std::uint64_t n) { return blocks_for(n > 0 ? n : 1, 256); }

template <class T>
hipError_t d3q19_mass(const Lattice3<typename T::R>& p, int phase, double* partial, double* out, hipStream_t s) {
    const std::uint32_t gc = blocks_for(p.cells, 256);
    const std::uint32_t gb = reduction_blocks(p.nb);
    hipLaunchKernelGGL((d3q19_mass_cells<T>), dim3(gc), dim3(256), 0, s, p, phase, partial);
    hipLaunchKernelGGL((d3q19_mass_boundary<T>), dim3(gb), dim3(256), 0, s, p, phase, partial + gc);
    hipLaunchKernelGGL((reduce_partials<1>), dim3(1), dim3(256), 0, s, partial, gc + gb, out);
    return hipGetLastError();
}

template <class T>
hipError_t d3q19_force_sum(const Lattice3<typename T::R>& p, int phase, const std::uint32_t* lcell,
                           const std::uint8_t* ldir, const double* lq, const std::uint32_t* lslot,
                           std::uint32_t nlinks, double* partial, double* out, hipStream_t s) {
    const std::uint32_t g = reduction_blocks(nlinks);
    hipLaunchKernelGGL((d3q19_force<T>), dim3(g), dim3(256), 0, s, p, phase, lcell, ldir, lq, lslot, nlinks,
                       partial);
    hipLaunchKernelGGL((reduce_partials<6>), dim3(1), dim3(256), 0, s, partial, g, out);
    return hipGetLastError();
}

#define LBM_INSTANTIATE3(T)                                                                                  \
    template hipError_t d3q19_step<T>(const Lattice3<T::R>&, bool, int, int, hipStream_t);                   \
    template hipError_t d3q19_step_bulk<T>(const Lattice3<T::R>&, bool, int, int, hipStream_t);              \
    template hipError_t d3q19_step_boundary<T>(const Lattice3<T::R>&, bool, hipStream_t);                    \
    template hipError_t d3q19_initialize<T>(const Lattice3<T::R>&, int, const double*, const double*,       \
                                            std::uint64_t, double, hipStream_t);                             \
    template hipError_t d3q19_populations<T>(const Lattice3<T::R>&, int, double*, hipStream_t);              \
    template hipError_t d3q19_moments<T>(const Lattice3<T::R>&, int, const double*, double*, hipStream_t);   \
    template hipError_t d3q19_mass<T>(const Lattice3<T::R>&, int, double*, double*, hipStream_t);            \
    template hipError_t d3q19_force_sum<T>(const Lattice3<T::R>&, int, const std::uint32_t*,                 \
                                           const std::uint8_t*, const double*, const std::uint32_t*,         \
                                           std::uint32_t, double*, double*, hipStream_t);

LBM_INSTANTIATE3(StoreF32)
LBM_INSTANTIATE3(StoreF16)
LBM_INSTANTIATE3(StoreBF16)
LBM_INSTANTIATE3(StoreF64)

} // namespace lbm
