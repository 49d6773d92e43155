// Device functions of the D3Q19 step, shared by the single-lattice kernels (lbm3d_kernels.hip)
// and the batched kernels (lbm_batch.hip): the collision, the regularised inlet, the absorbing
// layers, and the update of one bulk node and of one boundary node in the A-A scheme of
// lbm_device.h.
//
// Physics follows the reference step (src/aero/lbm3d/lbm3d_cpu.cpp:32-100,261-322 and
// src/render/shaders/lbm/lbm_d3q19_common.glsl:111-368): BGK with Guo forcing, the optional
// Smagorinsky closure, absorbing layers, halfway or interpolated (Bouzidi) bounce-back,
// slip / no-slip / moving walls, a regularised velocity inlet and a convective or
// zero-gradient outlet. The binary32 formats collide on populations shifted by their weights
// (the GLSL form); the binary64 format keeps the CPU solver's unshifted arithmetic so the
// on-device FP64 run reproduces it.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "lbm_device.h"

namespace lbm::d3 {

using L3 = D3Q19;

__device__ __forceinline__ int off3(int i, int nx, int plane) {
    return L3::cx(i) + nx * L3::cy(i) + plane * L3::cz(i);
}

// ---- population addresses (the blocked layout of Lattice3::f) ----------------------------------
// Element of A[dir][c] for the guarded cell index g = c + guard.
template <class R>
__device__ __forceinline__ std::uint32_t guarded_element(const Lattice3<R>& p, std::uint32_t g, std::uint32_t dir) {
    const std::uint32_t k = p.block_shift;
    return (__umul24(g >> k, 19U) + dir) << k | (g & ((1U << k) - 1U));
}
// Element of A[dir][cell + off] (a neighbour of a cell, or any cell within the guards).
template <class R>
__device__ __forceinline__ std::uint32_t pop(const Lattice3<R>& p, int dir, std::uint32_t cell, int off = 0) {
    return guarded_element(p, static_cast<std::uint32_t>(static_cast<int>(cell + p.guard) + off),
                           static_cast<std::uint32_t>(dir));
}

// The blocked layout seen from one wave: a uniform base at block `first` (the wave's first
// active lane's block, moved back by the guard blocks) and per-lane element offsets from it.
// A wave's cells and their neighbours span a few blocks around its own, so the offsets stay far
// below 2^32 bytes and every access is a scalar 64-bit base plus a 32-bit lane offset. Lanes must
// hold increasing cells (the first active lane then has the lowest block).
template <class R> struct WaveBlocks {
    std::uint32_t first;  // uniform block number of the base
    std::uint32_t mask;   // 2^k - 1
    std::uint32_t shift;  // k
    // Element offset of A[dir][c] from the base, g = c + guard.
    __device__ __forceinline__ std::uint32_t offset(std::uint32_t g, std::uint32_t dir) const {
        return (__umul24((g >> shift) - first, 19U) + dir) << shift | (g & mask);
    }
    // The element offset of A[0][c]; direction dir lies dir 2^k further (a uniform step).
    __device__ __forceinline__ std::uint32_t offset0(std::uint32_t g) const {
        return __umul24((g >> shift) - first, 19U) << shift | (g & mask);
    }
};
// The wave's blocks for a lane whose guarded cell is g0 (every lane of the wave active here);
// `back` blocks before the first lane's block (the guard blocks, for shifted addresses).
template <class R>
__device__ __forceinline__ WaveBlocks<R> wave_blocks(const Lattice3<R>& p, std::uint32_t g0, std::uint32_t back) {
    WaveBlocks<R> w;
    w.shift = p.block_shift;
    w.mask = (1U << p.block_shift) - 1U;
    w.first = __builtin_amdgcn_readfirstlane(g0 >> p.block_shift) - back;
    return w;
}

// Entry i of the storage offset table. The table is read-only while kernels run; reading it
// through the constant address space lets a uniform read become a scalar load, where a plain
// global pointer gives a vector load per thread (the kernel's own stores might alias it).
template <class R>
__device__ __forceinline__ R storage_offset(const Lattice3<R>& p, int i) {
    using ConstantR = const __attribute__((address_space(4))) R;
    return ((ConstantR*)p.shift)[i];
}

// A stored population of direction `dir` as the arithmetic value (f - w for the shifted formats),
// and back: formats with `offset` store it less shift[dir] (Lattice3::shift).
template <class T>
__device__ __forceinline__ typename T::R decode(const Lattice3<typename T::R>& p, typename T::S s, int dir) {
    if constexpr (T::offset) {
        return T::load(s, storage_offset(p, dir));
    } else {
        return T::load(s);
    }
}
template <class T>
__device__ __forceinline__ typename T::S encode(const Lattice3<typename T::R>& p, typename T::R v, int dir,
                                                std::uint32_t dither) {
    if constexpr (T::offset) {
        return T::pack(v, storage_offset(p, 19 + dir), dither);
    } else {
        return T::pack(v, dither);
    }
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
        // The collision of lbm_d3q19_common.glsl:264-326 (shifted equilibrium
        // w (drho + rho (3 cu + 4.5 cu^2 - 1.5 u^2)), BGK, Smagorinsky, Guo forcing, sponge
        // :358-368), evaluated for speed: opposite directions i, i + 1 (i odd) share
        // cu^2, so their equilibria are sym +- asym; the non-equilibrium stress is
        // P - rho u u - drho / 3 (the second moment of the equilibrium is exact on D3Q19);
        // reciprocal and square root use the hardware instructions (1 ulp); fused
        // multiply-adds are written out, so every kernel that inlines this evaluates it the
        // same way. The results agree with the reference GPU arithmetic to binary32 rounding.
        R ps[10];
        R pd[10];
#pragma unroll
        for (int k = 1; k < 10; ++k) {
            ps[k] = f[2 * k - 1] + f[2 * k];
            pd[k] = f[2 * k - 1] - f[2 * k];
        }
        // pairs: 1 x, 2 y, 3 z, 4 (+x+y), 5 (+x-y), 6 (+x+z), 7 (+x-z), 8 (+y+z), 9 (+y-z)
        const R drho = f[0] + (((ps[1] + ps[2]) + (ps[3] + ps[4])) + ((ps[5] + ps[6]) + (ps[7] + (ps[8] + ps[9]))));
        const R jx = (pd[1] + pd[4]) + (pd[5] + (pd[6] + pd[7]));
        const R jy = (pd[2] + pd[4]) - pd[5] + (pd[8] + pd[9]);
        const R jz = (pd[3] + pd[6]) - pd[7] + (pd[8] - pd[9]);
        const R rho = R(1) + drho;
        const R inv = __builtin_amdgcn_rcpf(rho);
        const R ux = __builtin_fmaf(R(0.5), p.force[0], jx) * inv;
        const R uy = __builtin_fmaf(R(0.5), p.force[1], jy) * inv;
        const R uz = __builtin_fmaf(R(0.5), p.force[2], jz) * inv;
        const R uu = __builtin_fmaf(ux, ux, __builtin_fmaf(uy, uy, uz * uz));
        R omega = p.omega;
        if (p.les) {
            const R third = drho * R(1.0 / 3.0);
            const R rx = rho * ux;
            const R ry = rho * uy;
            const R rz = rho * uz;
            const R qxx = __builtin_fmaf(-rx, ux, (ps[1] + ps[4]) + (ps[5] + (ps[6] + ps[7])) - third);
            const R qyy = __builtin_fmaf(-ry, uy, (ps[2] + ps[4]) + (ps[5] + (ps[8] + ps[9])) - third);
            const R qzz = __builtin_fmaf(-rz, uz, (ps[3] + ps[6]) + (ps[7] + (ps[8] + ps[9])) - third);
            const R qxy = __builtin_fmaf(-rx, uy, ps[4] - ps[5]);
            const R qxz = __builtin_fmaf(-rx, uz, ps[6] - ps[7]);
            const R qyz = __builtin_fmaf(-ry, uz, ps[8] - ps[9]);
            const R off = __builtin_fmaf(qxy, qxy, __builtin_fmaf(qxz, qxz, qyz * qyz));
            const R strain = __builtin_amdgcn_sqrtf(
                __builtin_fmaf(qxx, qxx, __builtin_fmaf(qyy, qyy, __builtin_fmaf(qzz, qzz, R(2) * off))));
            // 18 sqrt(2) = 25.455844122715710
            const R tau = R(0.5) * (p.tau + __builtin_amdgcn_sqrtf(__builtin_fmaf(
                                                p.tau, p.tau, R(25.455844122715710) * p.smag_c2 * strain * inv)));
            omega = __builtin_amdgcn_rcpf(tau);
        }
        const R base = __builtin_fmaf(R(-1.5) * rho, uu, drho);
        f[0] = __builtin_fmaf(omega, R(1.0 / 3.0) * base - f[0], f[0]);
        const R cu[10] = {R(0), ux, uy, uz, ux + uy, ux - uy, ux + uz, ux - uz, uy + uz, uy - uz};
#pragma unroll
        for (int k = 1; k < 10; ++k) {
            const R w = k < 4 ? R(1.0 / 18.0) : R(1.0 / 36.0);
            const R sym = __builtin_fmaf((R(4.5) * w) * rho * cu[k], cu[k], w * base);
            const R asym = (R(3) * w) * rho * cu[k];
            f[2 * k - 1] = __builtin_fmaf(omega, (sym + asym) - f[2 * k - 1], f[2 * k - 1]);
            f[2 * k] = __builtin_fmaf(omega, (sym - asym) - f[2 * k], f[2 * k]);
        }
        if (p.forced) {
            const R force_scale = R(1) - R(0.5) * omega;
#pragma unroll
            for (int i = 0; i < 19; ++i) {
                const R ex = R(L3::cx(i));
                const R ey = R(L3::cy(i));
                const R ez = R(L3::cz(i));
                const R c = ex * ux + ey * uy + ez * uz;
                f[i] += force_scale * R(L3::w(i)) *
                        (R(3) * ((ex - ux) * p.force[0] + (ey - uy) * p.force[1] + (ez - uz) * p.force[2]) +
                         R(9) * c * (ex * p.force[0] + ey * p.force[1] + ez * p.force[2]));
            }
        }
        if (sponge > R(0)) {
            const R vx = p.inlet[0];
            const R vy = p.inlet[1];
            const R vz = p.inlet[2];
            const R vv = vx * vx + vy * vy + vz * vz;
#pragma unroll
            for (int i = 0; i < 19; ++i) {
                const R c = R(L3::cx(i)) * vx + R(L3::cy(i)) * vy + R(L3::cz(i)) * vz;
                const R target = R(L3::w(i)) * (R(3) * c + R(4.5) * c * c - R(1.5) * vv);
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
    // the equilibrium is evaluated twice, identically, instead of being kept: 19 fewer live
    // registers where the bulk kernel inlines this for its inlet nodes
    const auto feq = [&](int i) -> R {
        const R cu = R(L3::cx(i)) * ux + R(L3::cy(i)) * uy + R(L3::cz(i)) * uz;
        if constexpr (T::shifted) {
            return R(L3::w(i)) * (drho + rho * (R(3) * cu + R(4.5) * cu * cu - R(1.5) * uu));
        } else {
            return L3::w(i) * rho * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * uu);
        }
    };
    R pi[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        const R ex = R(L3::cx(i));
        const R ey = R(L3::cy(i));
        const R ez = R(L3::cz(i));
        const R neq = f[i] - feq(i);
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
        f[i] = feq(i) + R(4.5) * R(L3::w(i)) *
                            ((ex * ex - cs2) * pi[0] + (ey * ey - cs2) * pi[1] + (ez * ez - cs2) * pi[2] +
                             R(2) * (ex * ey * pi[3] + ex * ez * pi[4] + ey * ez * pi[5]));
    }
}

// ---- one node of the step: bulk half ---------------------------------------------------------------------
// Every bulk node (all 18 neighbours inside the domain and fluid) streams plainly: 19 loads,
// the collision, 19 stores. kLocal selects the A step (odd n) or the B step (even n).
//
// The loads are split from the rest so that a kernel can issue them before it knows the
// node's class (the class byte and the 19 populations then travel together instead of one
// after the other). Any cell index below `cells` gives in-range addresses: the guard cells
// around the lattice keep the B-step sources of a non-bulk node, whose values are discarded,
// inside the allocation.
template <class T, bool kLocal>
__device__ __forceinline__ void bulk_load(const Lattice3<typename T::R>& p, std::uint32_t cell,
                                          typename T::R (&v)[19]) {
    using S = typename T::S;
    const S* __restrict__ f = static_cast<const S*>(p.f);
    const int nx = static_cast<int>(p.nx);
    const int plane = static_cast<int>(p.nx * p.ny);
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        v[i] = decode<T>(p, kLocal ? f[pop(p, i, cell)] : f[pop(p, L3::opp(i), cell, -off3(i, nx, plane))], i);
    }
}

template <class T, bool kLocal>
__device__ __forceinline__ void bulk_finish(const Lattice3<typename T::R>& p, std::uint32_t cell,
                                            typename T::R (&v)[19]) {
    using S = typename T::S;
    using R = typename T::R;
    S* __restrict__ f = static_cast<S*>(p.f);
    const int nx = static_cast<int>(p.nx);
    const int plane = static_cast<int>(p.nx * p.ny);
    R sponge = 0;
    if (p.sponge_strength > R(0)) {
        const std::uint32_t z = p.div_plane.div(cell);
        const std::uint32_t rest = cell - z * static_cast<std::uint32_t>(plane);
        const std::uint32_t y = p.div_nx.div(rest);
        sponge = sponge_sigma3(p, rest - y * p.nx, y, z);
    }
    collide3<T>(v, p, sponge);
    const std::uint32_t bits = T::dithered ? cell_dither(cell, slot_dither(*p.clock)) : 0U;
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        const S packed = encode<T>(p, v[i], i, direction_dither<T>(bits, i));
        if constexpr (kLocal) {
            f[pop(p, L3::opp(i), cell)] = packed;
        } else {
            f[pop(p, i, cell, off3(i, nx, plane))] = packed;
        }
    }
}

template <class T, bool kLocal>
__device__ __forceinline__ void bulk_node(const Lattice3<typename T::R>& p, std::uint32_t cell) {
    typename T::R v[19];
    bulk_load<T, kLocal>(p, cell, v);
    bulk_finish<T, kLocal>(p, cell, v);
}

// One of four constants by two flags, formed arithmetically: the compiler turns a select among
// constants into a table in memory, one dependent load per use.
__device__ __forceinline__ int pick4(int none, int y, int z, int yz, bool my, bool mz) {
    const int ky = my ? 1 : 0;
    const int kz = mz ? 1 : 0;
    return none + ky * (y - none) + kz * (z - none) + ky * kz * (yz - y - z + none);
}

// ---- side nodes ---------------------------------------------------------------------------------
// A side node streams in place like a bulk node. In the A step it reads A[i][x] and writes
// A[opp(i)][x], its own cell. In the B step pull i comes from A[opp(d)][s], where (s, d) is the
// source node and direction through the sides (resolve_source: a slip side keeps the row and
// mirrors the direction, a periodic side wraps), and output opp(i) is written to that same
// address. A population leaving through a slip side is thereby stored where its mirrored reader
// loads it in the next A step, a slot that no node streams into across the face; and the B-step
// reads and writes of a node are again one set of addresses that no other node touches. So side
// nodes need neither the stash nor source codes.

// Faces a side node lies on: bit 0 y = 0, bit 1 y = ny - 1, bit 2 z = 0, bit 3 z = nz - 1.
template <class R> __device__ __forceinline__ std::uint32_t side_faces(const Lattice3<R>& p, std::uint32_t cell) {
    const std::uint32_t z = p.div_plane.div(cell);
    const std::uint32_t y = p.div_nx.div(cell - z * (p.nx * p.ny));
    return (y == 0U ? 1U : 0U) | (y + 1U == p.ny ? 2U : 0U) | (z == 0U ? 4U : 0U) | (z + 1U == p.nz ? 8U : 0U);
}

// Element of the B-step address of pull i of node `cell` lying on `faces` (0: the plain address of
// a bulk node). Selects only: the result is formed without branches, so a run's loads issue
// together.
template <class R>
__device__ __forceinline__ std::uint32_t side_element(const Lattice3<R>& p, std::uint32_t cell, std::uint32_t faces, int i) {
    const int cx = L3::cx(i);
    const int cy = L3::cy(i);
    const int cz = L3::cz(i);
    const std::uint32_t ybit = cy == 1 ? 1U : (cy == -1 ? 2U : 0U);
    const std::uint32_t zbit = cz == 1 ? 4U : (cz == -1 ? 8U : 0U);
    const bool ycross = (faces & ybit) != 0U;
    const bool zcross = (faces & zbit) != 0U;
    const bool yslip = cy != 0 && p.side[cy == 1 ? 0 : 1] == kSlip;
    const bool zslip = cz != 0 && p.side[cz == 1 ? 2 : 3] == kSlip;
    const bool my = ycross && yslip;
    const bool mz = zcross && zslip;
    // source row and layer offsets: through a slip side the row is kept, through a periodic
    // side it wraps to the opposite face
    const int dy = ycross ? (yslip ? 0 : cy * (static_cast<int>(p.ny) - 1)) : -cy;
    const int dz = zcross ? (zslip ? 0 : cz * (static_cast<int>(p.nz) - 1)) : -cz;
    const int plane = pick4(L3::opp(i), L3::opp(L3::mirror_y(i)), L3::opp(L3::mirror_z(i)),
                            L3::opp(L3::mirror_z(L3::mirror_y(i))), my, mz);
    return pop(p, plane, cell, -cx + static_cast<int>(p.nx) * dy + static_cast<int>(p.nx * p.ny) * dz);
}

// The B step of one side node (the side kernel's threads, the fused schedule's lanes).
template <class T> __device__ __forceinline__ void side_node_b(const Lattice3<typename T::R>& p, std::uint32_t cell) {
    using S = typename T::S;
    using R = typename T::R;
    S* __restrict__ f = static_cast<S*>(p.f);
    const std::uint32_t faces = side_faces(p, cell);
    R v[19];
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        v[i] = decode<T>(p, f[side_element(p, cell, faces, i)], i);
    }
    R sponge = 0;
    if (p.sponge_strength > R(0)) {
        const std::uint32_t z = p.div_plane.div(cell);
        const std::uint32_t rest = cell - z * (p.nx * p.ny);
        const std::uint32_t y = p.div_nx.div(rest);
        sponge = sponge_sigma3(p, rest - y * p.nx, y, z);
    }
    collide3<T>(v, p, sponge);
    const std::uint32_t bits = T::dithered ? cell_dither(cell, slot_dither(*p.clock)) : 0U;
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        f[side_element(p, cell, faces, L3::opp(i))] = encode<T>(p, v[i], i, direction_dither<T>(bits, i));
    }
}

// Nodes updated in place by the bulk kernel (side nodes in A steps only; the side kernel updates
// them in B steps) and read through the lattice by their neighbours.
__device__ __forceinline__ bool streamed(std::uint8_t kind) {
    return kind == kBulkNode || kind == kSideNode || kind == kInletNode;
}
// Nodes the bulk kernel updates in B steps: their B-step addresses are the plain ones.
__device__ __forceinline__ bool plain_b(std::uint8_t kind) { return kind == kBulkNode || kind == kInletNode; }

// V consecutive stored populations of direction i (one block), as one memory access.
template <class T, int V>
__device__ __forceinline__ void load_run(const Lattice3<typename T::R>& p, const typename T::S* src, int i,
                                         typename T::R (&out)[V]) {
    typename T::S tmp[V];
    __builtin_memcpy(tmp, src, sizeof(tmp));
#pragma unroll
    for (int k = 0; k < V; ++k) {
        out[k] = decode<T>(p, tmp[k], i);
    }
}

// Direction i of V consecutive cells, whose dither bits are `bits`, as stored. The 16-bit
// formats round two cells into one 32-bit word (the second conversion writes the word's high
// half), so that a run's words need no packing.
template <class T, int V>
__device__ __forceinline__ void encode_run(const Lattice3<typename T::R>& p, const typename T::R (&in)[V],
                                           const std::uint32_t (&bits)[V], int i, typename T::S (&out)[V]) {
    if constexpr (sizeof(typename T::S) == 2 && V % 2 == 0) {
        std::uint32_t words[V / 2];
        const typename T::R scaled = storage_offset(p, 19 + i);
#pragma unroll
        for (int k = 0; k < V / 2; ++k) {
            words[k] = T::pack_pair(in[2 * k], in[2 * k + 1], scaled, direction_dither<T>(bits[2 * k], i),
                                    direction_dither<T>(bits[2 * k + 1], i));
        }
        __builtin_memcpy(out, words, sizeof(words));
    } else {
#pragma unroll
        for (int k = 0; k < V; ++k) {
            out[k] = encode<T>(p, in[k], i, direction_dither<T>(bits[k], i));
        }
    }
}

// Direction i of V consecutive cells, whose dither bits are `bits`.
template <class T, int V>
__device__ __forceinline__ void store_run(const Lattice3<typename T::R>& p, typename T::S* dst, const typename T::R (&in)[V],
                                          const std::uint32_t (&bits)[V], int i) {
    typename T::S tmp[V];
    encode_run<T, V>(p, in, bits, i, tmp);
    __builtin_memcpy(dst, tmp, sizeof(tmp));
}

// Element `index` of the populations counted from a wave's base: a uniform 64-bit base plus a
// 32-bit byte offset, the addressing of the scalar-base memory instructions (one offset register
// serves every direction of an A step).
template <class S> __device__ __forceinline__ S* element(S* base, std::uint32_t index) {
    using Byte = std::conditional_t<std::is_const_v<S>, const char, char>;
    return reinterpret_cast<S*>(reinterpret_cast<Byte*>(base) + index * static_cast<std::uint32_t>(sizeof(S)));
}

// A run of V cells displaced by a B-step shift starts at guarded cell g and may cross into the
// next block (the last V - 1 cells of a block), where its cells are no longer adjacent: such a
// lane moves them one by one.
template <class R, int V>
__device__ __forceinline__ bool crosses_block(const WaveBlocks<R>& w, std::uint32_t g) {
    return V > 1 && (g & w.mask) > w.mask + 1U - static_cast<std::uint32_t>(V);
}
// Whether any run of a wave crosses under the shift d (cells added to each run's start), from
// uniform values only. A wave's lanes hold consecutive runs starting at multiples of V, 64 V
// cells aligned to 64 V within one block (`group`: their position in the block rounded down).
// A run starting at g crosses iff (g + d) mod 2^k > 2^k - V; with g a multiple of V that needs d
// not a multiple of V and then fixes g mod 2^k to one value, 2^k - V - (d rounded down to a
// multiple of V).
template <class R, int V>
__device__ __forceinline__ bool wave_crosses(const WaveBlocks<R>& w, std::uint32_t group, int d) {
    if (V == 1 || (d & (V - 1)) == 0) {
        return false;
    }
    const std::uint32_t start = static_cast<std::uint32_t>(-V - (d & ~(V - 1))) & w.mask;
    return (start & ~(64U * V - 1U)) == group;
}

// Whether the wave's 64 V cells (`group`: their position in the block), displaced by the shift
// d, all stay inside one block, from uniform values only.
template <class R, int V>
__device__ __forceinline__ bool shift_inside(const WaveBlocks<R>& w, std::uint32_t group, int d) {
    return ((group + static_cast<std::uint32_t>(d)) & w.mask) <= w.mask + 1U - 64U * V;
}
// For such a shift, the element offset of A[dir][g + d] less the lane's own offset0(g), the
// same for every lane of the wave: the shift moves q = floor((group + d) / 2^k) blocks (q < 0:
// back), that is 19 q block lengths of the layout, of which q 2^k cells are covered by d itself.
// Unsigned arithmetic: the sum with offset0(g) is the offset modulo 2^32, which is exact.
template <class R>
__device__ __forceinline__ std::uint32_t shift_term(const WaveBlocks<R>& w, std::uint32_t group, int d, int dir) {
    const int q = (static_cast<int>(group) + d) >> w.shift;
    return (static_cast<std::uint32_t>(q * 18 + dir) << w.shift) + static_cast<std::uint32_t>(d);
}

// The B-step loads of a run of V cells, direction by direction: every direction's combined
// V-wide load (a crossing run's reads past its block's direction, inside the buffer: the guard
// holds V cells more than any shift), then, only in a wave that has a crossing run, the cells
// of the crossing runs one by one.
template <class T, int V>
__device__ __forceinline__ void load_shifted_all(const WaveBlocks<typename T::R>& w, const typename T::S* base,
                                                 std::uint32_t g0, std::uint32_t group, int nx, int plane,
                                                 typename T::S (&raw)[19][V]) {
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        const std::uint32_t g = static_cast<std::uint32_t>(static_cast<int>(g0) - off3(i, nx, plane));
        __builtin_memcpy(raw[i], element(base, w.offset(g, static_cast<std::uint32_t>(L3::opp(i)))), sizeof(raw[i]));
    }
    if constexpr (V > 1) {
        bool fix = false;
#pragma unroll
        for (int i = 0; i < 19; ++i) {
            fix = fix || wave_crosses<typename T::R, V>(w, group, -off3(i, nx, plane));
        }
        if (fix) {
#pragma unroll
            for (int i = 0; i < 19; ++i) {
                const std::uint32_t g = static_cast<std::uint32_t>(static_cast<int>(g0) - off3(i, nx, plane));
                if (crosses_block<typename T::R, V>(w, g)) {
#pragma unroll
                    for (int k = 1; k < V; ++k) {
                        raw[i][k] = *element(base, w.offset(g + static_cast<std::uint32_t>(k),
                                                            static_cast<std::uint32_t>(L3::opp(i))));
                    }
                }
            }
        }
    }
}

template <class T, int V>
__device__ __forceinline__ void store_shifted(const Lattice3<typename T::R>& p, const WaveBlocks<typename T::R>& w,
                                              typename T::S* base, std::uint32_t g, bool careful, int i,
                                              const typename T::R (&in)[V], const std::uint32_t (&bits)[V]) {
    typename T::S tmp[V];
    encode_run<T, V>(p, in, bits, i, tmp);
    if (careful && crosses_block<typename T::R, V>(w, g)) {
#pragma unroll
        for (int k = 0; k < V; ++k) {
            *element(base, w.offset(g + static_cast<std::uint32_t>(k), static_cast<std::uint32_t>(i))) = tmp[k];
        }
    } else {
        __builtin_memcpy(element(base, w.offset(g, static_cast<std::uint32_t>(i))), tmp, sizeof(tmp));
    }
}

// The bulk half of a step for the run of V cells starting at c0: each thread takes V
// consecutive cells and moves each population of the run as one V-wide access (19 loads and 19
// stores per thread whatever V is). A run whose cells are all updated stores whole runs; a
// mixed run stores its updated cells one by one, because the other cells' addresses belong to
// nodes the boundary kernel updates. The lanes of a wave take increasing c0, 64 V consecutive
// cells aligned to 64 V (the callers' blocks of threads are whole waves: multiples of 64, which
// lbm3d_set_launch and lbm_batch_set_launch enforce); the wave addressing and the per-wave test
// of block-edge runs rely on it.
template <class T, bool kLocal, int V>
__device__ __forceinline__ void bulk_run(const Lattice3<typename T::R>& p, std::uint32_t c0) {
    using S = typename T::S;
    using R = typename T::R;
    if (c0 >= p.cells) {
        return;
    }
    // The class array is padded with solid (0) entries beyond the last cell. The A step
    // updates bulk, side and inlet nodes (a side node's A step is a bulk node's); the B step bulk
    // and inlet nodes, side nodes having a kernel of their own then (side_node_b).
    std::uint8_t kind[V];
    __builtin_memcpy(kind, p.cls + c0, sizeof(kind));
    bool any = false;
    bool all = true;
#pragma unroll
    for (int c = 0; c < V; ++c) {
        const bool mine = kLocal ? streamed(kind[c]) : plain_b(kind[c]);
        any = any || mine;
        all = all && mine;
    }
    // A run's population loads either wait for its classes (a run without an update is skipped)
    // or are issued with them (the run then loads and stores nothing it does not update): the
    // second measured faster for the A step of the 4- and 8-byte formats and the B step of the
    // 16-bit ones (512x256x256: f32 A 915 -> 902 us, f16s B 565 -> 525 us), the first for the
    // other two (f32 B 975 against 1006 us under the tuned occupancy caps, f16s A 452 against 466).
    constexpr bool kLoadsWithClasses = sizeof(S) == 2 ? !kLocal : kLocal;
    if (!kLoadsWithClasses && !any) {
        return;
    }
    S* __restrict__ f = static_cast<S*>(p.f);
    const int nx = static_cast<int>(p.nx);
    const int plane = static_cast<int>(p.nx * p.ny);
    const std::uint32_t g0 = c0 + p.guard;
    // B-step sources and targets of a run reach nx ny + nx + 1 cells either way (and V - 1
    // further): the guard blocks behind the wave's first block keep their offsets positive.
    const WaveBlocks<R> w = wave_blocks(p, g0, kLocal ? 0U : p.guard >> p.block_shift);
    S* __restrict__ base = f + (static_cast<std::size_t>(w.first) * 19U << w.shift);
    const std::uint32_t own = w.offset0(g0);
    // the wave's runs lie in one 64 V-cell group of a block (blocks of threads are whole waves)
    const std::uint32_t group = __builtin_amdgcn_readfirstlane(g0) & w.mask & ~(64U * V - 1U);
    // B step: where the wave's cells stay inside one block under every direction's shift (the
    // load shifts -off3(i) are the store shifts off3(opp(i))), each lane's offsets are its own
    // offset plus uniform terms; other waves form them lane by lane.
    bool inside = true;
    if constexpr (!kLocal) {
#pragma unroll
        for (int i = 1; i < 19; ++i) {
            inside = inside && shift_inside<R, V>(w, group, -off3(i, nx, plane));
        }
    }
    R v[V][19];
    if (kLocal || inside) {
#pragma unroll
        for (int i = 0; i < 19; ++i) {
            R run[V];
            if constexpr (kLocal) {
                load_run<T, V>(p, element(base + (static_cast<std::size_t>(i) << w.shift), own), i, run);
            } else {
                load_run<T, V>(p, element(base, own + shift_term(w, group, -off3(i, nx, plane), L3::opp(i))), i, run);
            }
#pragma unroll
            for (int c = 0; c < V; ++c) {
                v[c][i] = run[c];
            }
        }
    } else if constexpr (V == 1) {
#pragma unroll
        for (int i = 0; i < 19; ++i) {
            R run[V];
            const std::uint32_t g = static_cast<std::uint32_t>(static_cast<int>(g0) - off3(i, nx, plane));
            load_run<T, V>(p, element(base, w.offset(g, static_cast<std::uint32_t>(L3::opp(i)))), i, run);
            v[0][i] = run[0];
        }
    } else {
        S raw[19][V];
        load_shifted_all<T, V>(w, base, g0, group, nx, plane, raw);
#pragma unroll
        for (int i = 0; i < 19; ++i) {
#pragma unroll
            for (int c = 0; c < V; ++c) {
                v[c][i] = decode<T>(p, raw[i][c], i);
            }
        }
    }
    const auto updated = [&](int c) { return kLocal ? streamed(kind[c]) : plain_b(kind[c]); };
#pragma unroll
    for (int c = 0; c < V; ++c) {
        if (updated(c)) {
            R sponge = 0;
            if (p.sponge_strength > R(0)) {
                const std::uint32_t cell = c0 + c;
                const std::uint32_t z = p.div_plane.div(cell);
                const std::uint32_t rest = cell - z * static_cast<std::uint32_t>(plane);
                const std::uint32_t y = p.div_nx.div(rest);
                sponge = sponge_sigma3(p, rest - y * p.nx, y, z);
            }
            if (kind[c] == kInletNode) {
                // the pulls from x = -1 (loaded from unowned addresses) are replaced by the
                // regularised inlet, as the boundary kernel does for a listed inlet node
                inlet_west3<T>(v[c], p);
            }
            collide3<T>(v[c], p, sponge);
        }
    }
    std::uint32_t bits[V];
    {
        const std::uint32_t epoch = T::dithered ? slot_dither(*p.clock) : 0U;
#pragma unroll
        for (int c = 0; c < V; ++c) {
            bits[c] = T::dithered ? cell_dither(c0 + static_cast<std::uint32_t>(c), epoch) : 0U;
        }
    }
    if (all && (kLocal || inside)) {
#pragma unroll
        for (int i = 0; i < 19; ++i) {
            R run[V];
#pragma unroll
            for (int c = 0; c < V; ++c) {
                run[c] = v[c][i];
            }
            if constexpr (kLocal) {
                store_run<T, V>(p, element(base + (static_cast<std::size_t>(L3::opp(i)) << w.shift), own), run, bits, i);
            } else {
                store_run<T, V>(p, element(base, own + shift_term(w, group, off3(i, nx, plane), i)), run, bits, i);
            }
        }
    } else if (all) {
        bool careful = false;  // a run of the wave crosses a block in some direction
#pragma unroll
        for (int i = 0; i < 19; ++i) {
            careful = careful || wave_crosses<R, V>(w, group, off3(i, nx, plane));
        }
#pragma unroll
        for (int i = 0; i < 19; ++i) {
            R run[V];
#pragma unroll
            for (int c = 0; c < V; ++c) {
                run[c] = v[c][i];
            }
            store_shifted<T, V>(p, w, base, static_cast<std::uint32_t>(static_cast<int>(g0) + off3(i, nx, plane)),
                                careful, i, run, bits);
        }
    } else {
#pragma unroll
        for (int c = 0; c < V; ++c) {
            if (updated(c)) {
                const std::uint32_t g = g0 + static_cast<std::uint32_t>(c);
#pragma unroll
                for (int i = 0; i < 19; ++i) {
                    const S packed = encode<T>(p, v[c][i], i, direction_dither<T>(bits[c], i));
                    if constexpr (kLocal) {
                        *element(base, w.offset(g, static_cast<std::uint32_t>(L3::opp(i)))) = packed;
                    } else {
                        *element(base, w.offset(static_cast<std::uint32_t>(static_cast<int>(g) + off3(i, nx, plane)),
                                                static_cast<std::uint32_t>(i))) = packed;
                    }
                }
            }
        }
    }
}

// Direction i after the mirrors a source code asks for, and its opposite (i is a constant of an
// unrolled loop).
__device__ __forceinline__ std::uint32_t mirrored(int i, std::uint32_t code) {
    return static_cast<std::uint32_t>(pick4(i, L3::mirror_y(i), L3::mirror_z(i), L3::mirror_z(L3::mirror_y(i)),
                                            (code & kCodeMirrorY) != 0U, (code & kCodeMirrorZ) != 0U));
}
// A stored population read as pull i through the mirrors of `code`: it was stored as the
// source's output of direction mirrored(i, code), so that direction's offset applies (the four
// offsets are selected, not indexed, so that no lane indexes the descriptor).
template <class T>
__device__ __forceinline__ typename T::R decode_mirrored(const Lattice3<typename T::R>& p, typename T::S s, int i,
                                                         std::uint32_t code) {
    if constexpr (T::offset) {
        const bool my = (code & kCodeMirrorY) != 0U;
        const bool mz = (code & kCodeMirrorZ) != 0U;
        const typename T::R shift =
            my ? (mz ? storage_offset(p, L3::mirror_z(L3::mirror_y(i))) : storage_offset(p, L3::mirror_y(i)))
               : (mz ? storage_offset(p, L3::mirror_z(i)) : storage_offset(p, i));
        return T::load(s, shift);
    } else {
        return T::load(s);
    }
}

__device__ __forceinline__ std::uint32_t mirrored_opp(int i, std::uint32_t code) {
    return static_cast<std::uint32_t>(pick4(L3::opp(i), L3::opp(L3::mirror_y(i)), L3::opp(L3::mirror_z(i)),
                                            L3::opp(L3::mirror_z(L3::mirror_y(i))), (code & kCodeMirrorY) != 0U,
                                            (code & kCodeMirrorZ) != 0U));
}

// ---- one node of the step: boundary half -----------------------------------------------------------------
// One boundary node. Sources come from the precomputed codes; values other than the natural
// pulls from bulk nodes come from the stash of the previous step (a boundary node's stash slot
// holds exactly what it stored into the lattice, so a boundary source is read there, densely,
// instead of from a row-scattered lattice address). The node stores into the lattice only the
// outputs a lattice reader will load (bwrite).
template <class T, bool kLocal>
__device__ __forceinline__ void boundary_node(const Lattice3<typename T::R>& p, std::uint32_t b) {
    using S = typename T::S;
    using R = typename T::R;
    S* __restrict__ f = static_cast<S*>(p.f);
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

    // Every value a rule needs is loaded first, each load depending only on the codes: per
    // direction the population from f (the natural pull, or the farther node of a short
    // bounce-back link), another node's stash value (mirrored and outlet sources), the wall
    // fraction and the outlet history, and the node's own stash slot when a bounce-back or wall
    // rule reads it. Loads a direction's rule does not need are not issued. The rules below then
    // compute exactly what they computed with their loads in line.
    std::uint32_t code[19];
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        code[i] = p.bcode[i * bs + b];
    }
    bool needs_own = false;
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        const int kind = static_cast<int>(code[i] & 7U);
        needs_own = needs_own || kind == kBounce || kind == kWall;
    }
    R own[19];
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        own[i] = needs_own ? st[i * bs + b] : R(0);
    }
    S fv[19];
    R sv[19];
    R q[19];
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        const int kind = static_cast<int>(code[i] & 7U);
        const std::uint32_t payload = code[i] >> kCodeShift;
        const bool far = kind == kBounce;  // the farther node of a short link (used when q < 1/2)
        const bool pull = kind == kNode;   // from a bulk or side node, through the lattice
        if constexpr (kLocal) {
            // A step: pulls are the node's own addresses, direction i or opposite(i) (a source
            // beyond a slip side stored its mirrored output here too)
            fv[i] = (pull || far) ? f[pop(p, far ? L3::opp(i) : i, cell)] : S(0);
        } else {
            // B step: A[opp(d)][s] for the source node s and its direction d after the mirrors
            const std::uint32_t src = far ? static_cast<std::uint32_t>(static_cast<int>(cell) + off3(i, nx, plane)) : payload;
            const std::uint32_t at = far ? static_cast<std::uint32_t>(i) : mirrored_opp(i, code[i]);
            fv[i] = (pull || far) ? f[pop(p, static_cast<int>(at), src)] : S(0);
        }
        const bool other = kind == kOutlet || kind == kStash;
        sv[i] = other ? st[mirrored(i, code[i]) * bs + payload] : R(0);
        q[i] = (far && p.bq != nullptr) ? p.bq[i * bs + b] : R(0.5);
    }
    R hist[19];
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        hist[i] = (L3::cx(i) < 0 && p.convective && static_cast<int>(code[i] & 7U) == kOutlet)
                      ? p.hist[L3::outlet_slot(i) * bs + b]
                      : R(0);
    }
    R wall_rho = 0;
    if (needs_own) {
        R s = 0;
#pragma unroll
        for (int k = 0; k < 19; ++k) {
            s += own[k];
        }
        wall_rho = T::shifted ? R(1) + s : s;
    }

    R v[19];
    // The convective outlet's new history values are written with the other stores after the
    // collision: a store among the rules would make every later wait for a load wait for it too
    // (one counter covers loads and stores).
    R hist_new[19];
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        const int kind = static_cast<int>(code[i] & 7U);
        const std::uint32_t payload = code[i] >> kCodeShift;
        hist_new[i] = R(0);
        if (kind == kNode) {
            v[i] = decode_mirrored<T>(p, fv[i], i, code[i]);
        } else if (kind == kStash) {
            v[i] = sv[i];
        } else if (kind == kBounce) {
            // reflected (lbm3d_internal.h:185-197)
            const R leaving = own[L3::opp(i)];
            const R qi = q[i];
            if (qi == R(0.5)) {
                v[i] = leaving;
            } else if (qi < R(0.5)) {
                // the farther node's output opp(i)
                v[i] = R(2) * qi * leaving + (R(1) - R(2) * qi) * decode<T>(p, fv[i], L3::opp(i));
            } else {
                v[i] = (leaving + (R(2) * qi - R(1)) * own[i]) / (R(2) * qi);
            }
        } else if (kind == kWall) {
            const R u = payload == 0U ? p.wall_u[0] : (payload == 1U ? p.wall_u[1] : (payload == 2U ? p.wall_u[2] : p.wall_u[3]));
            v[i] = own[L3::opp(i)] + R(6) * R(L3::w(i)) * wall_rho * R(L3::cx(i)) * u;
        } else if (kind == kInlet) {
            v[i] = R(0);
        } else {
            R up = sv[i];
            if (p.convective) {
                up = (hist[i] + p.outlet_speed * up) / (R(1) + p.outlet_speed);
                hist_new[i] = up;
            }
            v[i] = up;
        }
    }
    if (p.x_mode == 1 && x == 0) {
        inlet_west3<T>(v, p);
    }
    const R sponge = p.sponge_strength > R(0) ? sponge_sigma3(p, x, y, z) : R(0);
    collide3<T>(v, p, sponge);
    const std::uint32_t bits = T::dithered ? cell_dither(cell, slot_dither(*p.clock)) : 0U;
    const std::uint32_t wmask = p.bwrite[b];
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        if (L3::cx(i) < 0 && p.convective && static_cast<int>(code[i] & 7U) == kOutlet) {
            p.hist[L3::outlet_slot(i) * bs + b] = hist_new[i];
        }
    }
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        const S packed = encode<T>(p, v[i], i, direction_dither<T>(bits, i));
        sw[i * bs + b] = decode<T>(p, packed, i);
        if (((wmask >> i) & 1U) == 0U) {
            continue;
        }
        if constexpr (kLocal) {
            f[pop(p, L3::opp(i), cell)] = packed;
        } else {
            // The reader of output i and the direction it pulls: x + c_i through the sides
            // (periodic: wrapped; slip: the row kept and the direction mirrored, a reader that
            // is a side node, or a bulk node across a periodic side; other readers through a
            // side read the stash and have no bit in the mask). It loads it from A[d][reader] in
            // the next A step.
            int tx = static_cast<int>(x) + L3::cx(i);
            int ty = static_cast<int>(y) + L3::cy(i);
            int tz = static_cast<int>(z) + L3::cz(i);
            bool my = false;
            bool mz = false;
            bool inside = true;
            if (ty < 0 || ty >= static_cast<int>(p.ny)) {
                const int kind = ty < 0 ? p.side[0] : p.side[1];
                if (kind == kPeriodic) {
                    ty = ty < 0 ? ty + static_cast<int>(p.ny) : ty - static_cast<int>(p.ny);
                } else if (kind == kSlip) {
                    ty = static_cast<int>(y);
                    my = true;
                } else {
                    inside = false;
                }
            }
            if (tz < 0 || tz >= static_cast<int>(p.nz)) {
                const int kind = tz < 0 ? p.side[2] : p.side[3];
                if (kind == kPeriodic) {
                    tz = tz < 0 ? tz + static_cast<int>(p.nz) : tz - static_cast<int>(p.nz);
                } else if (kind == kSlip) {
                    tz = static_cast<int>(z);
                    mz = true;
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
                const int d = pick4(i, L3::mirror_y(i), L3::mirror_z(i), L3::mirror_z(L3::mirror_y(i)), my, mz);
                f[pop(p, d, static_cast<std::uint32_t>(tx + nx * ty + plane * tz))] = packed;
            }
        }
    }
}


} // namespace lbm::d3
