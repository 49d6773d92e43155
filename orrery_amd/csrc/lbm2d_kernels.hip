// D2Q9 kernels: the in-place stream-collide step (bulk and boundary halves), initialisation,
// FP64 readbacks and reductions. Physics of src/aero/lbm/lbm2d_cpu.cpp:28-188,352-404 and
// src/render/shaders/lbm/lbm_d2q9_common.glsl:55-209 (halfway bounce-back only, regularised
// Zou-He inlet, convective or zero-gradient outlet), with the addressing of lbm_device.h.
#include <hip/hip_runtime.h>

#include "lbm_device.h"
#include "lbm_kernels.h"

namespace lbm {

template <int K> __global__ void reduce_partials(const double* partial, std::uint32_t n, double* out);
__global__ void reduce_max(const double* partial, std::uint32_t n, double* out);

namespace {

using L2 = D2Q9;

__device__ __forceinline__ int off2(int i, int nx) { return L2::cx(i) + nx * L2::cy(i); }

// Entry i of the storage offset table (read-only while kernels run; read through the constant
// address space, so uniform reads are scalar loads).
template <class R> __device__ __forceinline__ R offset2(const Lattice2<R>& p, int i) {
    using ConstantR = const __attribute__((address_space(4))) R;
    return ((ConstantR*)p.shift)[i];
}
// A stored population of direction i as the arithmetic value, and back: the 16-bit formats
// store its offset from the free-stream equilibrium (Lattice2::shift), rounded to nearest.
template <class T>
__device__ __forceinline__ typename T::R decode2(const Lattice2<typename T::R>& p, typename T::S s, int i) {
    if constexpr (T::offset) {
        return T::load(s, offset2(p, i));
    } else {
        return T::load(s);
    }
}
template <class T>
__device__ __forceinline__ typename T::S encode2(const Lattice2<typename T::R>& p, typename T::R v, int i) {
    if constexpr (T::offset) {
        return T::nearest(v, offset2(p, 9 + i));
    } else {
        return T::nearest(v);
    }
}

template <class T>
__device__ __forceinline__ void collide2(typename T::R (&f)[9], const Lattice2<typename T::R>& p) {
    using R = typename T::R;
    if constexpr (T::shifted) {
        // lbm_d2q9_common.glsl:161-209
        const R drho = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8];
        const R jx = f[1] - f[3] + f[5] - f[6] - f[7] + f[8];
        const R jy = f[2] - f[4] + f[5] + f[6] - f[7] - f[8];
        const R rho = R(1) + drho;
        const R ux = (jx + R(0.5) * p.force[0]) / rho;
        const R uy = (jy + R(0.5) * p.force[1]) / rho;
        const R uu = ux * ux + uy * uy;
        R feq[9];
        R pxx = 0, pyy = 0, pxy = 0;
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            const R ex = R(L2::cx(i));
            const R ey = R(L2::cy(i));
            const R cu = ex * ux + ey * uy;
            feq[i] = R(L2::w(i)) * (drho + rho * (R(3) * cu + R(4.5) * cu * cu - R(1.5) * uu));
            const R neq = f[i] - feq[i];
            pxx += ex * ex * neq;
            pyy += ey * ey * neq;
            pxy += ex * ey * neq;
        }
        R omega = p.omega;
        if (p.les) {
            const R strain = sqrt(pxx * pxx + pyy * pyy + R(2) * pxy * pxy);
            const R tau = R(0.5) * (p.tau + sqrt(p.tau * p.tau + R(25.455844122715710) * p.smag_c2 * strain / rho));
            omega = R(1) / tau;
        }
        const R force_scale = R(1) - R(0.5) * omega;
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            R value = f[i] - omega * (f[i] - feq[i]);
            if (p.forced) {
                const R ex = R(L2::cx(i));
                const R ey = R(L2::cy(i));
                const R cu = ex * ux + ey * uy;
                value += force_scale * R(L2::w(i)) *
                         (R(3) * ((ex - ux) * p.force[0] + (ey - uy) * p.force[1]) +
                          R(9) * cu * (ex * p.force[0] + ey * p.force[1]));
            }
            f[i] = value;
        }
    } else {
        // lbm2d_cpu.cpp:141-188
        const double rho = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8];
        const double jx = f[1] - f[3] + f[5] - f[6] - f[7] + f[8];
        const double jy = f[2] - f[4] + f[5] + f[6] - f[7] - f[8];
        const double fx = p.force[0];
        const double fy = p.force[1];
        const double ux = (jx + 0.5 * fx) / rho;
        const double uy = (jy + 0.5 * fy) / rho;
        const double base = 1.0 - 1.5 * (ux * ux + uy * uy);
        const double cu[9] = {0.0, ux, uy, -ux, -uy, ux + uy, uy - ux, -ux - uy, ux - uy};
        double feq[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            feq[i] = L2::w(i) * rho * (base + cu[i] * (3.0 + 4.5 * cu[i]));
        }
        double tau = p.tau;
        if (p.les) {
            double pxx = 0.0, pyy = 0.0, pxy = 0.0;
#pragma unroll
            for (int i = 0; i < 9; ++i) {
                const double neq = f[i] - feq[i];
                const double ex = L2::cx(i);
                const double ey = L2::cy(i);
                pxx += ex * ex * neq;
                pyy += ey * ey * neq;
                pxy += ex * ey * neq;
            }
            const double strain = sqrt(pxx * pxx + pyy * pyy + 2.0 * pxy * pxy);
            const double c = p.smag_c;
            tau = 0.5 * (tau + sqrt(tau * tau + 18.0 * 1.4142135623730951 * c * c * strain / rho));
        }
        const double omega = 1.0 / tau;
        const double force_scale = 1.0 - 0.5 * omega;
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            double value = f[i] - omega * (f[i] - feq[i]);
            if (p.forced) {
                const double ex = L2::cx(i);
                const double ey = L2::cy(i);
                value += force_scale * L2::w(i) * (3.0 * ((ex - ux) * fx + (ey - uy) * fy) + 9.0 * cu[i] * (ex * fx + ey * fy));
            }
            f[i] = value;
        }
    }
}

// Regularised Zou-He inlet (zou_he_west, src/aero/lbm/lbm2d_internal.h:85-111; shifted form
// lbm_d2q9_common.glsl:55-87).
template <class T>
__device__ __forceinline__ void zou_he_west(typename T::R (&f)[9], const Lattice2<typename T::R>& p) {
    using R = typename T::R;
    const R ux = p.inlet[0];
    const R uy = p.inlet[1];
    R rho;
    R drho;
    if constexpr (T::shifted) {
        const R s = f[0] + f[2] + f[4] + R(2) * (f[3] + f[6] + f[7]);
        drho = (s + ux) / (R(1) - ux);
        rho = R(1) + drho;
    } else {
        rho = (f[0] + f[2] + f[4] + 2.0 * (f[3] + f[6] + f[7])) / (1.0 - ux);
        drho = rho;
    }
    f[1] = f[3] + (R(2) / R(3)) * rho * ux;
    f[5] = f[7] - R(0.5) * (f[2] - f[4]) + (R(1) / R(6)) * rho * ux + R(0.5) * rho * uy;
    f[8] = f[6] + R(0.5) * (f[2] - f[4]) + (R(1) / R(6)) * rho * ux - R(0.5) * rho * uy;
    const R uu = ux * ux + uy * uy;
    R feq[9];
    R pxx = 0, pyy = 0, pxy = 0;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const R ex = R(L2::cx(i));
        const R ey = R(L2::cy(i));
        const R cu = ex * ux + ey * uy;
        if constexpr (T::shifted) {
            feq[i] = R(L2::w(i)) * (drho + rho * (R(3) * cu + R(4.5) * cu * cu - R(1.5) * uu));
        } else {
            feq[i] = L2::w(i) * rho * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * uu);
        }
        const R neq = f[i] - feq[i];
        pxx += ex * ex * neq;
        pyy += ey * ey * neq;
        pxy += ex * ey * neq;
    }
    const R cs2 = R(1.0 / 3.0);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const R ex = R(L2::cx(i));
        const R ey = R(L2::cy(i));
        f[i] = feq[i] + R(4.5) * R(L2::w(i)) * ((ex * ex - cs2) * pxx + (ey * ey - cs2) * pyy + R(2) * ex * ey * pxy);
    }
}

template <class T, bool kLocal>
__device__ __forceinline__ void bulk_cell2(const Lattice2<typename T::R>& p, std::uint32_t cell) {
    using S = typename T::S;
    using R = typename T::R;
    S* __restrict__ f = static_cast<S*>(p.f);
    if (cell >= p.cells || p.cls[cell] != 1) {
        return;
    }
    const int nx = static_cast<int>(p.nx);
    const std::uint32_t stride = p.stride;
    R v[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        if constexpr (kLocal) {
            v[i] = decode2<T>(p, f[i * stride + cell], i);
        } else {
            v[i] = decode2<T>(p, f[L2::opp(i) * stride + static_cast<std::uint32_t>(static_cast<int>(cell) - off2(i, nx))], i);
        }
    }
    collide2<T>(v, p);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        if constexpr (kLocal) {
            f[L2::opp(i) * stride + cell] = encode2<T>(p, v[i], i);
        } else {
            f[i * stride + static_cast<std::uint32_t>(static_cast<int>(cell) + off2(i, nx))] = encode2<T>(p, v[i], i);
        }
    }
}

template <class T, bool kLocal>
__device__ __forceinline__ void boundary_node2(const Lattice2<typename T::R>& p, std::uint32_t b) {
    using S = typename T::S;
    using R = typename T::R;
    if (b >= p.nb) {
        return;
    }
    S* __restrict__ f = static_cast<S*>(p.f);
    const std::uint32_t stride = p.stride;
    const std::uint32_t bs = p.bstride;
    const int nx = static_cast<int>(p.nx);
    const std::uint32_t cell = p.bcell[b];
    const std::uint32_t y = p.div_nx.div(cell);
    const std::uint32_t x = cell - y * p.nx;
    const R* __restrict__ st = p.stash + (kLocal ? 9U * bs : 0U);
    R* __restrict__ sw = p.stash + (kLocal ? 0U : 9U * bs);
    R v[9];
    R wall_rho = R(-1);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const std::uint32_t code = p.bcode[i * bs + b];
        const int kind = static_cast<int>(code & 7U);
        const std::uint32_t payload = code >> kCodeShift;
        if (kind == kNode) {
            if ((code & kCodeMirrorY) == 0U) {
                if constexpr (kLocal) {
                    v[i] = decode2<T>(p, f[i * stride + cell], i);
                } else {
                    v[i] = decode2<T>(p, f[L2::opp(i) * stride + payload], i);
                }
            } else {
                v[i] = st[L2::mirror_y(i) * bs + payload];
            }
        } else if (kind == kBounce) {
            v[i] = st[L2::opp(i) * bs + b];
        } else if (kind == kWall) {
            if (wall_rho < R(0)) {
                R s = 0;
#pragma unroll
                for (int k = 0; k < 9; ++k) {
                    s += st[k * bs + b];
                }
                wall_rho = T::shifted ? R(1) + s : s;
            }
            v[i] = st[L2::opp(i) * bs + b] + R(6) * R(L2::w(i)) * wall_rho * R(L2::cx(i)) * (payload == 0U ? p.wall_u[0] : p.wall_u[1]);
        } else if (kind == kInlet) {
            v[i] = R(0);
        } else {
            const int d = (code & kCodeMirrorY) ? L2::mirror_y(i) : i;
            R up = st[d * bs + payload];
            if (p.convective) {
                R* h = p.hist + L2::outlet_slot(i) * bs + b;
                up = (*h + p.outlet_speed * up) / (R(1) + p.outlet_speed);
                *h = up;
            }
            v[i] = up;
        }
    }
    if (p.x_mode == 1 && x == 0) {
        zou_he_west<T>(v, p);
    }
    collide2<T>(v, p);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const S packed = encode2<T>(p, v[i], i);
        sw[i * bs + b] = decode2<T>(p, packed, i);
        if constexpr (kLocal) {
            f[L2::opp(i) * stride + cell] = packed;
        } else {
            int tx = static_cast<int>(x) + L2::cx(i);
            int ty = static_cast<int>(y) + L2::cy(i);
            bool inside = true;
            if (ty < 0 || ty >= static_cast<int>(p.ny)) {
                if ((ty < 0 ? p.side[0] : p.side[1]) == kPeriodic) {
                    ty = ty < 0 ? ty + static_cast<int>(p.ny) : ty - static_cast<int>(p.ny);
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
                f[i * stride + static_cast<std::uint32_t>(tx + nx * ty)] = packed;
            }
        }
    }
}

// The bulk half of a step: one thread per cell (block order renumbered per XCD).
template <class T, bool kLocal>
__global__ __launch_bounds__(256) void d2q9_bulk(Lattice2<typename T::R> p) {
    bulk_cell2<T, kLocal>(p, xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x);
}

// The boundary half: one thread per boundary node.
template <class T, bool kLocal>
__global__ __launch_bounds__(256) void d2q9_boundary(Lattice2<typename T::R> p) {
    boundary_node2<T, kLocal>(p, blockIdx.x * blockDim.x + threadIdx.x);
}

// Both halves in one launch: the first `bulk_blocks` blocks take the cells, the others the
// boundary nodes. Within a step every population address is read and written by one node only
// and the stash is double-buffered, so the two halves need no order between them; one launch
// per step is what a lattice small enough to live in cache is bound by.
template <class T, bool kLocal>
__global__ __launch_bounds__(256) void d2q9_step_kernel(Lattice2<typename T::R> p, std::uint32_t bulk_blocks) {
    if (blockIdx.x < bulk_blocks) {
        bulk_cell2<T, kLocal>(p, xcd_tile(blockIdx.x, bulk_blocks) * blockDim.x + threadIdx.x);
    } else {
        boundary_node2<T, kLocal>(p, (blockIdx.x - bulk_blocks) * blockDim.x + threadIdx.x);
    }
}

__device__ __forceinline__ double perturbation2(std::uint64_t seed, std::uint64_t cell, std::uint32_t component) {
    std::uint64_t z = seed + 0x9E3779B97F4A7C15ULL * (cell * 4U + component + 1U);
    z = (z ^ (z >> 30U)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27U)) * 0x94D049BB133111EBULL;
    z ^= z >> 31U;
    const double u = static_cast<double>(z >> 11U) * (1.0 / 9007199254740992.0);
    return 2.0 * u - 1.0;
}

template <class T>
__global__ __launch_bounds__(256) void d2q9_init(Lattice2<typename T::R> p, int mode, const double* planes, double u0x,
                                                 double u0y, std::uint64_t seed, double amplitude) {
    const std::uint32_t cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= p.cells) {
        return;
    }
    auto* f = static_cast<typename T::S*>(p.f);
    double rho = 1.0, ux = 0.0, uy = 0.0;
    if (p.cls[cell] != 0) {
        if (mode == 1) {
            rho = planes[cell];
            ux = planes[p.cells + cell];
            uy = planes[2U * p.cells + cell];
        } else if (mode == 2) {
            ux = u0x + amplitude * perturbation2(seed, cell, 0);
            uy = u0y + amplitude * perturbation2(seed, cell, 1);
        } else {
            ux = u0x;
            uy = u0y;
        }
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        // d2q9::equilibrium (d2q9.h:53-58)
        const double cu = static_cast<double>(L2::cx(i)) * ux + static_cast<double>(L2::cy(i)) * uy;
        const double uu = ux * ux + uy * uy;
        double value = L2::w(i) * rho * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * uu);
        if constexpr (T::shifted) {
            value -= L2::w(i);
        }
        if constexpr (T::offset) {
            value -= static_cast<double>(offset2(p, i));
        }
        f[L2::opp(i) * p.stride + cell] = T::from_double(value);
    }
}

template <class T>
__global__ __launch_bounds__(256) void d2q9_init_boundary(Lattice2<typename T::R> p) {
    const std::uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= p.nb) {
        return;
    }
    const auto* f = static_cast<const typename T::S*>(p.f);
    const std::uint32_t cell = p.bcell[b];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const typename T::R v = decode2<T>(p, f[L2::opp(i) * p.stride + cell], i);
        p.stash[i * p.bstride + b] = v;
        if (L2::cx(i) < 0) {
            p.hist[L2::outlet_slot(i) * p.bstride + b] = v;
        }
    }
}

template <class T>
__device__ __forceinline__ typename T::R bulk_value2(const Lattice2<typename T::R>& p, const typename T::S* f, int i,
                                                     std::uint32_t cell, int phase) {
    if (phase == 0) {
        return decode2<T>(p, f[L2::opp(i) * p.stride + cell], i);
    }
    return decode2<T>(p, f[i * p.stride + static_cast<std::uint32_t>(static_cast<int>(cell) + off2(i, static_cast<int>(p.nx)))], i);
}

template <class T>
__global__ __launch_bounds__(256) void d2q9_populations_cells(Lattice2<typename T::R> p, int phase, double* out) {
    const std::uint32_t cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= p.cells || p.cls[cell] == 2) {
        return;
    }
    const auto* f = static_cast<const typename T::S*>(p.f);
    const bool solid = p.cls[cell] == 0;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        double v = L2::w(i);
        if (!solid) {
            v = static_cast<double>(bulk_value2<T>(p, f, i, cell, phase)) + (T::shifted ? L2::w(i) : 0.0);
        }
        out[static_cast<std::size_t>(i) * p.cells + cell] = v;
    }
}

template <class T>
__global__ __launch_bounds__(256) void d2q9_populations_boundary(Lattice2<typename T::R> p, int phase, double* out) {
    const std::uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= p.nb) {
        return;
    }
    const auto* st = p.stash + (phase ? 9U * p.bstride : 0U);
    const std::uint32_t cell = p.bcell[b];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        out[static_cast<std::size_t>(i) * p.cells + cell] =
            static_cast<double>(st[i * p.bstride + b]) + (T::shifted ? L2::w(i) : 0.0);
    }
}

template <class T>
__device__ __forceinline__ void moments2(const double (&f)[9], double fx, double fy, double& rho, double& ux, double& uy) {
    double s = 0.0, jx = 0.0, jy = 0.0;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        s += f[i];
        jx += static_cast<double>(L2::cx(i)) * f[i];
        jy += static_cast<double>(L2::cy(i)) * f[i];
    }
    rho = T::shifted ? 1.0 + s : s;
    ux = (jx - 0.5 * fx) / rho;
    uy = (jy - 0.5 * fy) / rho;
}

template <class T>
__global__ __launch_bounds__(256) void d2q9_moments_cells(Lattice2<typename T::R> p, int phase, double fx, double fy,
                                                          double* out) {
    const std::uint32_t cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= p.cells || p.cls[cell] == 2) {
        return;
    }
    double rho = 1.0, ux = 0.0, uy = 0.0;
    if (p.cls[cell] == 1) {
        const auto* f = static_cast<const typename T::S*>(p.f);
        double v[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            v[i] = static_cast<double>(bulk_value2<T>(p, f, i, cell, phase));
        }
        moments2<T>(v, fx, fy, rho, ux, uy);
    }
    out[cell] = rho;
    out[p.cells + cell] = ux;
    out[2U * static_cast<std::size_t>(p.cells) + cell] = uy;
}

template <class T>
__global__ __launch_bounds__(256) void d2q9_moments_boundary(Lattice2<typename T::R> p, int phase, double fx,
                                                             double fy, double* out) {
    const std::uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= p.nb) {
        return;
    }
    const auto* st = p.stash + (phase ? 9U * p.bstride : 0U);
    double v[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        v[i] = static_cast<double>(st[i * p.bstride + b]);
    }
    double rho, ux, uy;
    moments2<T>(v, fx, fy, rho, ux, uy);
    const std::uint32_t cell = p.bcell[b];
    out[cell] = rho;
    out[p.cells + cell] = ux;
    out[2U * static_cast<std::size_t>(p.cells) + cell] = uy;
}

// Moments of listed nodes (the values macroscopic() reports there): a node with a boundary
// slot reads its stash entries, a bulk node its lattice entries, a solid node is at rest.
template <class T>
__global__ __launch_bounds__(256) void d2q9_sample(Lattice2<typename T::R> p, int phase, double fx, double fy,
                                                   const std::uint32_t* cells, const std::int32_t* slots,
                                                   std::uint32_t n, double* out) {
    const std::uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) {
        return;
    }
    const std::uint32_t cell = cells[k];
    const std::int32_t slot = slots[k];
    double rho = 1.0, ux = 0.0, uy = 0.0;
    double v[9];
    if (slot >= 0) {
        const auto* st = p.stash + (phase ? 9U * p.bstride : 0U);
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            v[i] = static_cast<double>(st[i * p.bstride + static_cast<std::uint32_t>(slot)]);
        }
        moments2<T>(v, fx, fy, rho, ux, uy);
    } else if (cell < p.cells && p.cls[cell] == 1) {
        const auto* f = static_cast<const typename T::S*>(p.f);
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            v[i] = static_cast<double>(bulk_value2<T>(p, f, i, cell, phase));
        }
        moments2<T>(v, fx, fy, rho, ux, uy);
    }
    out[k] = rho;
    out[n + k] = ux;
    out[2U * n + k] = uy;
}

// max_speed (lbm2d_common.cpp:168-176): hypot(u, v) of the moments of each fluid node.
template <class T>
__global__ __launch_bounds__(256) void d2q9_speed_cells(Lattice2<typename T::R> p, int phase, double fx, double fy,
                                                        double* partial) {
    const std::uint32_t cell = blockIdx.x * blockDim.x + threadIdx.x;
    double speed = 0.0;
    if (cell < p.cells && p.cls[cell] == 1) {
        const auto* f = static_cast<const typename T::S*>(p.f);
        double v[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            v[i] = static_cast<double>(bulk_value2<T>(p, f, i, cell, phase));
        }
        double rho, ux, uy;
        moments2<T>(v, fx, fy, rho, ux, uy);
        speed = hypot(ux, uy);
    }
    block_max(speed, partial);
}

template <class T>
__global__ __launch_bounds__(256) void d2q9_speed_boundary(Lattice2<typename T::R> p, int phase, double fx, double fy,
                                                           double* partial) {
    const std::uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    double speed = 0.0;
    if (b < p.nb) {
        const auto* st = p.stash + (phase ? 9U * p.bstride : 0U);
        double v[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            v[i] = static_cast<double>(st[i * p.bstride + b]);
        }
        double rho, ux, uy;
        moments2<T>(v, fx, fy, rho, ux, uy);
        speed = hypot(ux, uy);
    }
    block_max(speed, partial);
}

template <int K>
__device__ __forceinline__ void block_sum2(double (&v)[K], double* partial) {
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

// Momentum exchange of the halfway wall (lbm2d_cpu.cpp:265-272, add_link_force lbm2d.h:145-154):
// 2 c_o f*_o per link, moment about the origin at x - c_j / 2.
template <class T>
__global__ __launch_bounds__(256) void d2q9_force(Lattice2<typename T::R> p, int phase, const std::uint32_t* lcell,
                                                  const std::uint8_t* ldir, const std::uint32_t* lslot,
                                                  std::uint32_t nlinks, double* partial) {
    const std::uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    double v[3] = {0, 0, 0};
    if (k < nlinks) {
        const auto* st = p.stash + (phase ? 9U * p.bstride : 0U);
        const std::uint32_t cell = lcell[k];
        const int j = ldir[k];
        const int o = L2::opp(j);
        const double outgoing = static_cast<double>(st[o * p.bstride + lslot[k]]) + (T::shifted ? L2::w(o) : 0.0);
        const double fx = 2.0 * static_cast<double>(L2::cx(o)) * outgoing;
        const double fy = 2.0 * static_cast<double>(L2::cy(o)) * outgoing;
        const std::uint32_t x = cell % p.nx;
        const std::uint32_t y = cell / p.nx;
        const double rx = static_cast<double>(x) - 0.5 * L2::cx(j);
        const double ry = static_cast<double>(y) - 0.5 * L2::cy(j);
        v[0] = fx;
        v[1] = fy;
        v[2] = rx * fy - ry * fx;
    }
    block_sum2<3>(v, partial);
}

template <class T>
__global__ __launch_bounds__(256) void d2q9_mass_cells(Lattice2<typename T::R> p, int phase, double* partial) {
    const std::uint32_t cell = blockIdx.x * blockDim.x + threadIdx.x;
    double v[1] = {0.0};
    if (cell < p.cells && p.cls[cell] == 1) {
        const auto* f = static_cast<const typename T::S*>(p.f);
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            s += static_cast<double>(bulk_value2<T>(p, f, i, cell, phase)) + (T::shifted ? L2::w(i) : 0.0);
        }
        v[0] = s;
    }
    block_sum2<1>(v, partial);
}

template <class T>
__global__ __launch_bounds__(256) void d2q9_mass_boundary(Lattice2<typename T::R> p, int phase, double* partial) {
    const std::uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    double v[1] = {0.0};
    if (b < p.nb) {
        const auto* st = p.stash + (phase ? 9U * p.bstride : 0U);
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            s += static_cast<double>(st[i * p.bstride + b]) + (T::shifted ? L2::w(i) : 0.0);
        }
        v[0] = s;
    }
    block_sum2<1>(v, partial);
}

inline std::uint32_t blocks2(std::uint64_t n, std::uint32_t block) {
    return static_cast<std::uint32_t>((n + block - 1) / block);
}

} // namespace

template <class T>
hipError_t d2q9_step(const Lattice2<typename T::R>& p, bool local, int block, hipStream_t s) {
    const std::uint32_t bulk = blocks2(p.cells, static_cast<std::uint32_t>(block));
    const std::uint32_t grid = bulk + blocks2(p.nb, static_cast<std::uint32_t>(block));
    if (local) {
        hipLaunchKernelGGL((d2q9_step_kernel<T, true>), dim3(grid), dim3(block), 0, s, p, bulk);
    } else {
        hipLaunchKernelGGL((d2q9_step_kernel<T, false>), dim3(grid), dim3(block), 0, s, p, bulk);
    }
    return hipGetLastError();
}

template <class T>
hipError_t d2q9_step_split(const Lattice2<typename T::R>& p, bool local, int block, hipStream_t s) {
    const std::uint32_t grid = blocks2(p.cells, static_cast<std::uint32_t>(block));
    if (local) {
        hipLaunchKernelGGL((d2q9_bulk<T, true>), dim3(grid), dim3(block), 0, s, p);
    } else {
        hipLaunchKernelGGL((d2q9_bulk<T, false>), dim3(grid), dim3(block), 0, s, p);
    }
    if (p.nb > 0) {
        const std::uint32_t gb = blocks2(p.nb, 256);
        if (local) {
            hipLaunchKernelGGL((d2q9_boundary<T, true>), dim3(gb), dim3(256), 0, s, p);
        } else {
            hipLaunchKernelGGL((d2q9_boundary<T, false>), dim3(gb), dim3(256), 0, s, p);
        }
    }
    return hipGetLastError();
}

template <class T>
hipError_t d2q9_initialize(const Lattice2<typename T::R>& p, int mode, const double* planes, const double u0[2],
                           std::uint64_t seed, double amplitude, hipStream_t s) {
    hipLaunchKernelGGL((d2q9_init<T>), dim3(blocks2(p.cells, 256)), dim3(256), 0, s, p, mode, planes, u0[0], u0[1],
                       seed, amplitude);
    if (p.nb > 0) {
        hipLaunchKernelGGL((d2q9_init_boundary<T>), dim3(blocks2(p.nb, 256)), dim3(256), 0, s, p);
    }
    return hipGetLastError();
}

template <class T>
hipError_t d2q9_populations(const Lattice2<typename T::R>& p, int phase, double* out, hipStream_t s) {
    hipLaunchKernelGGL((d2q9_populations_cells<T>), dim3(blocks2(p.cells, 256)), dim3(256), 0, s, p, phase, out);
    if (p.nb > 0) {
        hipLaunchKernelGGL((d2q9_populations_boundary<T>), dim3(blocks2(p.nb, 256)), dim3(256), 0, s, p, phase, out);
    }
    return hipGetLastError();
}

template <class T>
hipError_t d2q9_moments(const Lattice2<typename T::R>& p, int phase, const double force[2], double* out,
                        hipStream_t s) {
    hipLaunchKernelGGL((d2q9_moments_cells<T>), dim3(blocks2(p.cells, 256)), dim3(256), 0, s, p, phase, force[0],
                       force[1], out);
    if (p.nb > 0) {
        hipLaunchKernelGGL((d2q9_moments_boundary<T>), dim3(blocks2(p.nb, 256)), dim3(256), 0, s, p, phase,
                           force[0], force[1], out);
    }
    return hipGetLastError();
}

template <class T>
hipError_t d2q9_sample(const Lattice2<typename T::R>& p, int phase, const double force[2], const std::uint32_t* cells,
                       const std::int32_t* slots, std::uint32_t n, double* out, hipStream_t s) {
    if (n == 0) {
        return hipSuccess;
    }
    hipLaunchKernelGGL((d2q9_sample<T>), dim3(blocks2(n, 256)), dim3(256), 0, s, p, phase, force[0], force[1], cells,
                       slots, n, out);
    return hipGetLastError();
}

template <class T>
hipError_t d2q9_max_speed(const Lattice2<typename T::R>& p, int phase, const double force[2], double* partial,
                          double* result, hipStream_t s) {
    const std::uint32_t nc = blocks2(p.cells, 256);
    hipLaunchKernelGGL((d2q9_speed_cells<T>), dim3(nc), dim3(256), 0, s, p, phase, force[0], force[1], partial);
    std::uint32_t n = nc;
    if (p.nb > 0) {
        const std::uint32_t nb = blocks2(p.nb, 256);
        hipLaunchKernelGGL((d2q9_speed_boundary<T>), dim3(nb), dim3(256), 0, s, p, phase, force[0], force[1],
                           partial + nc);
        n += nb;
    }
    hipLaunchKernelGGL(reduce_max, dim3(1), dim3(256), 0, s, partial, n, result);
    return hipGetLastError();
}

template <class T>
hipError_t d2q9_mass(const Lattice2<typename T::R>& p, int phase, double* partial, double* out, hipStream_t s) {
    const std::uint32_t gc = blocks2(p.cells, 256);
    const std::uint32_t gb = blocks2(p.nb > 0 ? p.nb : 1, 256);
    hipLaunchKernelGGL((d2q9_mass_cells<T>), dim3(gc), dim3(256), 0, s, p, phase, partial);
    hipLaunchKernelGGL((d2q9_mass_boundary<T>), dim3(gb), dim3(256), 0, s, p, phase, partial + gc);
    hipLaunchKernelGGL((reduce_partials<1>), dim3(1), dim3(256), 0, s, partial, gc + gb, out);
    return hipGetLastError();
}

template <class T>
hipError_t d2q9_force_sum(const Lattice2<typename T::R>& p, int phase, const std::uint32_t* lcell,
                          const std::uint8_t* ldir, const std::uint32_t* lslot, std::uint32_t nlinks, double* partial,
                          double* out, hipStream_t s) {
    const std::uint32_t g = blocks2(nlinks > 0 ? nlinks : 1, 256);
    hipLaunchKernelGGL((d2q9_force<T>), dim3(g), dim3(256), 0, s, p, phase, lcell, ldir, lslot, nlinks, partial);
    hipLaunchKernelGGL((reduce_partials<3>), dim3(1), dim3(256), 0, s, partial, g, out);
    return hipGetLastError();
}

#define LBM_INSTANTIATE2(T)                                                                                     \
    template hipError_t d2q9_step<T>(const Lattice2<T::R>&, bool, int, hipStream_t);                            \
    template hipError_t d2q9_step_split<T>(const Lattice2<T::R>&, bool, int, hipStream_t);                      \
    template hipError_t d2q9_initialize<T>(const Lattice2<T::R>&, int, const double*, const double*,            \
                                           std::uint64_t, double, hipStream_t);                                 \
    template hipError_t d2q9_populations<T>(const Lattice2<T::R>&, int, double*, hipStream_t);                  \
    template hipError_t d2q9_moments<T>(const Lattice2<T::R>&, int, const double*, double*, hipStream_t);       \
    template hipError_t d2q9_sample<T>(const Lattice2<T::R>&, int, const double*, const std::uint32_t*,         \
                                       const std::int32_t*, std::uint32_t, double*, hipStream_t);               \
    template hipError_t d2q9_max_speed<T>(const Lattice2<T::R>&, int, const double*, double*, double*,          \
                                          hipStream_t);                                                         \
    template hipError_t d2q9_mass<T>(const Lattice2<T::R>&, int, double*, double*, hipStream_t);                \
    template hipError_t d2q9_force_sum<T>(const Lattice2<T::R>&, int, const std::uint32_t*, const std::uint8_t*, \
                                          const std::uint32_t*, std::uint32_t, double*, double*, hipStream_t);

LBM_INSTANTIATE2(StoreF32)
LBM_INSTANTIATE2(StoreF16)
LBM_INSTANTIATE2(StoreBF16)
LBM_INSTANTIATE2(StoreF64)

} // namespace lbm
