// D3Q19 kernels of one lattice: the fused in-place stream-collide step (bulk and boundary
// halves, device functions in lbm3d_step.h), state initialisation, and the FP64 readbacks and
// reductions (populations, moments, mass, force).
#include <hip/hip_runtime.h>

#include "lbm3d_step.h"
#include "lbm_device.h"
#include "lbm_kernels.h"

namespace lbm {

namespace {

using L3 = D3Q19;
using d3::off3;

// ---- the step kernels -------------------------------------------------------------------------
// Bulk: every bulk node (all 18 neighbours inside the domain and fluid) streams plainly, V
// consecutive cells per thread. Boundary: one thread per boundary node. kLocal selects the A
// step (odd step count) or the B step (even).
// The bulk half of a step (d3::bulk_run in lbm3d_step.h): each thread takes V consecutive cells.
template <class T, bool kLocal, int V>
__global__ __launch_bounds__(256) void d3q19_bulk(Lattice3<typename T::R> p) {
    d3::bulk_run<T, kLocal, V>(p, (blockIdx.x * blockDim.x + threadIdx.x) * V);
}

// The whole step in one launch, in cell order: each wave takes 64 consecutive cells, runs the
// bulk update on its bulk lanes and then the boundary update on its boundary lanes. A boundary
// node's slot is the number of boundary nodes before its 64-cell run (tile_slot) plus its rank
// among the wave's boundary lanes. The boundary nodes of a row then share cache lines with the
// bulk nodes around them, where a separate boundary pass would fetch a line per population for
// each node of the inlet and outlet faces (which lie one row apart in memory).
template <class T, bool kLocal>
__global__ __launch_bounds__(256) void d3q19_fused(Lattice3<typename T::R> p) {
    using R = typename T::R;
    const std::uint32_t cell = blockIdx.x * blockDim.x + threadIdx.x;
    const bool inside = cell < p.cells;
    const std::uint32_t c = inside ? cell : p.cells - 1;
    const std::uint8_t kind = inside ? p.cls[c] : std::uint8_t(0);
    if (kLocal || kind != kSideNode) {
        R v[19];
        d3::bulk_load<T, kLocal>(p, c, v);
        if (d3::streamed(kind)) {
            if (kind == kInletNode) {
                d3::inlet_west3<T>(v, p);
            }
            d3::bulk_finish<T, kLocal>(p, c, v);  // in the A step a side node streams as a bulk node
        }
    } else {
        d3::side_node_b<T>(p, c);
    }
    const std::uint64_t mask = __ballot(kind == 2);
    if (kind == 2) {
        const std::uint32_t rank = __builtin_amdgcn_mbcnt_hi(static_cast<std::uint32_t>(mask >> 32),
                                                             __builtin_amdgcn_mbcnt_lo(static_cast<std::uint32_t>(mask), 0U));
        d3::boundary_node<T, kLocal>(p, p.tile_slot[c >> 6] + rank);
    }
}

// B step of the side nodes (in A steps the bulk kernel updates them).
template <class T> __global__ __launch_bounds__(256) void d3q19_side(Lattice3<typename T::R> p) {
    const std::uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < p.ns) {
        d3::side_node_b<T>(p, p.scell[s]);
    }
}

template <class T, bool kLocal>
__global__ __launch_bounds__(256) void d3q19_boundary(Lattice3<typename T::R> p) {
    const std::uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < p.nb) {
        d3::boundary_node<T, kLocal>(p, b);
    }
}

__global__ void d3q19_tick_kernel(std::uint32_t* clock) {
    if (threadIdx.x == 0) {
        *clock += 1U;
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
        if constexpr (T::offset) {
            value -= static_cast<double>(p.shift[i]);
        }
        f[d3::pop(p, L3::opp(i), cell)] = T::from_double(value);
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
        const typename T::R v = d3::decode<T>(p, f[d3::pop(p, L3::opp(i), cell)], i);
        p.stash[i * p.bstride + b] = v;
        if (L3::cx(i) < 0) {
            p.hist[L3::outlet_slot(i) * p.bstride + b] = v;
        }
    }
}

// ---- readbacks -------------------------------------------------------------------------------
// Post-collision population i of a bulk or side node (class `kind`) in the layout of
// `phase` = n & 1: after an A step its own cell's A[opp(i)]; after a B step the address its
// reader loads, which is the B-step address of the node's pull opp(i).
template <class T>
__device__ __forceinline__ typename T::R bulk_value(const Lattice3<typename T::R>& p, const typename T::S* f,
                                                    int i, std::uint32_t cell, int phase, std::uint8_t kind) {
    if (phase == 0) {
        return d3::decode<T>(p, f[d3::pop(p, L3::opp(i), cell)], i);
    }
    if (kind == kSideNode) {
        return d3::decode<T>(p, f[d3::side_element(p, cell, d3::side_faces(p, cell), L3::opp(i))], i);
    }
    const int off = off3(i, static_cast<int>(p.nx), static_cast<int>(p.nx * p.ny));
    return d3::decode<T>(p, f[d3::pop(p, i, cell, off)], i);
}

// Populations with the weights added back, [19][cells] doubles. Pass 1 over cells covers
// solid, bulk and side nodes; pass 2 over the boundary list fills boundary nodes from the stash.
template <class T>
__global__ __launch_bounds__(256) void d3q19_populations_cells(Lattice3<typename T::R> p, int phase, double* out) {
    const std::uint32_t cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= p.cells) {
        return;
    }
    const std::uint8_t c = p.cls[cell];
    if (c == kListedNode) {
        return;
    }
    const auto* f = static_cast<const typename T::S*>(p.f);
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        double v;
        if (c == kSolidNode) {
            v = L3::w(i);
        } else {
            v = static_cast<double>(bulk_value<T>(p, f, i, cell, phase, c));
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
    if (c == kListedNode) {
        return;
    }
    double rho = 1.0, ux = 0.0, uy = 0.0, uz = 0.0;
    if (d3::streamed(c)) {
        const auto* f = static_cast<const typename T::S*>(p.f);
        double v[19];
#pragma unroll
        for (int i = 0; i < 19; ++i) {
            v[i] = static_cast<double>(bulk_value<T>(p, f, i, cell, phase, c));
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
                farther = static_cast<double>(
                    d3::decode<T>(p, f[d3::pop(p, j, cell, off)], o));
            } else {
                farther = static_cast<double>(d3::decode<T>(p, f[d3::pop(p, o, cell)], o));
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

// Mass: pass over cells (bulk and side nodes) and pass over the boundary list, each to block
// partials.
template <class T>
__global__ __launch_bounds__(256) void d3q19_mass_cells(Lattice3<typename T::R> p, int phase, double* partial) {
    const std::uint32_t cell = blockIdx.x * blockDim.x + threadIdx.x;
    double v[1] = {0.0};
    const std::uint8_t c = cell < p.cells ? p.cls[cell] : std::uint8_t(kSolidNode);
    if (d3::streamed(c)) {
        const auto* f = static_cast<const typename T::S*>(p.f);
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < 19; ++i) {
            s += static_cast<double>(bulk_value<T>(p, f, i, cell, phase, c)) + (T::shifted ? L3::w(i) : 0.0);
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

// Moments of listed nodes (the values macroscopic() reports there): a node with a boundary
// slot reads its stash entries, a bulk or side node its lattice entries, anything else is at
// rest.
template <class T>
__global__ __launch_bounds__(256) void d3q19_sample(Lattice3<typename T::R> p, int phase, double fx, double fy, double fz,
                                                    const std::uint32_t* cells, const std::int32_t* slots, std::uint32_t n,
                                                    double* out) {
    const std::uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) {
        return;
    }
    const std::uint32_t cell = cells[k];
    const std::int32_t slot = slots[k];
    const double force[3] = {fx, fy, fz};
    double rho = 1.0, ux = 0.0, uy = 0.0, uz = 0.0;
    double v[19];
    if (slot >= 0) {
        const auto* st = p.stash + (phase ? 19U * p.bstride : 0U);
#pragma unroll
        for (int i = 0; i < 19; ++i) {
            v[i] = static_cast<double>(st[i * p.bstride + static_cast<std::uint32_t>(slot)]);
        }
        moments_of<T>(v, force, rho, ux, uy, uz);
    } else if (cell < p.cells && d3::streamed(p.cls[cell])) {
        const auto* f = static_cast<const typename T::S*>(p.f);
        const std::uint8_t c = p.cls[cell];
#pragma unroll
        for (int i = 0; i < 19; ++i) {
            v[i] = static_cast<double>(bulk_value<T>(p, f, i, cell, phase, c));
        }
        moments_of<T>(v, force, rho, ux, uy, uz);
    }
    out[k] = rho;
    out[n + k] = ux;
    out[2U * n + k] = uy;
    out[3U * n + k] = uz;
}

// max_speed (lbm3d_common.cpp:200-211): sqrt(u^2 + v^2 + w^2) of the moments of each fluid node.
template <class T>
__global__ __launch_bounds__(256) void d3q19_speed_cells(Lattice3<typename T::R> p, int phase, double fx, double fy,
                                                         double fz, double* partial) {
    const std::uint32_t cell = blockIdx.x * blockDim.x + threadIdx.x;
    double speed = 0.0;
    const std::uint8_t c = cell < p.cells ? p.cls[cell] : std::uint8_t(kSolidNode);
    if (d3::streamed(c)) {
        const auto* f = static_cast<const typename T::S*>(p.f);
        double v[19];
#pragma unroll
        for (int i = 0; i < 19; ++i) {
            v[i] = static_cast<double>(bulk_value<T>(p, f, i, cell, phase, c));
        }
        const double force[3] = {fx, fy, fz};
        double rho, ux, uy, uz;
        moments_of<T>(v, force, rho, ux, uy, uz);
        speed = sqrt(ux * ux + uy * uy + uz * uz);
    }
    block_max(speed, partial);
}

template <class T>
__global__ __launch_bounds__(256) void d3q19_speed_boundary(Lattice3<typename T::R> p, int phase, double fx, double fy,
                                                            double fz, double* partial) {
    const std::uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    double speed = 0.0;
    if (b < p.nb) {
        const auto* st = p.stash + (phase ? 19U * p.bstride : 0U);
        double v[19];
#pragma unroll
        for (int i = 0; i < 19; ++i) {
            v[i] = static_cast<double>(st[i * p.bstride + b]);
        }
        const double force[3] = {fx, fy, fz};
        double rho, ux, uy, uz;
        moments_of<T>(v, force, rho, ux, uy, uz);
        speed = sqrt(ux * ux + uy * uy + uz * uz);
    }
    block_max(speed, partial);
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

// Final stage of max_speed: the largest of `n` partial maxima.
__global__ __launch_bounds__(256) void reduce_max(const double* partial, std::uint32_t n, double* out) {
    __shared__ double lds[256];
    double m = 0.0;
    for (std::uint32_t k = threadIdx.x; k < n; k += blockDim.x) {
        m = fmax(m, partial[k]);
    }
    lds[threadIdx.x] = m;
    __syncthreads();
    for (unsigned w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            lds[threadIdx.x] = fmax(lds[threadIdx.x], lds[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        *out = lds[0];
    }
}

// ---- launchers ---------------------------------------------------------------------------------
namespace {
inline std::uint32_t blocks_for(std::uint64_t n, std::uint32_t block) {
    return static_cast<std::uint32_t>((n + block - 1) / block);
}

// `lds`: bytes of LDS each workgroup reserves without using them, which caps how many
// workgroups share a compute unit (0: no cap).
template <class T, bool kLocal>
hipError_t step_bulk(const Lattice3<typename T::R>& p, int block, int cpt, std::uint32_t lds, hipStream_t s) {
    const std::uint32_t per = static_cast<std::uint32_t>(block) * static_cast<std::uint32_t>(cpt);
    const std::uint32_t grid = blocks_for(p.cells, per);
    switch (cpt) {
        case 2:
            hipLaunchKernelGGL((d3q19_bulk<T, kLocal, 2>), dim3(grid), dim3(block), lds, s, p);
            break;
        case 4:
            hipLaunchKernelGGL((d3q19_bulk<T, kLocal, 4>), dim3(grid), dim3(block), lds, s, p);
            break;
        default:
            hipLaunchKernelGGL((d3q19_bulk<T, kLocal, 1>), dim3(grid), dim3(block), lds, s, p);
            break;
    }
    return hipGetLastError();
}
} // namespace

template <class T>
hipError_t d3q19_step_bulk(const Lattice3<typename T::R>& p, bool local, int block, int cpt, std::uint32_t lds,
                           hipStream_t s) {
    return local ? step_bulk<T, true>(p, block, cpt, lds, s) : step_bulk<T, false>(p, block, cpt, lds, s);
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

template <class T> hipError_t d3q19_step_side(const Lattice3<typename T::R>& p, bool local, hipStream_t s) {
    if (local || p.ns == 0) {
        return hipSuccess;
    }
    hipLaunchKernelGGL((d3q19_side<T>), dim3(blocks_for(p.ns, 256)), dim3(256), 0, s, p);
    return hipGetLastError();
}

hipError_t d3q19_tick(std::uint32_t* clock, hipStream_t s) {
    hipLaunchKernelGGL(d3q19_tick_kernel, dim3(1), dim3(64), 0, s, clock);
    return hipGetLastError();
}

template <class T> hipError_t d3q19_step_fused(const Lattice3<typename T::R>& p, bool local, hipStream_t s) {
    const std::uint32_t grid = blocks_for(p.cells, 256);
    if (local) {
        hipLaunchKernelGGL((d3q19_fused<T, true>), dim3(grid), dim3(256), 0, s, p);
    } else {
        hipLaunchKernelGGL((d3q19_fused<T, false>), dim3(grid), dim3(256), 0, s, p);
    }
    return hipGetLastError();
}

template <class T>
hipError_t d3q19_step(const Lattice3<typename T::R>& p, bool local, int block, int cpt, std::uint32_t lds,
                      hipStream_t s) {
    hipError_t e = d3q19_step_bulk<T>(p, local, block, cpt, lds, s);
    if (e == hipSuccess) {
        e = d3q19_step_side<T>(p, local, s);
    }
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

template <class T>
hipError_t d3q19_sample(const Lattice3<typename T::R>& p, int phase, const double force[3], const std::uint32_t* cells,
                        const std::int32_t* slots, std::uint32_t n, double* out, hipStream_t s) {
    if (n == 0) {
        return hipSuccess;
    }
    hipLaunchKernelGGL((d3q19_sample<T>), dim3(blocks_for(n, 256)), dim3(256), 0, s, p, phase, force[0], force[1],
                       force[2], cells, slots, n, out);
    return hipGetLastError();
}

template <class T>
hipError_t d3q19_max_speed(const Lattice3<typename T::R>& p, int phase, const double force[3], double* partial,
                           double* result, hipStream_t s) {
    const std::uint32_t nc = blocks_for(p.cells, 256);
    hipLaunchKernelGGL((d3q19_speed_cells<T>), dim3(nc), dim3(256), 0, s, p, phase, force[0], force[1], force[2],
                       partial);
    std::uint32_t n = nc;
    if (p.nb > 0) {
        const std::uint32_t nb = blocks_for(p.nb, 256);
        hipLaunchKernelGGL((d3q19_speed_boundary<T>), dim3(nb), dim3(256), 0, s, p, phase, force[0], force[1],
                           force[2], partial + nc);
        n += nb;
    }
    hipLaunchKernelGGL(reduce_max, dim3(1), dim3(256), 0, s, partial, n, result);
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
    template hipError_t d3q19_step<T>(const Lattice3<T::R>&, bool, int, int, std::uint32_t, hipStream_t);    \
    template hipError_t d3q19_step_bulk<T>(const Lattice3<T::R>&, bool, int, int, std::uint32_t, hipStream_t); \
    template hipError_t d3q19_step_boundary<T>(const Lattice3<T::R>&, bool, hipStream_t);                    \
    template hipError_t d3q19_step_side<T>(const Lattice3<T::R>&, bool, hipStream_t);                        \
    template hipError_t d3q19_step_fused<T>(const Lattice3<T::R>&, bool, hipStream_t);                       \
    template hipError_t d3q19_initialize<T>(const Lattice3<T::R>&, int, const double*, const double*,       \
                                            std::uint64_t, double, hipStream_t);                             \
    template hipError_t d3q19_populations<T>(const Lattice3<T::R>&, int, double*, hipStream_t);              \
    template hipError_t d3q19_moments<T>(const Lattice3<T::R>&, int, const double*, double*, hipStream_t);   \
    template hipError_t d3q19_sample<T>(const Lattice3<T::R>&, int, const double*, const std::uint32_t*,      \
                                        const std::int32_t*, std::uint32_t, double*, hipStream_t);            \
    template hipError_t d3q19_max_speed<T>(const Lattice3<T::R>&, int, const double*, double*, double*,       \
                                           hipStream_t);                                                      \
    template hipError_t d3q19_mass<T>(const Lattice3<T::R>&, int, double*, double*, hipStream_t);            \
    template hipError_t d3q19_force_sum<T>(const Lattice3<T::R>&, int, const std::uint32_t*,                 \
                                           const std::uint8_t*, const double*, const std::uint32_t*,         \
                                           std::uint32_t, double*, double*, hipStream_t);

LBM_INSTANTIATE3(StoreF32)
LBM_INSTANTIATE3(StoreF16)
LBM_INSTANTIATE3(StoreBF16)
LBM_INSTANTIATE3(StoreF64)

} // namespace lbm
