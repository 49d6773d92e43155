// Batched D3Q19 engine: many resident lattices of uneven size advanced together. Every launch
// covers all lattices of one storage format: a block-to-lattice table (prefix sums of each
// lattice's blocks) lets one grid sweep every lattice, and each lattice keeps its own A-A parity
// in device memory, flipped by a one-block kernel after every step, so a chunk of
// `sample_every` steps plus the force sample is parameter-free and is captured once into a
// hipGraph and replayed. Per-lattice arithmetic, including the FP64 force reduction (fixed
// chunks of that lattice's links, then a fixed-order sum), never depends on the other
// lattices, so a lattice's results are the same alone and in any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "lbm3d_step.h"
#include "lbm_device.h"
#include "lbm_host.h"
#include "lbm_internal.h"
#include "lbm_kernels.h"

using lbm::host::fail;
using lbm::host::ok;

namespace lbm {

namespace {

using L3 = D3Q19;

// Link table of one lattice for the force sample.
struct LinkTable {
    const std::uint32_t* cell;
    const std::uint8_t* dir;
    const double* q;
    const std::uint32_t* slot;
    std::uint32_t count;
};

// Largest k with start[k] <= tile (start[0] = 0, start strictly covers [0, total)).
__device__ __forceinline__ std::uint32_t owner(const std::uint32_t* __restrict__ start, std::uint32_t n,
                                               std::uint32_t tile) {
    std::uint32_t lo = 0;
    std::uint32_t hi = n; // start[n] = total > tile
    while (hi - lo > 1) {
        const std::uint32_t mid = (lo + hi) >> 1;
        if (start[mid] <= tile) {
            lo = mid;
        } else {
            hi = mid;
        }
    }
    return lo;
}

template <class T, int V>
__global__ __launch_bounds__(256) void batch_bulk(const Lattice3<typename T::R>* __restrict__ lat,
                                                  const std::uint32_t* __restrict__ start, std::uint32_t n,
                                                  const std::uint32_t* __restrict__ parity) {
    const std::uint32_t l = owner(start, n, blockIdx.x);
    const Lattice3<typename T::R>& p = lat[l];
    const std::uint32_t c0 = ((blockIdx.x - start[l]) * blockDim.x + threadIdx.x) * V;
    if ((parity[l] & 1U) != 0) {
        d3::bulk_run<T, true, V>(p, c0);
    } else {
        d3::bulk_run<T, false, V>(p, c0);
    }
}

template <class T>
__global__ __launch_bounds__(256) void batch_boundary(const Lattice3<typename T::R>* __restrict__ lat,
                                                      const std::uint32_t* __restrict__ start, std::uint32_t n,
                                                      const std::uint32_t* __restrict__ parity) {
    const std::uint32_t l = owner(start, n, blockIdx.x);
    const Lattice3<typename T::R>& p = lat[l];
    const std::uint32_t b = (blockIdx.x - start[l]) * blockDim.x + threadIdx.x;
    if (b >= p.nb) {
        return;
    }
    if ((parity[l] & 1U) != 0) {
        d3::boundary_node<T, true>(p, b);
    } else {
        d3::boundary_node<T, false>(p, b);
    }
}

// B step of the side nodes of the lattices whose step is a B step (the bulk kernel updates them
// in A steps).
template <class T>
__global__ __launch_bounds__(256) void batch_side(const Lattice3<typename T::R>* __restrict__ lat,
                                                  const std::uint32_t* __restrict__ start, std::uint32_t n,
                                                  const std::uint32_t* __restrict__ parity) {
    const std::uint32_t l = owner(start, n, blockIdx.x);
    const Lattice3<typename T::R>& p = lat[l];
    const std::uint32_t s = (blockIdx.x - start[l]) * blockDim.x + threadIdx.x;
    if (s < p.ns && (parity[l] & 1U) == 0U) {
        d3::side_node_b<T>(p, p.scell[s]);
    }
}

// parity[l] holds lattice l's step count (mod 2^32): bit 0 its A-A parity, the whole word the
// step clock of the 16-bit formats' rounding dither.
__global__ void batch_advance(std::uint32_t* parity, std::uint32_t n) {
    for (std::uint32_t l = threadIdx.x; l < n; l += blockDim.x) {
        parity[l] += 1U;
    }
}

template <int K>
__device__ __forceinline__ void block_sum_b(double (&v)[K], double* partial) {
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

// Momentum exchange of link k of lattice l (the arithmetic of d3q19_force in lbm3d_kernels.hip).
template <class T>
__global__ __launch_bounds__(256) void batch_force(const Lattice3<typename T::R>* __restrict__ lat,
                                                   const LinkTable* __restrict__ links,
                                                   const std::uint32_t* __restrict__ start, std::uint32_t n,
                                                   const std::uint32_t* __restrict__ parity, double* partial) {
    const std::uint32_t l = owner(start, n, blockIdx.x);
    const Lattice3<typename T::R>& p = lat[l];
    const LinkTable& t = links[l];
    const std::uint32_t k = (blockIdx.x - start[l]) * blockDim.x + threadIdx.x;
    const int phase = static_cast<int>(parity[l] & 1U);
    double v[6] = {0, 0, 0, 0, 0, 0};
    if (k < t.count) {
        const auto* f = static_cast<const typename T::S*>(p.f);
        const std::uint32_t cell = t.cell[k];
        const int j = t.dir[k];
        const double q = t.q[k];
        const double qr = static_cast<double>(static_cast<typename T::R>(q));
        const std::uint32_t b = t.slot[k];
        const auto* st = p.stash + (phase ? 19U * p.bstride : 0U);
        const int o = L3::opp(j);
        const double out = static_cast<double>(st[o * p.bstride + b]) + (T::shifted ? L3::w(o) : 0.0);
        double back;
        if (qr == 0.5) {
            back = out;
        } else if (qr < 0.5) {
            double farther;
            if (phase == 0) {
                const int off = d3::off3(j, static_cast<int>(p.nx), static_cast<int>(p.nx * p.ny));
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
    block_sum_b<6>(v, partial);
}

// One block per lattice: the fixed-order sum of its link-block partials (the arithmetic of
// reduce_partials<6>), written to out[l].
__global__ __launch_bounds__(256) void batch_force_reduce(const double* partial, const std::uint32_t* start,
                                                          double* out) {
    __shared__ double lds[256][6];
    const std::uint32_t l = blockIdx.x;
    const std::uint32_t first = start[l];
    const std::uint32_t count = start[l + 1] - first;
    double s[6] = {0, 0, 0, 0, 0, 0};
    for (std::uint32_t r = threadIdx.x; r < count; r += blockDim.x) {
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            s[k] += partial[static_cast<std::size_t>(first + r) * 6 + k];
        }
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        lds[threadIdx.x][k] = s[k];
    }
    __syncthreads();
    for (unsigned width = blockDim.x / 2; width > 0; width >>= 1) {
        if (threadIdx.x < width) {
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                lds[threadIdx.x][k] += lds[threadIdx.x + width][k];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) {
        out[static_cast<std::size_t>(l) * 6 + threadIdx.x] = lds[0][threadIdx.x];
    }
}

inline std::uint32_t ceil_div(std::uint64_t a, std::uint64_t b) { return static_cast<std::uint32_t>((a + b - 1) / b); }

// Device arrays of one storage group (lattices of one population format).
struct Group {
    int storage = LBM_STORE_F32;
    std::vector<int> members; // indices into Batch::entries
    void* lat = nullptr;      // Lattice3<R>[n]
    LinkTable* links = nullptr;
    std::uint32_t* bulk_start = nullptr;
    std::uint32_t* bnd_start = nullptr;
    std::uint32_t* side_start = nullptr;
    std::uint32_t* link_start = nullptr;
    std::uint32_t* parity = nullptr;
    double* partial = nullptr;
    double* force = nullptr; // [n][6]
    std::uint32_t bulk_blocks = 0;
    std::uint32_t bnd_blocks = 0;
    std::uint32_t side_blocks = 0;
    std::uint32_t link_blocks = 0;
    int block = 256;
    int cpt = 1;
    std::uint32_t lds = 0; // dynamic LDS reserved per bulk block: the occupancy cap
};

} // namespace

} // namespace lbm

struct lbm_batch {
    struct Entry {
        int id = 0;
        lbm3d* h = nullptr;
        std::size_t bytes = 0;
        std::vector<double> time;
        std::vector<double> force; // 6 per sample
    };
    hipStream_t stream = nullptr;
    std::size_t budget = 0;
    std::size_t used = 0;
    std::uint32_t sample_every = 10;
    int next_id = 1;
    std::vector<Entry> entries;
    std::vector<lbm::Group> groups;
    bool dirty = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<void*> device; // allocations of the groups
    double* pinned = nullptr;
    std::size_t pinned_doubles = 0;
    std::map<int, std::pair<int, int>> launch; // storage -> (block, cpt)
    std::map<int, int> occupancy;              // storage -> bulk groups per CU (0: no cap)
    std::uint32_t lds_per_cu = 0;
    std::uint32_t lds_per_block = 0;
};

namespace {

using lbm::Group;

int hip_fail_b(const char* what, hipError_t e) {
    return fail(LBM_EDEVICE, std::string("lbm batch: ") + what + ": " + hipGetErrorString(e));
}

void release_groups(lbm_batch& b) {
    if (b.exec != nullptr) {
        (void)hipGraphExecDestroy(b.exec);
        b.exec = nullptr;
    }
    if (b.graph != nullptr) {
        (void)hipGraphDestroy(b.graph);
        b.graph = nullptr;
    }
    for (void* p : b.device) {
        (void)hipFree(p);
    }
    b.device.clear();
    b.groups.clear();
}

template <typename P> int dalloc(lbm_batch& b, P** out, std::size_t bytes) {
    void* p = nullptr;
    const hipError_t e = hipMalloc(&p, bytes == 0 ? 8 : bytes);
    if (e != hipSuccess) {
        return fail(LBM_ENOMEM, std::string("lbm batch: device tables: ") + hipGetErrorString(e));
    }
    b.device.push_back(p);
    *out = static_cast<P*>(p);
    return ok();
}

template <class T> int fill_group(lbm_batch& b, Group& g) {
    using R = typename T::R;
    const std::size_t n = g.members.size();
    std::vector<lbm::Lattice3<R>> lat(n);
    std::vector<lbm::LinkTable> links(n);
    std::vector<std::uint32_t> bulk(n + 1, 0), bnd(n + 1, 0), side(n + 1, 0), lnk(n + 1, 0), parity(n, 0);
    const std::uint32_t per = static_cast<std::uint32_t>(g.block * g.cpt);
    for (std::size_t k = 0; k < n; ++k) {
        lbm3d* h = b.entries[static_cast<std::size_t>(g.members[k])].h;
        lat[k] = lbm::internal::descriptor3<T>(*h);
        const lbm::internal::LinkView v = lbm::internal::link_view(*h);
        links[k] = lbm::LinkTable{v.cell, v.dir, v.q, v.slot, v.count};
        bulk[k + 1] = bulk[k] + lbm::ceil_div(lat[k].cells, per);
        bnd[k + 1] = bnd[k] + std::max<std::uint32_t>(1, lbm::ceil_div(lat[k].nb, 256));
        side[k + 1] = side[k] + std::max<std::uint32_t>(1, lbm::ceil_div(lat[k].ns, 256));
        lnk[k + 1] = lnk[k] + std::max<std::uint32_t>(1, lbm::ceil_div(v.count, 256));
        parity[k] = static_cast<std::uint32_t>(lbm3d_steps(h)); // the step clock too (mod 2^32)
    }
    g.bulk_blocks = bulk[n];
    g.bnd_blocks = bnd[n];
    g.side_blocks = side[n];
    g.link_blocks = lnk[n];
    int rc = LBM_OK;
    // (the descriptors' clocks point into g.parity once it is allocated, below)
    if ((rc = dalloc(b, reinterpret_cast<char**>(&g.lat), n * sizeof(lbm::Lattice3<R>))) ||
        (rc = dalloc(b, &g.links, n * sizeof(lbm::LinkTable))) || (rc = dalloc(b, &g.bulk_start, (n + 1) * 4)) ||
        (rc = dalloc(b, &g.bnd_start, (n + 1) * 4)) || (rc = dalloc(b, &g.side_start, (n + 1) * 4)) ||
        (rc = dalloc(b, &g.link_start, (n + 1) * 4)) ||
        (rc = dalloc(b, &g.parity, n * 4)) || (rc = dalloc(b, &g.partial, static_cast<std::size_t>(g.link_blocks) * 48)) ||
        (rc = dalloc(b, &g.force, n * 48))) {
        return rc;
    }
    for (std::size_t k = 0; k < n; ++k) {
        lat[k].clock = g.parity + k;
    }
    const auto up = [&](void* dst, const void* src, std::size_t bytes) {
        return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, b.stream);
    };
    hipError_t e = up(g.lat, lat.data(), n * sizeof(lbm::Lattice3<R>));
    if (e == hipSuccess) e = up(g.links, links.data(), n * sizeof(lbm::LinkTable));
    if (e == hipSuccess) e = up(g.bulk_start, bulk.data(), (n + 1) * 4);
    if (e == hipSuccess) e = up(g.bnd_start, bnd.data(), (n + 1) * 4);
    if (e == hipSuccess) e = up(g.side_start, side.data(), (n + 1) * 4);
    if (e == hipSuccess) e = up(g.link_start, lnk.data(), (n + 1) * 4);
    if (e == hipSuccess) e = up(g.parity, parity.data(), n * 4);
    if (e == hipSuccess) e = hipStreamSynchronize(b.stream);
    return e == hipSuccess ? ok() : hip_fail_b("table upload", e);
}

template <class T> void enqueue_step(Group& g, hipStream_t s) {
    using R = typename T::R;
    const auto* lat = static_cast<const lbm::Lattice3<R>*>(g.lat);
    const auto n = static_cast<std::uint32_t>(g.members.size());
    switch (g.cpt) {
        case 2:
            hipLaunchKernelGGL((lbm::batch_bulk<T, 2>), dim3(g.bulk_blocks), dim3(g.block), g.lds, s, lat, g.bulk_start, n,
                               g.parity);
            break;
        case 4:
            hipLaunchKernelGGL((lbm::batch_bulk<T, 4>), dim3(g.bulk_blocks), dim3(g.block), g.lds, s, lat, g.bulk_start, n,
                               g.parity);
            break;
        default:
            hipLaunchKernelGGL((lbm::batch_bulk<T, 1>), dim3(g.bulk_blocks), dim3(g.block), g.lds, s, lat, g.bulk_start, n,
                               g.parity);
            break;
    }
    hipLaunchKernelGGL((lbm::batch_side<T>), dim3(g.side_blocks), dim3(256), 0, s, lat, g.side_start, n, g.parity);
    hipLaunchKernelGGL((lbm::batch_boundary<T>), dim3(g.bnd_blocks), dim3(256), 0, s, lat, g.bnd_start, n, g.parity);
    hipLaunchKernelGGL(lbm::batch_advance, dim3(1), dim3(256), 0, s, g.parity, n);
}

template <class T> void enqueue_force(Group& g, hipStream_t s) {
    using R = typename T::R;
    const auto* lat = static_cast<const lbm::Lattice3<R>*>(g.lat);
    const auto n = static_cast<std::uint32_t>(g.members.size());
    hipLaunchKernelGGL((lbm::batch_force<T>), dim3(g.link_blocks), dim3(256), 0, s, lat, g.links, g.link_start, n,
                       g.parity, g.partial);
    hipLaunchKernelGGL(lbm::batch_force_reduce, dim3(n), dim3(256), 0, s, g.partial, g.link_start, g.force);
}

template <class Fn> void for_store(int storage, Fn&& fn) {
    switch (storage) {
        case LBM_STORE_F16S:
            fn.template operator()<lbm::StoreF16>();
            break;
        case LBM_STORE_BF16S:
            fn.template operator()<lbm::StoreBF16>();
            break;
        case LBM_STORE_F64:
            fn.template operator()<lbm::StoreF64>();
            break;
        default:
            fn.template operator()<lbm::StoreF32>();
            break;
    }
}

// Rebuilds the groups and captures one chunk (sample_every steps + the force sample) into a graph.
int rebuild(lbm_batch& b) {
    release_groups(b);
    std::map<int, Group> by_store;
    for (std::size_t k = 0; k < b.entries.size(); ++k) {
        const int st = lbm::internal::storage_of(*b.entries[k].h);
        Group& g = by_store[st];
        g.storage = st;
        g.cpt = lbm::internal::default_cells_per_thread(st);
        g.members.push_back(static_cast<int>(k));
    }
    for (auto& [st, g] : by_store) {
        const auto it = b.launch.find(st);
        if (it != b.launch.end()) {
            g.block = it->second.first;
            g.cpt = it->second.second;
        }
        const auto cap = b.occupancy.find(st);
        if (cap != b.occupancy.end() && cap->second > 0) {
            g.lds = b.lds_per_cu / static_cast<std::uint32_t>(cap->second) / 512U * 512U;
        }
        b.groups.push_back(g);
    }
    for (Group& g : b.groups) {
        int rc = LBM_OK;
        for_store(g.storage, [&]<class T>() { rc = fill_group<T>(b, g); });
        if (rc != LBM_OK) {
            return rc;
        }
    }
    if (b.entries.empty()) {
        b.dirty = false;
        return ok();
    }
    hipError_t e = hipStreamBeginCapture(b.stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        return hip_fail_b("begin capture", e);
    }
    for (std::uint32_t s = 0; s < b.sample_every; ++s) {
        for (Group& g : b.groups) {
            for_store(g.storage, [&]<class T>() { enqueue_step<T>(g, b.stream); });
        }
    }
    for (Group& g : b.groups) {
        for_store(g.storage, [&]<class T>() { enqueue_force<T>(g, b.stream); });
    }
    e = hipStreamEndCapture(b.stream, &b.graph);
    if (e != hipSuccess) {
        return hip_fail_b("end capture", e);
    }
    e = hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        b.exec = nullptr;
        return hip_fail_b("graph instantiate", e);
    }
    b.dirty = false; // only a complete capture clears it: any failure above rebuilds next run
    return ok();
}

// Host side of the chunks [0, chunks) that ran: samples (in group order) into the lattices'
// histories, then the step counts.
void collect(lbm_batch& b, std::uint64_t chunks) {
    const std::size_t n = b.entries.size();
    for (std::uint64_t c = 0; c < chunks; ++c) {
        std::size_t offset = 0;
        for (lbm::Group& g : b.groups) {
            for (std::size_t k = 0; k < g.members.size(); ++k) {
                lbm_batch::Entry& entry = b.entries[static_cast<std::size_t>(g.members[k])];
                const double* v = b.pinned + (c * n + offset + k) * 6;
                entry.time.push_back(static_cast<double>(lbm3d_steps(entry.h) + (c + 1) * b.sample_every));
                entry.force.insert(entry.force.end(), v, v + 6);
            }
            offset += g.members.size();
        }
    }
    for (auto& entry : b.entries) {
        lbm::internal::advance_steps(*entry.h, chunks * b.sample_every);
    }
}

// Frees a lattice of the batch: the graph and the group tables (whose descriptors point into
// the lattice) go first, after the stream has drained.
void drop_entry(lbm_batch& b, std::vector<lbm_batch::Entry>::iterator it) {
    (void)hipStreamSynchronize(b.stream);
    release_groups(b);
    b.used -= it->bytes;
    it->h->in_batch = false;
    lbm3d_destroy(it->h);
    b.entries.erase(it);
    b.dirty = true;
}

lbm_batch::Entry* find(lbm_batch& b, int id) {
    for (auto& e : b.entries) {
        if (e.id == id) {
            return &e;
        }
    }
    return nullptr;
}

} // namespace

extern "C" {

int lbm_batch_create(size_t budget_bytes, uint32_t sample_every, lbm_batch** out) {
    if (out == nullptr || sample_every == 0) {
        return fail(LBM_EINVAL, "lbm batch: null output or zero sample interval");
    }
    auto* b = new (std::nothrow) lbm_batch();
    if (b == nullptr) {
        return fail(LBM_ENOMEM, "lbm batch: host allocation failed");
    }
    b->budget = budget_bytes;
    b->sample_every = sample_every;
    int device = 0;
    int lds_cu = 0;
    int lds_block = 0;
    hipError_t e = hipGetDevice(&device);
    if (e == hipSuccess) {
        e = hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, device);
    }
    if (e == hipSuccess) {
        e = hipDeviceGetAttribute(&lds_block, hipDeviceAttributeMaxSharedMemoryPerBlock, device);
    }
    b->lds_per_cu = static_cast<std::uint32_t>(lds_cu > 0 ? lds_cu : 0);
    b->lds_per_block = static_cast<std::uint32_t>(lds_block > 0 ? lds_block : 0);
    if (e == hipSuccess) {
        e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    }
    if (e != hipSuccess) {
        delete b;
        return hip_fail_b("device query or stream", e);
    }
    *out = b;
    return ok();
}

int lbm_batch_destroy(lbm_batch* b) {
    if (b == nullptr) {
        return ok();
    }
    (void)hipStreamSynchronize(b->stream);
    release_groups(*b);
    for (auto& e : b->entries) {
        e.h->in_batch = false;
        lbm3d_destroy(e.h);
    }
    if (b->pinned != nullptr) {
        (void)hipHostFree(b->pinned);
    }
    (void)hipStreamDestroy(b->stream);
    delete b;
    return ok();
}

int lbm_batch_required_bytes(const lbm3d_config* config, const uint8_t* solid, const double* distance,
                             int32_t storage, size_t* bytes) {
    if (config == nullptr || bytes == nullptr) {
        return fail(LBM_EINVAL, "lbm batch: null argument");
    }
    if (std::string p = lbm::host::problem3(*config, solid, distance); !p.empty()) {
        return fail(LBM_EINVAL, p);
    }
    *bytes = lbm::internal::required_bytes3(*config, solid, distance, storage);
    return ok();
}

int lbm_batch_add(lbm_batch* b, const lbm3d_config* config, const uint8_t* solid, const double* distance,
                  int32_t storage, uint64_t seed, double amplitude, int32_t* id) {
    if (b == nullptr || config == nullptr || id == nullptr) {
        return fail(LBM_EINVAL, "lbm batch: null argument");
    }
    if (std::string p = lbm::host::problem3(*config, solid, distance); !p.empty()) {
        return fail(LBM_EINVAL, p);
    }
    const std::size_t need = lbm::internal::required_bytes3(*config, solid, distance, storage);
    if (b->budget != 0 && b->used + need > b->budget) {
        return fail(LBM_ENOMEM, "lbm batch: the lattice needs " + lbm::host::megabytes(need) + ", " +
                                    lbm::host::megabytes(b->budget - b->used) + " of the budget are free");
    }
    lbm3d* h = nullptr;
    int rc = lbm3d_create(config, solid, distance, storage, &h);
    if (rc != LBM_OK) {
        return rc;
    }
    if (amplitude != 0.0) {
        rc = lbm3d_initialize_perturbed(h, seed, amplitude);
        if (rc != LBM_OK) {
            lbm3d_destroy(h);
            return rc;
        }
    }
    h->in_batch = true;
    lbm_batch::Entry e;
    e.id = b->next_id++;
    e.h = h;
    e.bytes = need;
    b->used += need;
    b->entries.push_back(std::move(e));
    b->dirty = true;
    *id = b->entries.back().id;
    return ok();
}

int lbm_batch_remove(lbm_batch* b, int32_t id) {
    if (b == nullptr) {
        return fail(LBM_EINVAL, "lbm batch: null handle");
    }
    for (auto it = b->entries.begin(); it != b->entries.end(); ++it) {
        if (it->id == id) {
            drop_entry(*b, it);
            return ok();
        }
    }
    return fail(LBM_EINVAL, "lbm batch: unknown lattice id");
}

int lbm_batch_run(lbm_batch* b, uint64_t steps) {
    if (b == nullptr) {
        return fail(LBM_EINVAL, "lbm batch: null handle");
    }
    if (steps % b->sample_every != 0) {
        return fail(LBM_EINVAL, "lbm batch: steps must be a multiple of the sample interval");
    }
    if (b->entries.empty() || steps == 0) {
        return ok();
    }
    for (auto& e : b->entries) {
        if (int rc = lbm3d_synchronize(e.h); rc != LBM_OK) {
            return rc;
        }
    }
    if (b->dirty) {
        if (int rc = rebuild(*b); rc != LBM_OK) {
            return rc;
        }
    }
    const std::uint64_t chunks = steps / b->sample_every;
    const std::size_t n = b->entries.size();
    const std::size_t need = chunks * n * 6;
    if (need > b->pinned_doubles) {
        if (b->pinned != nullptr) {
            (void)hipHostFree(b->pinned);
            b->pinned = nullptr;
        }
        if (hipHostMalloc(reinterpret_cast<void**>(&b->pinned), need * sizeof(double), hipHostMallocDefault) !=
            hipSuccess) {
            b->pinned_doubles = 0;
            return fail(LBM_ENOMEM, "lbm batch: pinned sample buffer");
        }
        b->pinned_doubles = need;
    }
    // A chunk is one graph replay plus the copies of its force samples; `queued` counts the
    // chunks whose replay was enqueued (the device parity flips with each of them).
    std::uint64_t queued = 0;
    hipError_t e = hipSuccess;
    for (std::uint64_t c = 0; c < chunks && e == hipSuccess; ++c) {
        e = hipGraphLaunch(b->exec, b->stream);
        if (e != hipSuccess) {
            break;
        }
        ++queued;
        std::size_t offset = 0;
        for (lbm::Group& g : b->groups) {
            const std::size_t gn = g.members.size();
            if (e == hipSuccess) {
                e = hipMemcpyAsync(b->pinned + (c * n + offset) * 6, g.force, gn * 48, hipMemcpyDeviceToHost,
                                   b->stream);
            }
            offset += gn;
        }
    }
    const hipError_t drained = hipStreamSynchronize(b->stream);
    if (e != hipSuccess || drained != hipSuccess) {
        // Keep the host step counts on the device parity: the replays that were enqueued ran
        // (if the stream drained). The tables are uploaded again before the next run.
        if (drained == hipSuccess) {
            collect(*b, queued);
        }
        b->dirty = true;
        return hip_fail_b(e != hipSuccess ? "graph launch" : "run", e != hipSuccess ? e : drained);
    }
    collect(*b, chunks);
    return ok();
}

int lbm_batch_run_timed(lbm_batch* b, uint64_t steps, lbm_batch_timing* timing) {
    if (b == nullptr || timing == nullptr) {
        return fail(LBM_EINVAL, "lbm batch: null argument");
    }
    if (steps % b->sample_every != 0) {
        return fail(LBM_EINVAL, "lbm batch: steps must be a multiple of the sample interval");
    }
    *timing = lbm_batch_timing{};
    if (b->entries.empty() || steps == 0) {
        return ok();
    }
    for (auto& e : b->entries) {
        if (int rc = lbm3d_synchronize(e.h); rc != LBM_OK) {
            return rc;
        }
    }
    if (b->dirty) {
        if (int rc = rebuild(*b); rc != LBM_OK) {
            return rc;
        }
    }
    const std::uint64_t chunks = steps / b->sample_every;
    const std::size_t n = b->entries.size();
    const std::size_t need = chunks * n * 6;
    if (need > b->pinned_doubles) {
        if (b->pinned != nullptr) {
            (void)hipHostFree(b->pinned);
            b->pinned = nullptr;
        }
        if (hipHostMalloc(reinterpret_cast<void**>(&b->pinned), need * sizeof(double), hipHostMallocDefault) !=
            hipSuccess) {
            b->pinned_doubles = 0;
            return fail(LBM_ENOMEM, "lbm batch: pinned sample buffer");
        }
        b->pinned_doubles = need;
    }
    // Marks: per step and group, before the bulk kernel, after the side kernel (before the
    // boundary kernel) and after the boundary kernel; per chunk,
    // before and after the force kernels; one at the very end.
    const std::size_t groups = b->groups.size();
    const std::size_t per_chunk = b->sample_every * groups * 3U + 2U;
    std::vector<hipEvent_t> ev(chunks * per_chunk + 1, nullptr);
    hipError_t e = hipSuccess;
    for (auto& x : ev) {
        if (e == hipSuccess) {
            e = hipEventCreate(&x);
        }
    }
    std::uint64_t queued = 0;
    std::size_t m = 0;
    for (std::uint64_t c = 0; c < chunks && e == hipSuccess; ++c) {
        for (std::uint32_t s = 0; s < b->sample_every; ++s) {
            for (lbm::Group& g : b->groups) {
                const int st = g.storage;
                (void)hipEventRecord(ev[m++], b->stream);
                for_store(st, [&]<class T>() {
                    using R = typename T::R;
                    const auto* lat = static_cast<const lbm::Lattice3<R>*>(g.lat);
                    const auto count = static_cast<std::uint32_t>(g.members.size());
                    switch (g.cpt) {
                        case 2:
                            hipLaunchKernelGGL((lbm::batch_bulk<T, 2>), dim3(g.bulk_blocks), dim3(g.block), g.lds,
                                               b->stream, lat, g.bulk_start, count, g.parity);
                            break;
                        case 4:
                            hipLaunchKernelGGL((lbm::batch_bulk<T, 4>), dim3(g.bulk_blocks), dim3(g.block), g.lds,
                                               b->stream, lat, g.bulk_start, count, g.parity);
                            break;
                        default:
                            hipLaunchKernelGGL((lbm::batch_bulk<T, 1>), dim3(g.bulk_blocks), dim3(g.block), g.lds,
                                               b->stream, lat, g.bulk_start, count, g.parity);
                            break;
                    }
                    hipLaunchKernelGGL((lbm::batch_side<T>), dim3(g.side_blocks), dim3(256), 0, b->stream, lat,
                                       g.side_start, count, g.parity);
                    (void)hipEventRecord(ev[m++], b->stream);
                    hipLaunchKernelGGL((lbm::batch_boundary<T>), dim3(g.bnd_blocks), dim3(256), 0, b->stream, lat,
                                       g.bnd_start, count, g.parity);
                    (void)hipEventRecord(ev[m++], b->stream);
                    hipLaunchKernelGGL(lbm::batch_advance, dim3(1), dim3(256), 0, b->stream, g.parity, count);
                });
            }
        }
        (void)hipEventRecord(ev[m++], b->stream);
        for (lbm::Group& g : b->groups) {
            for_store(g.storage, [&]<class T>() { enqueue_force<T>(g, b->stream); });
        }
        (void)hipEventRecord(ev[m++], b->stream);
        std::size_t offset = 0;
        for (lbm::Group& g : b->groups) {
            const std::size_t gn = g.members.size();
            if (e == hipSuccess) {
                e = hipMemcpyAsync(b->pinned + (c * n + offset) * 6, g.force, gn * 48, hipMemcpyDeviceToHost,
                                   b->stream);
            }
            offset += gn;
        }
        if (e == hipSuccess) {
            e = hipGetLastError();
        }
        if (e == hipSuccess) {
            ++queued;
        }
    }
    if (e == hipSuccess) {
        e = hipEventRecord(ev[m], b->stream);
    }
    const hipError_t drained = hipStreamSynchronize(b->stream);
    if (e == hipSuccess && drained == hipSuccess) {
        const auto span = [&](std::size_t i, std::size_t j) {
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, ev[i], ev[j]);
            return static_cast<double>(ms);
        };
        std::size_t k = 0;
        double force_total = 0.0;
        for (std::uint64_t c = 0; c < chunks; ++c) {
            for (std::uint32_t s = 0; s < b->sample_every; ++s) {
                for (lbm::Group& g : b->groups) {
                    timing->bulk_ms[g.storage] += span(k, k + 1);
                    timing->boundary_ms[g.storage] += span(k + 1, k + 2);
                    k += 3;
                }
            }
            force_total += span(k, k + 1);
            k += 2;
        }
        timing->step_ms = span(0, m) / static_cast<double>(steps);
        timing->force_ms = force_total / static_cast<double>(chunks);
        for (int st = 0; st < 4; ++st) {
            timing->bulk_ms[st] /= static_cast<double>(steps);
            timing->boundary_ms[st] /= static_cast<double>(steps);
        }
        for (const lbm::Group& g : b->groups) {
            for (const int member : g.members) {
                std::size_t bulk = 0;
                std::size_t side = 0;
                std::size_t listed = 0;
                (void)lbm3d_stream_counts(b->entries[static_cast<std::size_t>(member)].h, &bulk, &side, &listed);
                timing->bulk_nodes[g.storage] += bulk + side;
                timing->boundary_nodes[g.storage] += listed;
            }
        }
    }
    for (auto& x : ev) {
        if (x != nullptr) {
            (void)hipEventDestroy(x);
        }
    }
    if (e != hipSuccess || drained != hipSuccess) {
        if (drained == hipSuccess) {
            collect(*b, queued);
        }
        b->dirty = true;
        return hip_fail_b("timed run", e != hipSuccess ? e : drained);
    }
    collect(*b, chunks);
    return ok();
}

int lbm_batch_count(const lbm_batch* b, size_t* count) {
    if (b == nullptr || count == nullptr) {
        return fail(LBM_EINVAL, "lbm batch: null argument");
    }
    *count = b->entries.size();
    return ok();
}

int lbm_batch_lattice(lbm_batch* b, int32_t id, lbm3d** h) {
    if (b == nullptr || h == nullptr) {
        return fail(LBM_EINVAL, "lbm batch: null argument");
    }
    lbm_batch::Entry* e = find(*b, id);
    if (e == nullptr) {
        return fail(LBM_EINVAL, "lbm batch: unknown lattice id");
    }
    *h = e->h;
    return ok();
}

int lbm_batch_history(lbm_batch* b, int32_t id, double* time, double* force, size_t capacity, size_t* count) {
    if (b == nullptr) {
        return fail(LBM_EINVAL, "lbm batch: null handle");
    }
    lbm_batch::Entry* e = find(*b, id);
    if (e == nullptr) {
        return fail(LBM_EINVAL, "lbm batch: unknown lattice id");
    }
    const std::size_t n = e->time.size();
    for (std::size_t k = 0; k < n && k < capacity; ++k) {
        if (time != nullptr) time[k] = e->time[k];
        if (force != nullptr) std::memcpy(force + 6 * k, e->force.data() + 6 * k, 6 * sizeof(double));
    }
    if (count != nullptr) {
        *count = n;
    }
    return ok();
}

int lbm_batch_bytes(const lbm_batch* b, size_t* used, size_t* budget) {
    if (b == nullptr) {
        return fail(LBM_EINVAL, "lbm batch: null handle");
    }
    if (used != nullptr) *used = b->used;
    if (budget != nullptr) *budget = b->budget;
    return ok();
}

int lbm_batch_set_launch(lbm_batch* b, int32_t storage, int32_t block, int32_t cells_per_thread) {
    if (b == nullptr || (block != 64 && block != 128 && block != 192 && block != 256) ||
        (cells_per_thread != 1 && cells_per_thread != 2 && cells_per_thread != 4)) {
        return fail(LBM_EINVAL, "lbm batch: block must be 64..256 (multiple of 64), cells per thread 1, 2 or 4");
    }
    b->launch[storage] = {block, cells_per_thread};
    b->dirty = true;
    return ok();
}

int lbm_batch_set_occupancy(lbm_batch* b, int32_t storage, int32_t groups_per_cu) {
    if (b == nullptr || storage < 0 || storage > 3) {
        return fail(LBM_EINVAL, "lbm batch: null handle or unknown storage");
    }
    if (groups_per_cu < 0 || groups_per_cu > 32) {
        return fail(LBM_EINVAL, "lbm batch: workgroups per compute unit must be 0 (no cap) to 32");
    }
    if (groups_per_cu > 0 &&
        (b->lds_per_cu == 0 || b->lds_per_cu / static_cast<std::uint32_t>(groups_per_cu) / 512U * 512U >
                                   b->lds_per_block)) {
        return fail(LBM_EINVAL, "lbm batch: the device's LDS cannot express a cap of " +
                                    std::to_string(groups_per_cu) + " workgroups per compute unit");
    }
    b->occupancy[storage] = groups_per_cu;
    b->dirty = true;
    return ok();
}

int lbm_batch_autotune(lbm_batch* b, uint32_t steps, lbm_launch_choice* best) {
    if (b == nullptr || best == nullptr) {
        return fail(LBM_EINVAL, "lbm batch: null argument");
    }
    if (b->stream != nullptr) {
        const hipError_t e = hipStreamSynchronize(b->stream);
        if (e != hipSuccess) {
            return hip_fail_b("autotune", e);
        }
    }
    for (int st = 0; st < 4; ++st) {
        best[st] = lbm_launch_choice{LBM_SCHEDULE_SERIAL, 0, 0, 0, 0, 0.0};
        // The largest lattice of this storage stands for the group, copied into a scratch
        // handle (halfway walls: the bulk kernel, which the shape tunes, is the same).
        const lbm_batch::Entry* pick = nullptr;
        for (const auto& e : b->entries) {
            if (e.h->storage == st && (pick == nullptr || e.h->cells > pick->h->cells)) {
                pick = &e;
            }
        }
        if (pick == nullptr) {
            continue;
        }
        lbm3d_config c = pick->h->cfg;
        c.solid_wall = LBM_SOLID_HALFWAY;
        lbm3d* scratch = nullptr;
        int rc = lbm3d_create(&c, pick->h->solid.data(), nullptr, st, &scratch);
        if (rc == LBM_OK) {
            rc = lbm3d_autotune(scratch, steps, 1, nullptr, 0, nullptr, &best[st]);
        }
        if (scratch != nullptr) {
            (void)lbm3d_destroy(scratch);
        }
        if (rc == LBM_OK) {
            rc = lbm_batch_set_launch(b, st, best[st].block, best[st].cells_per_thread);
        }
        if (rc == LBM_OK) {
            // The caps tuned on one lattice are not applied: on a batch of uneven lattices they
            // measured slower (48 wing sections: 4 groups per CU 101.8 s, uncapped 99.6 s).
            best[st].groups_a = 0;
            best[st].groups_b = 0;
            rc = lbm_batch_set_occupancy(b, st, 0);
        }
        if (rc != LBM_OK) {
            return rc;
        }
    }
    return ok();
}

int lbm_batch_set_inlet(lbm_batch* b, int32_t id, const double velocity[3]) {
    if (b == nullptr) {
        return fail(LBM_EINVAL, "lbm batch: null handle");
    }
    lbm_batch::Entry* e = find(*b, id);
    if (e == nullptr) {
        return fail(LBM_EINVAL, "lbm batch: unknown lattice id");
    }
    e->h->in_batch = false;
    const int rc = lbm3d_set_inlet(e->h, velocity);
    e->h->in_batch = true;
    if (rc == LBM_OK) {
        b->dirty = true; // the descriptors carry the inlet velocity
    }
    return rc;
}

} // extern "C"
