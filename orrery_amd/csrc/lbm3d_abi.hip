// C ABI of the D3Q19 device solver (include/orrery_lbm.h): the handle that replaces GpuLbm3D
// and make_gpu_solver (src/aero/lbm3d/lbm3d_gpu.cpp:151-513).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "lbm_device.h"
#include "lbm_host.h"
#include "lbm_internal.h"
#include "lbm_kernels.h"

using lbm::host::fail;
using lbm::host::ok;

namespace {

using lbm::internal::lattice_bytes;
using lbm::internal::real_bytes;
using lbm::internal::store_bytes;

int hip_fail(const char* what, hipError_t e) {
    return fail(LBM_EDEVICE, std::string("lbm3d gpu: ") + what + ": " + hipGetErrorString(e));
}

template <class T> lbm::Lattice3<typename T::R> descriptor(const lbm3d& h) { return lbm::internal::descriptor3<T>(h); }

// Runs `fn.template operator()<T>()` for the storage format of the handle.
template <class Fn> int with_store(const lbm3d& h, Fn&& fn) {
    switch (h.storage) {
        case LBM_STORE_F16S:
            return fn.template operator()<lbm::StoreF16>();
        case LBM_STORE_BF16S:
            return fn.template operator()<lbm::StoreBF16>();
        case LBM_STORE_F64:
            return fn.template operator()<lbm::StoreF64>();
        default:
            return fn.template operator()<lbm::StoreF32>();
    }
}

int alloc(lbm3d& h, void** out, std::size_t bytes, const char* what) {
    *out = nullptr;
    if (bytes == 0) {
        return ok();
    }
    const hipError_t e = hipMalloc(out, bytes);
    if (e != hipSuccess) {
        *out = nullptr;
        return fail(LBM_ENOMEM, std::string("lbm3d gpu: ") + what + ": " + hipGetErrorString(e));
    }
    h.owned.push_back(*out);
    h.device_bytes += bytes;
    return ok();
}

template <typename V> int upload(lbm3d& h, void* dst, const std::vector<V>& src, const char* what) {
    if (src.empty()) {
        return ok();
    }
    const hipError_t e = hipMemcpyAsync(dst, src.data(), src.size() * sizeof(V), hipMemcpyHostToDevice, h.stream);
    return e == hipSuccess ? ok() : hip_fail(what, e);
}

void release(lbm3d* h) {
    if (h == nullptr) {
        return;
    }
    if (h->stream != nullptr) {
        (void)hipStreamSynchronize(h->stream);
    }
    for (void* p : h->owned) {
        (void)hipFree(p);
    }
    if (h->side != nullptr) {
        (void)hipStreamSynchronize(h->side);
        (void)hipStreamDestroy(h->side);
    }
    if (h->fork != nullptr) {
        (void)hipEventDestroy(h->fork);
    }
    if (h->join != nullptr) {
        (void)hipEventDestroy(h->join);
    }
    if (h->stream != nullptr) {
        (void)hipStreamDestroy(h->stream);
    }
    delete h;
}

// LDS each bulk workgroup reserves for the handle's cap on workgroups per compute unit (A step
// when `local`), rounded down to 512-byte allocation granules: `groups` fit in a compute unit's
// LDS, `groups + 1` do not. 0 without a cap.
std::uint32_t bulk_lds(const lbm3d& h, bool local) {
    const int groups = h.groups[local ? 0 : 1];
    if (groups <= 0 || h.lds_per_cu == 0) {
        return 0;
    }
    return h.lds_per_cu / static_cast<std::uint32_t>(groups) / 512U * 512U;
}

// One step under the handle's schedule: one fused launch; the bulk and boundary kernels one
// after the other; or the two kernels concurrently on two streams.
// The end of a step: dithered formats advance their step clock (after both halves of the
// step, on the handle's stream).
template <class T> hipError_t end_step(lbm3d& h) {
    if constexpr (T::dithered) {
        return lbm::d3q19_tick(h.clock, h.stream);
    }
    return hipSuccess;
}

template <class T> hipError_t step_once(lbm3d& h, const lbm::Lattice3<typename T::R>& p, bool local) {
    hipError_t e = hipSuccess;
    if (h.schedule == LBM_SCHEDULE_FUSED) {
        e = lbm::d3q19_step_fused<T>(p, local, h.stream);
    } else if (h.schedule == LBM_SCHEDULE_SERIAL || p.nb == 0) {
        e = lbm::d3q19_step<T>(p, local, h.block, local ? h.cpt : h.cpt_b, bulk_lds(h, local), h.stream);
    } else {
        e = hipEventRecord(h.fork, h.stream);
        if (e == hipSuccess) e = hipStreamWaitEvent(h.side, h.fork, 0);
        if (e == hipSuccess) e = lbm::d3q19_step_boundary<T>(p, local, h.side);
        if (e == hipSuccess) e = lbm::d3q19_step_bulk<T>(p, local, h.block, local ? h.cpt : h.cpt_b, bulk_lds(h, local), h.stream);
        if (e == hipSuccess) e = lbm::d3q19_step_side<T>(p, local, h.stream);
        if (e == hipSuccess) e = hipEventRecord(h.join, h.side);
        if (e == hipSuccess) e = hipStreamWaitEvent(h.stream, h.join, 0);
    }
    return e == hipSuccess ? end_step<T>(h) : e;
}

int sync(lbm3d& h, const char* what) {
    const hipError_t e = hipStreamSynchronize(h.stream);
    return e == hipSuccess ? ok() : hip_fail(what, e);
}

int initialize_mode(lbm3d* h, int mode, const double* host_planes, std::uint64_t seed, double amplitude) {
    if (h == nullptr) {
        return fail(LBM_EINVAL, "lbm3d: null handle");
    }
    const bool inflow = h->cfg.x_boundary == LBM_X_INFLOW_OUTFLOW;
    const double u0[3] = {inflow ? h->cfg.inlet_ux : 0.0, inflow ? h->cfg.inlet_uy : 0.0,
                          inflow ? h->cfg.inlet_uz : 0.0};
    double* planes = nullptr;
    if (mode == 1) {
        hipError_t e = hipMalloc(reinterpret_cast<void**>(&planes), 4U * h->cells * sizeof(double));
        if (e != hipSuccess) {
            return fail(LBM_ENOMEM, std::string("lbm3d gpu: initial state buffer: ") + hipGetErrorString(e));
        }
        e = hipMemcpyAsync(planes, host_planes, 4U * h->cells * sizeof(double), hipMemcpyHostToDevice, h->stream);
        if (e != hipSuccess) {
            (void)hipFree(planes);
            return hip_fail("initial state upload", e);
        }
    }
    const int rc = with_store(*h, [&]<class T>() -> int {
        hipError_t e = lbm::d3q19_initialize<T>(descriptor<T>(*h), mode, planes, u0, seed, amplitude, h->stream);
        if (e == hipSuccess) {
            e = hipMemsetAsync(h->clock, 0, sizeof(std::uint32_t), h->stream);
        }
        return e == hipSuccess ? ok() : hip_fail("initialize", e);
    });
    const int done = sync(*h, "initialize");
    if (planes != nullptr) {
        (void)hipFree(planes);
    }
    if (rc != LBM_OK) {
        return rc;
    }
    if (done != LBM_OK) {
        return done;
    }
    h->steps = 0;
    h->inlet[0] = h->cfg.inlet_ux;
    h->inlet[1] = h->cfg.inlet_uy;
    h->inlet[2] = h->cfg.inlet_uz;
    return ok();
}

} // namespace

// A lattice held by a batch is stepped by the batch alone (lbm_internal.h, lbm3d::in_batch).
static int refuse_in_batch(const char* what) {
    return fail(LBM_EINVAL, std::string("lbm3d: ") + what + ": the lattice belongs to a batch (use the lbm_batch_* calls)");
}

extern "C" {

int lbm3d_create(const lbm3d_config* config, const uint8_t* solid, const double* distance, int32_t storage,
                 lbm3d** out) {
    if (out == nullptr || config == nullptr) {
        return fail(LBM_EINVAL, "lbm3d: null argument");
    }
    *out = nullptr;
    if (storage < LBM_STORE_F32 || storage > LBM_STORE_F64) {
        return fail(LBM_EINVAL, "lbm3d gpu: unknown population storage");
    }
    const lbm3d_config& c = *config;
    // validate() (lbm3d_common.cpp:159-161)
    if (std::string p = lbm::host::problem3(c, solid, distance); !p.empty()) {
        return fail(LBM_EINVAL, p);
    }
    const std::size_t cells = c.nx * c.ny * c.nz;
    const lbm::internal::PopulationLayout layout = lbm::internal::population_layout(c.nx, c.ny, c.nz, storage);
    // Every population index the kernels form must fit in 32 bits (lbm3d_gpu.cpp:492-497), and
    // the byte offsets of the bulk kernel's waves (a few blocks either side of a plane's
    // neighbours, at most 2 (nx ny + nx + 1) + 4 2^shift cells of 19 populations) too.
    const std::uint64_t wave_span = (2U * static_cast<std::uint64_t>(layout.guard) + (std::uint64_t{4} << layout.shift)) * 19U *
                                    static_cast<std::uint64_t>(store_bytes(storage));
    if (layout.elements > std::numeric_limits<std::uint32_t>::max() || wave_span > std::numeric_limits<std::uint32_t>::max()) {
        return fail(LBM_EINVAL, "lbm3d gpu: lattice too large for 32-bit indices");
    }
    if (storage != LBM_STORE_F64) {
        if (std::string p = lbm::host::binary32_problem3(c); !p.empty()) {
            return fail(LBM_EINVAL, p);
        }
    }
    auto* h = new (std::nothrow) lbm3d();
    if (h == nullptr) {
        return fail(LBM_ENOMEM, "lbm3d: host allocation failed");
    }
    h->cfg = c;
    h->storage = storage;
    h->cpt = lbm::internal::default_cells_per_thread(storage);
    h->cpt_b = h->cpt;
    h->cells = cells;
    h->layout = layout;
    h->solid.assign(solid, solid + cells);
    h->tables = lbm::host::tables3(c, solid, distance);
    const bool interpolated = c.solid_wall == LBM_SOLID_INTERPOLATED;
    // check_memory (lbm3d_gpu.cpp:131-149): refuse before allocating.
    const std::size_t need = lattice_bytes(cells, layout, h->tables, storage, interpolated);
    std::size_t available = 0;
    if (int rc = lbm::host::device_memory(&available, nullptr); rc != LBM_OK) {
        delete h;
        return rc;
    }
    if (need > available) {
        delete h;
        return fail(LBM_ENOMEM, "lbm3d gpu: the lattice needs " + lbm::host::megabytes(need) +
                                    " of device memory, " + lbm::host::megabytes(available) +
                                    " are available; use fewer cells");
    }
    hipError_t e = hipGetDevice(&h->device);
    int lds_cu = 0;
    int lds_block = 0;
    if (e == hipSuccess) {
        e = hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, h->device);
    }
    if (e == hipSuccess) {
        e = hipDeviceGetAttribute(&lds_block, hipDeviceAttributeMaxSharedMemoryPerBlock, h->device);
    }
    h->lds_per_cu = static_cast<std::uint32_t>(lds_cu > 0 ? lds_cu : 0);
    h->lds_per_block = static_cast<std::uint32_t>(lds_block > 0 ? lds_block : 0);
    // One stream; the second stream of the two-stream schedule is created when that schedule is
    // selected. An idle second stream is not free: a first lattice of a process whose handle
    // also held one ran its B steps 4-5% slower (512x256x256 f32 bulk kernel, B step 1027-1030
    // against 983 us without it), an effect of how the process's streams map to the device's
    // hardware queues.
    if (e == hipSuccess) {
        e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    }
    if (e == hipSuccess) {
        e = hipEventCreateWithFlags(&h->fork, hipEventDisableTiming);
    }
    if (e == hipSuccess) {
        e = hipEventCreateWithFlags(&h->join, hipEventDisableTiming);
    }
    if (e != hipSuccess) {
        release(h);
        return hip_fail("stream", e);
    }
    const lbm::host::Tables& t = h->tables;
    const std::size_t sb = store_bytes(storage);
    const std::size_t rb = real_bytes(storage);
    const std::size_t nb = t.bcell.size();
    const std::size_t nl = t.links.size();
    const std::size_t red = std::max<std::size_t>(lbm::reduction_blocks(cells) + lbm::reduction_blocks(nb),
                                                  6U * lbm::reduction_blocks(nl));
    int rc = LBM_OK;
    void* tmp = nullptr;
    const std::size_t cls_len = lbm::internal::class_length(cells);
    if ((rc = alloc(*h, &h->f, layout.elements * sb, "population buffer")) ||
        (rc = alloc(*h, &tmp, cls_len, "node class buffer")) || ((h->cls = static_cast<std::uint8_t*>(tmp)), false) ||
        (rc = alloc(*h, &tmp, std::max<std::size_t>(nb, 1) * 4U, "boundary list")) ||
        ((h->bcell = static_cast<std::uint32_t*>(tmp)), false) ||
        (rc = alloc(*h, &tmp, t.tile_slot.size() * 4U, "boundary slot table")) ||
        ((h->tile_slot = static_cast<std::uint32_t*>(tmp)), false) ||
        (rc = alloc(*h, &tmp, std::max<std::size_t>(t.scell.size(), 1) * 4U, "side list")) ||
        ((h->scell = static_cast<std::uint32_t*>(tmp)), false) ||
        (rc = alloc(*h, &h->offsets, 38U * rb, "storage offsets")) ||
        (rc = alloc(*h, &tmp, 19U * static_cast<std::size_t>(t.bstride) * 4U, "source codes")) ||
        ((h->bcode = static_cast<std::uint32_t*>(tmp)), false) ||
        (rc = alloc(*h, &tmp, static_cast<std::size_t>(t.bstride) * 4U, "lattice write masks")) ||
        ((h->bwrite = static_cast<std::uint32_t*>(tmp)), false) ||
        (interpolated && (rc = alloc(*h, &h->bq, 19U * static_cast<std::size_t>(t.bstride) * rb, "wall buffer"))) ||
        (rc = alloc(*h, &h->stash, 2U * 19U * static_cast<std::size_t>(t.bstride) * rb, "boundary stash")) ||
        (rc = alloc(*h, &h->hist, 5U * static_cast<std::size_t>(t.bstride) * rb, "outlet buffer")) ||
        (rc = alloc(*h, &tmp, std::max<std::size_t>(nl, 1) * 4U, "link buffer")) ||
        ((h->lcell = static_cast<std::uint32_t*>(tmp)), false) ||
        (rc = alloc(*h, &tmp, std::max<std::size_t>(nl, 1), "link buffer")) ||
        ((h->ldir = static_cast<std::uint8_t*>(tmp)), false) ||
        (rc = alloc(*h, &tmp, std::max<std::size_t>(nl, 1) * 8U, "link buffer")) ||
        ((h->lq = static_cast<double*>(tmp)), false) ||
        (rc = alloc(*h, &tmp, std::max<std::size_t>(nl, 1) * 4U, "link buffer")) ||
        ((h->lslot = static_cast<std::uint32_t*>(tmp)), false) ||
        (rc = alloc(*h, &tmp, (red + 8U) * 8U, "reduction buffer")) || ((h->partial = static_cast<double*>(tmp)), false)) {
        release(h);
        return rc;
    }
    h->result = h->partial + red;
    h->clock = reinterpret_cast<std::uint32_t*>(h->result + 7); // results use at most 6 doubles
    // Upload the tables.
    std::vector<std::uint32_t> lcell(nl);
    std::vector<std::uint8_t> ldir(nl);
    std::vector<double> lq(nl);
    for (std::size_t k = 0; k < nl; ++k) {
        lcell[k] = t.links[k].cell;
        ldir[k] = static_cast<std::uint8_t>(t.links[k].direction);
        lq[k] = t.links[k].q;
    }
    // Storage offsets of the 16-bit formats (zeros for the others), as bytes of the arithmetic type;
    // kept until the stream is drained below.
    std::vector<unsigned char> offsets(38U * rb);
    (void)with_store(*h, [&]<class T>() -> int {
        const std::vector<typename T::R> v = lbm::internal::storage_offsets<T>(h->cfg);
        std::memcpy(offsets.data(), v.data(), offsets.size());
        return LBM_OK;
    });
    // The class array is padded with solid entries, so the bulk kernel reads the classes of a
    // whole run of cells at once.
    if (cls_len > cells) {
        const hipError_t pe = hipMemsetAsync(h->cls + cells, 0, cls_len - cells, h->stream);
        if (pe != hipSuccess) {
            release(h);
            return hip_fail("node class padding", pe);
        }
    }
    if ((rc = upload(*h, h->cls, t.cls, "node class upload")) || (rc = upload(*h, h->bcell, t.bcell, "boundary upload")) ||
        (rc = upload(*h, h->tile_slot, t.tile_slot, "boundary slot upload")) ||
        (rc = upload(*h, h->scell, t.scell, "side list upload")) ||
        (rc = upload(*h, h->offsets, offsets, "storage offsets upload")) ||
        (rc = upload(*h, h->bcode, t.bcode, "code upload")) || (rc = upload(*h, h->bwrite, t.bwrite, "mask upload")) ||
        (rc = upload(*h, h->lcell, lcell, "link upload")) ||
        (rc = upload(*h, h->ldir, ldir, "link upload")) || (rc = upload(*h, h->lq, lq, "link upload")) ||
        (rc = upload(*h, h->lslot, t.lslot, "link upload"))) {
        release(h);
        return rc;
    }
    if (interpolated) {
        if (storage == LBM_STORE_F64) {
            rc = upload(*h, h->bq, t.bq, "wall upload");
        } else {
            std::vector<float> q32(t.bq.begin(), t.bq.end());
            rc = upload(*h, h->bq, q32, "wall upload");
        }
        if (rc != LBM_OK) {
            release(h);
            return rc;
        }
    }
    // Zero the history and the stash so that no stale bits are ever read.
    if ((e = hipMemsetAsync(h->hist, 0, 5U * static_cast<std::size_t>(t.bstride) * rb, h->stream)) != hipSuccess ||
        (e = hipMemsetAsync(h->stash, 0, 2U * 19U * static_cast<std::size_t>(t.bstride) * rb, h->stream)) !=
            hipSuccess) {
        release(h);
        return hip_fail("clear", e);
    }
    if ((rc = initialize_mode(h, 0, nullptr, 0, 0.0)) != LBM_OK) {
        release(h);
        return rc;
    }
    *out = h;
    return ok();
}

int lbm3d_destroy(lbm3d* h) {
    if (h != nullptr && h->in_batch) {
        return refuse_in_batch("destroy");
    }
    release(h);
    return ok();
}

int lbm3d_initialize(lbm3d* h, const double* moments) {
    if (h != nullptr && h->in_batch) {
        return refuse_in_batch("initialize");
    }
    return initialize_mode(h, moments != nullptr ? 1 : 0, moments, 0, 0.0);
}

int lbm3d_initialize_perturbed(lbm3d* h, uint64_t seed, double amplitude) {
    if (h != nullptr && h->in_batch) {
        return refuse_in_batch("initialize");
    }
    if (!std::isfinite(amplitude)) {
        return fail(LBM_EINVAL, "lbm3d: the perturbation amplitude must be finite");
    }
    return initialize_mode(h, 2, nullptr, seed, amplitude);
}

int lbm3d_step(lbm3d* h, uint64_t count) {
    if (h == nullptr) {
        return fail(LBM_EINVAL, "lbm3d: null handle");
    }
    if (h->in_batch) {
        return refuse_in_batch("step");
    }
    return with_store(*h, [&]<class T>() -> int {
        const auto p = descriptor<T>(*h);
        const auto launch = [&](bool local) { return step_once<T>(*h, p, local); };
        if (h->schedule == LBM_SCHEDULE_STREAMS) {
            for (std::uint64_t n = 0; n < count; ++n) {
                const hipError_t e = launch((h->steps & 1U) != 0);
                if (e != hipSuccess) {
                    return hip_fail("step", e);
                }
                ++h->steps;
            }
            return ok();
        }
        std::vector<char> key;
        lbm::key_append(key, p);
        lbm::key_append(key, h->schedule);
        lbm::key_append(key, h->block);
        lbm::key_append(key, h->cpt);
        lbm::key_append(key, h->cpt_b);
        lbm::key_append(key, h->groups);
        const hipError_t e = h->graph.run(h->stream, count, h->steps, key, launch);
        return e == hipSuccess ? ok() : hip_fail("step", e);
    });
}

int lbm3d_set_graphs(lbm3d* h, int32_t enabled) {
    if (h == nullptr) {
        return fail(LBM_EINVAL, "lbm3d: null handle");
    }
    h->graph.enable(enabled != 0);
    return ok();
}

uint64_t lbm3d_steps(const lbm3d* h) { return h != nullptr ? h->steps : 0; }

int lbm3d_set_inlet(lbm3d* h, const double velocity[3]) {
    if (h == nullptr || velocity == nullptr) {
        return fail(LBM_EINVAL, "lbm3d: null argument");
    }
    if (h->in_batch) {
        return refuse_in_batch("set_inlet");
    }
    if (std::string p = lbm::host::inlet_problem3(h->cfg, velocity); !p.empty()) {
        return fail(LBM_EINVAL, p);
    }
    for (int a = 0; a < 3; ++a) {
        h->inlet[a] = velocity[a];
    }
    return ok();
}

int lbm3d_get_inlet(const lbm3d* h, double velocity[3]) {
    if (h == nullptr || velocity == nullptr) {
        return fail(LBM_EINVAL, "lbm3d: null argument");
    }
    for (int a = 0; a < 3; ++a) {
        velocity[a] = h->inlet[a];
    }
    return ok();
}

int lbm3d_force(lbm3d* h, double force[3], double moment[3]) {
    if (h == nullptr || force == nullptr || moment == nullptr) {
        return fail(LBM_EINVAL, "lbm3d: null argument");
    }
    const std::size_t nl = h->tables.links.size();
    if (nl == 0) {
        for (int a = 0; a < 3; ++a) {
            force[a] = 0.0;
            moment[a] = 0.0;
        }
        return ok();
    }
    const int phase = static_cast<int>(h->steps & 1U);
    int rc = with_store(*h, [&]<class T>() -> int {
        const hipError_t e = lbm::d3q19_force_sum<T>(descriptor<T>(*h), phase, h->lcell, h->ldir, h->lq, h->lslot,
                                                     static_cast<std::uint32_t>(nl), h->partial, h->result, h->stream);
        return e == hipSuccess ? ok() : hip_fail("force", e);
    });
    if (rc != LBM_OK) {
        return rc;
    }
    double v[6];
    hipError_t e = hipMemcpyAsync(v, h->result, sizeof(v), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) {
        e = hipStreamSynchronize(h->stream);
    }
    if (e != hipSuccess) {
        return hip_fail("force readback", e);
    }
    for (int a = 0; a < 3; ++a) {
        force[a] = v[a];
        moment[a] = v[3 + a];
    }
    return ok();
}

int lbm3d_mass(lbm3d* h, double* mass) {
    if (h == nullptr || mass == nullptr) {
        return fail(LBM_EINVAL, "lbm3d: null argument");
    }
    const int phase = static_cast<int>(h->steps & 1U);
    int rc = with_store(*h, [&]<class T>() -> int {
        const hipError_t e = lbm::d3q19_mass<T>(descriptor<T>(*h), phase, h->partial, h->result, h->stream);
        return e == hipSuccess ? ok() : hip_fail("mass", e);
    });
    if (rc != LBM_OK) {
        return rc;
    }
    hipError_t e = hipMemcpyAsync(mass, h->result, sizeof(double), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) {
        e = hipStreamSynchronize(h->stream);
    }
    return e == hipSuccess ? ok() : hip_fail("mass readback", e);
}

int lbm3d_macroscopic(lbm3d* h, double* rho, double* ux, double* uy, double* uz) {
    if (h == nullptr) {
        return fail(LBM_EINVAL, "lbm3d: null handle");
    }
    double* planes = nullptr;
    const std::size_t bytes = 4U * h->cells * sizeof(double);
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&planes), bytes);
    if (e != hipSuccess) {
        return fail(LBM_ENOMEM, std::string("lbm3d gpu: moment buffer: ") + hipGetErrorString(e));
    }
    const int phase = static_cast<int>(h->steps & 1U);
    const double force[3] = {h->cfg.body_force_x, h->cfg.body_force_y, h->cfg.body_force_z};
    int rc = with_store(*h, [&]<class T>() -> int {
        const hipError_t err = lbm::d3q19_moments<T>(descriptor<T>(*h), phase, force, planes, h->stream);
        return err == hipSuccess ? ok() : hip_fail("moments", err);
    });
    double* outs[4] = {rho, ux, uy, uz};
    for (int k = 0; k < 4 && rc == LBM_OK; ++k) {
        if (outs[k] != nullptr) {
            e = hipMemcpyAsync(outs[k], planes + k * h->cells, h->cells * sizeof(double), hipMemcpyDeviceToHost,
                               h->stream);
            if (e != hipSuccess) {
                rc = hip_fail("moment readback", e);
            }
        }
    }
    const int s = sync(*h, "moment readback");
    (void)hipFree(planes);
    return rc != LBM_OK ? rc : s;
}

int lbm3d_sample(lbm3d* h, const uint32_t* cells, size_t n, double* rho, double* ux, double* uy, double* uz) {
    if (h == nullptr || (n > 0 && cells == nullptr)) {
        return fail(LBM_EINVAL, "lbm3d: null argument");
    }
    if (n == 0) {
        return ok();
    }
    if (n > std::numeric_limits<std::int32_t>::max()) {
        return fail(LBM_EINVAL, "lbm3d: too many sample nodes");
    }
    // Boundary slot of each node (the boundary list is in cell order), or -1.
    const std::vector<std::uint32_t>& bcell = h->tables.bcell;
    std::vector<std::int32_t> slots(n);
    for (std::size_t k = 0; k < n; ++k) {
        if (cells[k] >= h->cells) {
            return fail(LBM_EINVAL, "lbm3d: sample node outside the lattice");
        }
        const auto it = std::lower_bound(bcell.begin(), bcell.end(), cells[k]);
        slots[k] = it != bcell.end() && *it == cells[k] ? static_cast<std::int32_t>(it - bcell.begin()) : -1;
    }
    char* buffer = nullptr;
    const std::size_t bytes = n * (4U + 4U + 4U * sizeof(double));
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&buffer), bytes);
    if (e != hipSuccess) {
        return fail(LBM_ENOMEM, std::string("lbm3d gpu: sample buffer: ") + hipGetErrorString(e));
    }
    auto* dcells = reinterpret_cast<std::uint32_t*>(buffer);
    auto* dslots = reinterpret_cast<std::int32_t*>(buffer + 4U * n);
    auto* dout = reinterpret_cast<double*>(buffer + 8U * n);
    e = hipMemcpyAsync(dcells, cells, 4U * n, hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) {
        e = hipMemcpyAsync(dslots, slots.data(), 4U * n, hipMemcpyHostToDevice, h->stream);
    }
    int rc = e == hipSuccess ? LBM_OK : hip_fail("sample upload", e);
    if (rc == LBM_OK) {
        const int phase = static_cast<int>(h->steps & 1U);
        const double force[3] = {h->cfg.body_force_x, h->cfg.body_force_y, h->cfg.body_force_z};
        rc = with_store(*h, [&]<class T>() -> int {
            const hipError_t err = lbm::d3q19_sample<T>(descriptor<T>(*h), phase, force, dcells, dslots,
                                                        static_cast<std::uint32_t>(n), dout, h->stream);
            return err == hipSuccess ? ok() : hip_fail("sample", err);
        });
    }
    double* outs[4] = {rho, ux, uy, uz};
    for (int k = 0; k < 4 && rc == LBM_OK; ++k) {
        if (outs[k] != nullptr) {
            e = hipMemcpyAsync(outs[k], dout + k * n, n * sizeof(double), hipMemcpyDeviceToHost, h->stream);
            if (e != hipSuccess) {
                rc = hip_fail("sample readback", e);
            }
        }
    }
    const int sy = sync(*h, "sample readback");
    (void)hipFree(buffer);
    return rc != LBM_OK ? rc : sy;
}

} // extern "C"

// ---- lbm3d_sample_fixed: lbm3d_sample without allocations -----------------------------------
// The per-call buffer of lbm3d_sample (node list, boundary slots, four output planes) lives in
// one file-scope device array instead. Each device holds its own copy; its address on the
// current device is looked up once (hipGetSymbolAddress) and kept per device. The host names
// `scratch` only as a symbol handle: kernels, copies and fills take the looked-up address.
namespace {

constexpr std::size_t kSampleMax = 1U << 16;  // nodes per call
constexpr std::size_t kSampleBytes = kSampleMax * (4U + 4U + 4U * sizeof(double));
static __device__ __align__(256) unsigned char scratch[kSampleBytes];

constexpr int kMaxDevices = 64;
std::once_flag scratch_once[kMaxDevices];
void* scratch_ptr[kMaxDevices] = {};
hipError_t scratch_status[kMaxDevices] = {};
std::mutex scratch_lock[kMaxDevices];  // one call at a time per device shares its array

// Boundary slot of each listed node (bcell is sorted: the boundary list is in cell order), or
// -1: the host loop of lbm3d_sample, as a binary search on the device.
__global__ __launch_bounds__(256) void sample_slots(const std::uint32_t* __restrict__ bcell, std::uint32_t nb,
                                                    const std::uint32_t* __restrict__ cells, std::uint32_t n,
                                                    std::int32_t* __restrict__ slots) {
    const std::uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) {
        return;
    }
    const std::uint32_t cell = cells[k];
    std::uint32_t lo = 0;
    std::uint32_t hi = nb;
    while (lo < hi) {
        const std::uint32_t mid = lo + ((hi - lo) >> 1);
        if (bcell[mid] < cell) {
            lo = mid + 1;
        } else {
            hi = mid;
        }
    }
    slots[k] = lo < nb && bcell[lo] == cell ? static_cast<std::int32_t>(lo) : -1;
}

} // namespace

extern "C" {

int lbm3d_sample_fixed(lbm3d* h, const uint32_t* cells, size_t n, double* rho, double* ux, double* uy, double* uz) {
    if (h == nullptr || (n > 0 && cells == nullptr)) {
        return fail(LBM_EINVAL, "lbm3d: null argument");
    }
    if (n == 0) {
        return ok();
    }
    if (n > kSampleMax) {
        return fail(LBM_EINVAL, "lbm3d: lbm3d_sample_fixed takes at most " + std::to_string(kSampleMax) +
                                    " nodes per call (" + std::to_string(n) + " given); use lbm3d_sample");
    }
    for (std::size_t k = 0; k < n; ++k) {
        if (cells[k] >= h->cells) {
            return fail(LBM_EINVAL, "lbm3d: sample node outside the lattice");
        }
    }
    if (h->device < 0 || h->device >= kMaxDevices) {
        return fail(LBM_EINVAL, "lbm3d: lbm3d_sample_fixed supports device ordinals below " +
                                    std::to_string(kMaxDevices));
    }
    const int dev = h->device;
    std::call_once(scratch_once[dev], [dev] {
        int current = 0;
        hipError_t e = hipGetDevice(&current);
        if (e == hipSuccess && current != dev) {
            e = hipSetDevice(dev);
        }
        if (e == hipSuccess) {
            e = hipGetSymbolAddress(&scratch_ptr[dev], HIP_SYMBOL(scratch));
        }
        if (current != dev) {
            (void)hipSetDevice(current);
        }
        scratch_status[dev] = e;
    });
    if (scratch_status[dev] != hipSuccess || scratch_ptr[dev] == nullptr) {
        return hip_fail("sample scratch address", scratch_status[dev] != hipSuccess ? scratch_status[dev]
                                                                                    : hipErrorInvalidSymbol);
    }
    const std::lock_guard<std::mutex> guard(scratch_lock[dev]);
    char* base = static_cast<char*>(scratch_ptr[dev]);
    auto* dcells = reinterpret_cast<std::uint32_t*>(base);
    auto* dslots = reinterpret_cast<std::int32_t*>(base + 4U * kSampleMax);
    auto* dout = reinterpret_cast<double*>(base + 8U * kSampleMax);
    const auto count = static_cast<std::uint32_t>(n);
    hipError_t e = hipMemcpyAsync(dcells, cells, 4U * n, hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(sample_slots, dim3(static_cast<std::uint32_t>((n + 255U) / 256U)), dim3(256), 0, h->stream,
                           h->bcell, static_cast<std::uint32_t>(h->tables.bcell.size()), dcells, count, dslots);
        e = hipGetLastError();
    }
    int rc = e == hipSuccess ? LBM_OK : hip_fail("sample upload", e);
    if (rc == LBM_OK) {
        const int phase = static_cast<int>(h->steps & 1U);
        const double force[3] = {h->cfg.body_force_x, h->cfg.body_force_y, h->cfg.body_force_z};
        rc = with_store(*h, [&]<class T>() -> int {
            const hipError_t err = lbm::d3q19_sample<T>(descriptor<T>(*h), phase, force, dcells, dslots, count, dout,
                                                        h->stream);
            return err == hipSuccess ? ok() : hip_fail("sample", err);
        });
    }
    double* outs[4] = {rho, ux, uy, uz};
    for (int k = 0; k < 4 && rc == LBM_OK; ++k) {
        if (outs[k] != nullptr) {
            e = hipMemcpyAsync(outs[k], dout + k * n, n * sizeof(double), hipMemcpyDeviceToHost, h->stream);
            if (e != hipSuccess) {
                rc = hip_fail("sample readback", e);
            }
        }
    }
    const int sy = sync(*h, "sample readback");
    return rc != LBM_OK ? rc : sy;
}

int lbm3d_max_speed(lbm3d* h, double* speed) {
    if (h == nullptr || speed == nullptr) {
        return fail(LBM_EINVAL, "lbm3d: null argument");
    }
    const int phase = static_cast<int>(h->steps & 1U);
    const double force[3] = {h->cfg.body_force_x, h->cfg.body_force_y, h->cfg.body_force_z};
    int rc = with_store(*h, [&]<class T>() -> int {
        const hipError_t e = lbm::d3q19_max_speed<T>(descriptor<T>(*h), phase, force, h->partial, h->result, h->stream);
        return e == hipSuccess ? ok() : hip_fail("max speed", e);
    });
    if (rc != LBM_OK) {
        return rc;
    }
    hipError_t e = hipMemcpyAsync(speed, h->result, sizeof(double), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) {
        e = hipStreamSynchronize(h->stream);
    }
    return e == hipSuccess ? ok() : hip_fail("max speed readback", e);
}

int lbm3d_populations(lbm3d* h, double* f) {
    if (h == nullptr || f == nullptr) {
        return fail(LBM_EINVAL, "lbm3d: null argument");
    }
    double* buffer = nullptr;
    const std::size_t bytes = 19U * h->cells * sizeof(double);
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&buffer), bytes);
    if (e != hipSuccess) {
        return fail(LBM_ENOMEM, std::string("lbm3d gpu: population readback buffer: ") + hipGetErrorString(e));
    }
    const int phase = static_cast<int>(h->steps & 1U);
    int rc = with_store(*h, [&]<class T>() -> int {
        const hipError_t err = lbm::d3q19_populations<T>(descriptor<T>(*h), phase, buffer, h->stream);
        return err == hipSuccess ? ok() : hip_fail("populations", err);
    });
    if (rc == LBM_OK) {
        e = hipMemcpyAsync(f, buffer, bytes, hipMemcpyDeviceToHost, h->stream);
        if (e != hipSuccess) {
            rc = hip_fail("population readback", e);
        }
    }
    const int s = sync(*h, "population readback");
    (void)hipFree(buffer);
    return rc != LBM_OK ? rc : s;
}

int lbm3d_links(const lbm3d* h, uint32_t* cell, uint32_t* direction, double* q, size_t capacity, size_t* count) {
    if (h == nullptr) {
        return fail(LBM_EINVAL, "lbm3d: null handle");
    }
    const auto& links = h->tables.links;
    for (std::size_t k = 0; k < links.size() && k < capacity; ++k) {
        if (cell != nullptr) cell[k] = links[k].cell;
        if (direction != nullptr) direction[k] = links[k].direction;
        if (q != nullptr) q[k] = links[k].q;
    }
    if (count != nullptr) {
        *count = links.size();
    }
    return ok();
}

int lbm3d_device_bytes(const lbm3d* h, size_t* bytes) {
    if (h == nullptr || bytes == nullptr) {
        return fail(LBM_EINVAL, "lbm3d: null argument");
    }
    *bytes = h->device_bytes;
    return ok();
}

int lbm3d_synchronize(lbm3d* h) {
    if (h == nullptr) {
        return fail(LBM_EINVAL, "lbm3d: null handle");
    }
    return sync(*h, "synchronize");
}

int lbm3d_set_launch(lbm3d* h, int32_t block, int32_t cells_per_thread) {
    if (h == nullptr) {
        return fail(LBM_EINVAL, "lbm3d: null handle");
    }
    if (block != 0 && (block < 64 || block > 256 || block % 64 != 0)) {
        return fail(LBM_EINVAL, "lbm3d: block size must be 64, 128, 192 or 256");
    }
    if (cells_per_thread != 0 && cells_per_thread != 1 && cells_per_thread != 2 && cells_per_thread != 4) {
        return fail(LBM_EINVAL, "lbm3d: cells per thread must be 1, 2 or 4");
    }
    if (block != 0) h->block = block;
    if (cells_per_thread != 0) {
        h->cpt = cells_per_thread;
        h->cpt_b = cells_per_thread;
    }
    return ok();
}

int lbm3d_set_cells_per_thread(lbm3d* h, int32_t a_step, int32_t b_step) {
    if (h == nullptr) {
        return fail(LBM_EINVAL, "lbm3d: null handle");
    }
    for (const int32_t v : {a_step, b_step}) {
        if (v != 0 && v != 1 && v != 2 && v != 4) {
            return fail(LBM_EINVAL, "lbm3d: cells per thread must be 1, 2 or 4");
        }
    }
    if (a_step != 0) h->cpt = a_step;
    if (b_step != 0) h->cpt_b = b_step;
    return ok();
}

int lbm3d_get_cells_per_thread(const lbm3d* h, int32_t* a_step, int32_t* b_step) {
    if (h == nullptr) {
        return fail(LBM_EINVAL, "lbm3d: null handle");
    }
    if (a_step != nullptr) *a_step = h->cpt;
    if (b_step != nullptr) *b_step = h->cpt_b;
    return ok();
}

int lbm3d_set_schedule(lbm3d* h, int32_t schedule) {
    if (h == nullptr) {
        return fail(LBM_EINVAL, "lbm3d: null handle");
    }
    if (schedule != LBM_SCHEDULE_FUSED && schedule != LBM_SCHEDULE_SERIAL && schedule != LBM_SCHEDULE_STREAMS) {
        return fail(LBM_EINVAL, "lbm3d: unknown step schedule");
    }
    if (schedule == LBM_SCHEDULE_STREAMS && h->side == nullptr) {
        const hipError_t e = hipStreamCreateWithFlags(&h->side, hipStreamNonBlocking);
        if (e != hipSuccess) {
            h->side = nullptr;
            return hip_fail("stream", e);
        }
    }
    h->schedule = schedule;
    return ok();
}

int lbm3d_step_bytes(const lbm3d* h, double* bytes) {
    if (h == nullptr || bytes == nullptr) {
        return fail(LBM_EINVAL, "lbm3d: null argument");
    }
    *bytes = 2.0 * 19.0 * static_cast<double>(store_bytes(h->storage)) * static_cast<double>(h->cells);
    return ok();
}

int lbm3d_counts(const lbm3d* h, size_t* bulk, size_t* boundary, size_t* solid, size_t* links) {
    if (h == nullptr) {
        return fail(LBM_EINVAL, "lbm3d: null handle");
    }
    std::size_t n[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (const std::uint8_t c : h->tables.cls) {
        ++n[c & 7U];
    }
    if (bulk != nullptr) *bulk = n[1];
    if (boundary != nullptr) *boundary = n[2] + n[3] + n[4];
    if (solid != nullptr) *solid = n[0];
    if (links != nullptr) *links = h->tables.links.size();
    return ok();
}

int lbm3d_stream_counts(const lbm3d* h, size_t* bulk, size_t* side, size_t* listed) {
    if (h == nullptr) {
        return fail(LBM_EINVAL, "lbm3d: null handle");
    }
    std::size_t n[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (const std::uint8_t c : h->tables.cls) {
        ++n[c & 7U];
    }
    // inlet nodes stream like bulk nodes
    if (bulk != nullptr) *bulk = n[1] + n[4];
    if (side != nullptr) *side = n[3];
    if (listed != nullptr) *listed = n[2];
    return ok();
}

int lbm3d_time_steps(lbm3d* h, uint64_t count, int32_t per_kernel, double* step_ms, double* bulk_ms,
                     double* boundary_ms) {
    if (h == nullptr) {
        return fail(LBM_EINVAL, "lbm3d: null handle");
    }
    if (h->in_batch) {
        return refuse_in_batch("time_steps");
    }
    const std::size_t marks = per_kernel ? 3U * count + 1U : 2U;
    std::vector<hipEvent_t> ev(marks, nullptr);
    int rc = LBM_OK;
    for (auto& e : ev) {
        if (hipEventCreate(&e) != hipSuccess) {
            rc = fail(LBM_EDEVICE, "lbm3d gpu: event creation failed");
            break;
        }
    }
    if (rc == LBM_OK) {
        hipError_t e = hipEventRecord(ev[0], h->stream);
        for (std::uint64_t n = 0; n < count && e == hipSuccess && rc == LBM_OK; ++n) {
            const bool local = (h->steps & 1U) != 0;
            rc = with_store(*h, [&]<class T>() -> int {
                const auto p = descriptor<T>(*h);
                if (!per_kernel) {
                    const hipError_t err = step_once<T>(*h, p, local);
                    return err == hipSuccess ? ok() : hip_fail("timed step", err);
                }
                hipError_t err = lbm::d3q19_step_bulk<T>(p, local, h->block, local ? h->cpt : h->cpt_b,
                                                         bulk_lds(*h, local), h->stream);
                if (err == hipSuccess && per_kernel) {
                    err = hipEventRecord(ev[3U * n + 1U], h->stream);
                }
                if (err == hipSuccess && per_kernel) {
                    err = hipEventRecord(ev[3U * n + 2U], h->stream);
                }
                if (err == hipSuccess) {
                    err = lbm::d3q19_step_side<T>(p, local, h->stream);
                }
                if (err == hipSuccess) {
                    err = lbm::d3q19_step_boundary<T>(p, local, h->stream);
                }
                if (err == hipSuccess && per_kernel) {
                    err = hipEventRecord(ev[3U * n + 3U], h->stream);
                }
                if (err == hipSuccess) {
                    err = end_step<T>(*h);
                }
                return err == hipSuccess ? ok() : hip_fail("timed step", err);
            });
            ++h->steps;
        }
        if (e == hipSuccess && rc == LBM_OK && !per_kernel) {
            e = hipEventRecord(ev[1], h->stream);
        }
        if (e == hipSuccess) {
            e = hipEventSynchronize(ev[marks - 1]);
        }
        if (e != hipSuccess && rc == LBM_OK) {
            rc = hip_fail("timing", e);
        }
        if (rc == LBM_OK) {
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, ev[0], ev[marks - 1]);
            if (step_ms != nullptr) *step_ms = ms;
            double bulk = 0.0;
            double bnd = 0.0;
            if (per_kernel) {
                for (std::uint64_t n = 0; n < count; ++n) {
                    float a = 0.0f;
                    float b = 0.0f;
                    (void)hipEventElapsedTime(&a, ev[3U * n], ev[3U * n + 1U]);
                    (void)hipEventElapsedTime(&b, ev[3U * n + 2U], ev[3U * n + 3U]);
                    bulk += a;
                    bnd += b;
                }
            }
            if (bulk_ms != nullptr) *bulk_ms = bulk;
            if (boundary_ms != nullptr) *boundary_ms = bnd;
        }
    }
    for (auto& e : ev) {
        if (e != nullptr) {
            (void)hipEventDestroy(e);
        }
    }
    return rc;
}

int lbm3d_time_kernels(lbm3d* h, uint64_t count, double ms[6]) {
    if (h == nullptr || ms == nullptr) {
        return fail(LBM_EINVAL, "lbm3d: null argument");
    }
    if (h->in_batch) {
        return refuse_in_batch("time_kernels");
    }
    // Marks around each kernel: [bulk start, bulk end = side start, side end = boundary start,
    // boundary end] per step.
    std::vector<hipEvent_t> ev(4U * count, nullptr);
    int rc = LBM_OK;
    for (auto& e : ev) {
        if (hipEventCreate(&e) != hipSuccess) {
            rc = fail(LBM_EDEVICE, "lbm3d gpu: event creation failed");
            break;
        }
    }
    for (std::uint64_t n = 0; n < count && rc == LBM_OK; ++n) {
        const bool local = (h->steps & 1U) != 0;
        rc = with_store(*h, [&]<class T>() -> int {
            const auto p = descriptor<T>(*h);
            hipError_t err = hipEventRecord(ev[4U * n], h->stream);
            if (err == hipSuccess) {
                err = lbm::d3q19_step_bulk<T>(p, local, h->block, local ? h->cpt : h->cpt_b, bulk_lds(*h, local),
                                              h->stream);
            }
            if (err == hipSuccess) err = hipEventRecord(ev[4U * n + 1U], h->stream);
            if (err == hipSuccess) err = lbm::d3q19_step_side<T>(p, local, h->stream);
            if (err == hipSuccess) err = hipEventRecord(ev[4U * n + 2U], h->stream);
            if (err == hipSuccess) err = lbm::d3q19_step_boundary<T>(p, local, h->stream);
            if (err == hipSuccess) err = hipEventRecord(ev[4U * n + 3U], h->stream);
            if (err == hipSuccess) err = end_step<T>(*h);
            return err == hipSuccess ? ok() : hip_fail("timed kernels", err);
        });
        if (rc == LBM_OK) {
            ++h->steps;
        }
    }
    if (rc == LBM_OK && count > 0) {
        const hipError_t e = hipEventSynchronize(ev[4U * count - 1U]);
        if (e != hipSuccess) {
            rc = hip_fail("timed kernels", e);
        }
    }
    if (rc == LBM_OK) {
        double sum[6] = {0, 0, 0, 0, 0, 0};
        std::uint64_t launches[2] = {0, 0};
        std::uint64_t first = h->steps - count;
        for (std::uint64_t n = 0; n < count; ++n) {
            const int b_step = ((first + n) & 1U) == 0 ? 1 : 0; // even step count before: a B step
            for (int k = 0; k < 3; ++k) {
                float t = 0.0f;
                (void)hipEventElapsedTime(&t, ev[4U * n + k], ev[4U * n + k + 1U]);
                sum[2 * k + b_step] += t;
            }
            ++launches[b_step];
        }
        for (int k = 0; k < 6; ++k) {
            const std::uint64_t l = launches[k & 1];
            ms[k] = l > 0 ? sum[k] / static_cast<double>(l) : 0.0;
        }
    }
    for (auto& e : ev) {
        if (e != nullptr) {
            (void)hipEventDestroy(e);
        }
    }
    return rc;
}

int lbm3d_set_occupancy(lbm3d* h, int32_t groups_a, int32_t groups_b) {
    if (h == nullptr) {
        return fail(LBM_EINVAL, "lbm3d: null handle");
    }
    for (const int32_t g : {groups_a, groups_b}) {
        if (g < 0 || g > 32) {
            return fail(LBM_EINVAL, "lbm3d: workgroups per compute unit must be 0 (no cap) to 32");
        }
        if (g > 0 && (h->lds_per_cu == 0 || h->lds_per_cu / static_cast<std::uint32_t>(g) / 512U * 512U >
                                                   h->lds_per_block)) {
            return fail(LBM_EINVAL, "lbm3d: the device's LDS cannot express a cap of " + std::to_string(g) +
                                        " workgroups per compute unit");
        }
    }
    h->groups[0] = groups_a;
    h->groups[1] = groups_b;
    return ok();
}

int lbm3d_get_occupancy(const lbm3d* h, int32_t* groups_a, int32_t* groups_b) {
    if (h == nullptr) {
        return fail(LBM_EINVAL, "lbm3d: null handle");
    }
    if (groups_a != nullptr) *groups_a = h->groups[0];
    if (groups_b != nullptr) *groups_b = h->groups[1];
    return ok();
}

int lbm3d_get_launch(const lbm3d* h, int32_t* schedule, int32_t* block, int32_t* cells_per_thread) {
    if (h == nullptr) {
        return fail(LBM_EINVAL, "lbm3d: null handle");
    }
    if (schedule != nullptr) *schedule = h->schedule;
    if (block != nullptr) *block = h->block;
    if (cells_per_thread != nullptr) *cells_per_thread = h->cpt;
    return ok();
}

int lbm3d_autotune(lbm3d* h, uint32_t steps, int32_t serial_only, lbm_launch_choice* tried, size_t capacity,
                   size_t* count, lbm_launch_choice* best) {
    if (h == nullptr) {
        return fail(LBM_EINVAL, "lbm3d: null handle");
    }
    if (h->in_batch) {
        return refuse_in_batch("autotune");
    }
    if (steps == 0) {
        return fail(LBM_EINVAL, "lbm3d: autotune needs at least one timed step");
    }
    const std::uint64_t timed = steps + (steps & 1U); // whole A-B pairs
    // The state a step reads and writes: populations (with their guard cells), the boundary
    // stash and the outlet history. A device copy of it brings the lattice back after timing.
    const std::size_t sb = store_bytes(h->storage);
    const std::size_t rb = real_bytes(h->storage);
    const std::size_t f_bytes = h->layout.elements * sb;
    const std::size_t stash_bytes = 2U * 19U * static_cast<std::size_t>(h->tables.bstride) * rb;
    const std::size_t hist_bytes = 5U * static_cast<std::size_t>(h->tables.bstride) * rb;
    char* f_base = static_cast<char*>(h->f);
    void* snap = nullptr;
    if (hipMalloc(&snap, f_bytes + stash_bytes + hist_bytes + sizeof(std::uint32_t)) != hipSuccess) {
        (void)hipGetLastError();
        return fail(LBM_ENOMEM, "lbm3d: autotune needs a device copy of the lattice state (" +
                                    lbm::host::megabytes(f_bytes + stash_bytes + hist_bytes) + ")");
    }
    char* s = static_cast<char*>(snap);
    hipError_t e = hipMemcpyAsync(s, f_base, f_bytes, hipMemcpyDeviceToDevice, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(s + f_bytes, h->stash, stash_bytes, hipMemcpyDeviceToDevice, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(s + f_bytes + stash_bytes, h->hist, hist_bytes, hipMemcpyDeviceToDevice, h->stream);
    if (e == hipSuccess) {
        e = hipMemcpyAsync(s + f_bytes + stash_bytes + hist_bytes, h->clock, sizeof(std::uint32_t),
                           hipMemcpyDeviceToDevice, h->stream);
    }
    if (e != hipSuccess) {
        (void)hipFree(snap);
        return hip_fail("autotune snapshot", e);
    }
    const int saved_schedule = h->schedule;
    const int saved_block = h->block;
    const int saved_cpt = h->cpt;
    const int saved_cpt_b = h->cpt_b;
    const int saved_groups[2] = {h->groups[0], h->groups[1]};
    const std::uint64_t saved_steps = h->steps;
    std::vector<lbm_launch_choice> cand;
    for (const int block : {256, 192, 128, 64}) {
        for (const int cpt : {1, 2, 4}) {
            cand.push_back({LBM_SCHEDULE_SERIAL, block, cpt, 0, 0, 0.0});
        }
    }
    if (!serial_only) {
        cand.push_back({LBM_SCHEDULE_FUSED, 256, 1, 0, 0, 0.0});
    }
    int rc = LBM_OK;
    const auto time_candidate = [&](lbm_launch_choice& c) {
        h->schedule = c.schedule;
        h->block = c.block;
        h->cpt = c.cells_per_thread;
        h->cpt_b = c.cells_per_thread;
        h->groups[0] = c.groups_a;
        h->groups[1] = c.groups_b;
        double ms = 0.0;
        rc = lbm3d_time_steps(h, 2, 0, &ms, nullptr, nullptr); // warm-up
        if (rc == LBM_OK) {
            rc = lbm3d_time_steps(h, timed, 0, &ms, nullptr, nullptr);
        }
        c.ms_per_step = ms / static_cast<double>(timed);
        return rc == LBM_OK;
    };
    std::size_t final_begin = 0; // candidates before it take no part in the choice
    const auto fastest = [&]() {
        std::size_t k = final_begin;
        for (std::size_t i = final_begin + 1; i < cand.size(); ++i) {
            if (cand[i].ms_per_step < cand[k].ms_per_step) {
                k = i;
            }
        }
        return k;
    };
    for (auto& c : cand) {
        if (!time_candidate(c)) {
            break;
        }
    }
    // Occupancy caps for the bulk kernel of the fastest serial shape: fewer resident
    // workgroups per compute unit stream the 19 planes with fewer conflicts. The A-step cap,
    // then the B-step cap under the best A-step cap, then the A-step cap again under that.
    if (rc == LBM_OK && cand[fastest()].schedule == LBM_SCHEDULE_SERIAL && h->lds_per_cu > 0) {
        for (int pass = 0; pass < 3 && rc == LBM_OK; ++pass) {
            const int side = pass & 1;
            const lbm_launch_choice base = cand[fastest()];
            for (const int g : {7, 6, 5, 4, 3}) {
                lbm_launch_choice c = base;
                (side == 0 ? c.groups_a : c.groups_b) = g;
                if (h->lds_per_cu / static_cast<std::uint32_t>(g) / 512U * 512U > h->lds_per_block) {
                    continue;
                }
                if (!time_candidate(c)) {
                    break;
                }
                cand.push_back(c);
            }
        }
    }
    // The fastest few, timed once more in turn: the first timings of candidates a percent apart
    // are within each other's noise, and the fastest of many noisy timings is biased low.
    if (rc == LBM_OK && cand.size() > 1) {
        std::vector<std::size_t> order(cand.size());
        for (std::size_t i = 0; i < order.size(); ++i) {
            order[i] = i;
        }
        std::sort(order.begin(), order.end(),
                  [&](std::size_t a, std::size_t b) { return cand[a].ms_per_step < cand[b].ms_per_step; });
        std::vector<lbm_launch_choice> finals;
        for (std::size_t i = 0; i < order.size() && finals.size() < 4; ++i) {
            const lbm_launch_choice& c = cand[order[i]];
            bool seen = false;
            for (const auto& f : finals) {
                seen = seen || (f.schedule == c.schedule && f.block == c.block &&
                                f.cells_per_thread == c.cells_per_thread && f.groups_a == c.groups_a &&
                                f.groups_b == c.groups_b);
            }
            if (!seen) {
                finals.push_back(c);
            }
        }
        // The default launch competes in the final round too, so that the choice is never one
        // that the final round timed slower than the default.
        const int default_cpt = lbm::internal::default_cells_per_thread(h->storage);
        bool has_default = false;
        for (const auto& f : finals) {
            has_default = has_default || (f.schedule == LBM_SCHEDULE_SERIAL && f.block == 256 &&
                                          f.cells_per_thread == default_cpt && f.groups_a == 0 && f.groups_b == 0);
        }
        if (!has_default) {
            finals.push_back({LBM_SCHEDULE_SERIAL, 256, default_cpt, 0, 0, 0.0});
        }
        for (auto& c : finals) {
            if (!time_candidate(c)) {
                break;
            }
            cand.push_back(c);
        }
        if (rc == LBM_OK) {
            final_begin = cand.size() - finals.size(); // only the final round competes
        }
    }
    e = hipMemcpyAsync(f_base, s, f_bytes, hipMemcpyDeviceToDevice, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(h->stash, s + f_bytes, stash_bytes, hipMemcpyDeviceToDevice, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(h->hist, s + f_bytes + stash_bytes, hist_bytes, hipMemcpyDeviceToDevice, h->stream);
    if (e == hipSuccess) {
        e = hipMemcpyAsync(h->clock, s + f_bytes + stash_bytes + hist_bytes, sizeof(std::uint32_t),
                           hipMemcpyDeviceToDevice, h->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(snap);
    h->steps = saved_steps;
    h->graph.reset();
    if (e != hipSuccess) {
        return hip_fail("autotune restore", e);
    }
    if (rc != LBM_OK) {
        h->schedule = saved_schedule;
        h->block = saved_block;
        h->cpt = saved_cpt;
        h->cpt_b = saved_cpt_b;
        h->groups[0] = saved_groups[0];
        h->groups[1] = saved_groups[1];
        return rc;
    }
    const std::size_t k = fastest();
    h->schedule = cand[k].schedule;
    h->block = cand[k].block;
    h->cpt = cand[k].cells_per_thread;
    h->cpt_b = cand[k].cells_per_thread;
    h->groups[0] = cand[k].groups_a;
    h->groups[1] = cand[k].groups_b;
    if (best != nullptr) *best = cand[k];
    if (tried != nullptr) {
        for (std::size_t i = 0; i < cand.size() && i < capacity; ++i) {
            tried[i] = cand[i];
        }
    }
    if (count != nullptr) *count = cand.size();
    return ok();
}

} // extern "C"
