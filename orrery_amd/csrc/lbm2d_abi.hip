// C ABI of the D2Q9 device solver (include/orrery_lbm.h): the counterpart of GpuLbm2D and
// make_gpu_solver for Lbm2DConfig (src/aero/lbm/lbm2d_gpu.cpp:68-378).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <new>
#include <string>
#include <vector>

#include "lbm_device.h"
#include "lbm_graph.h"
#include "lbm_host.h"
#include "lbm_kernels.h"

using lbm::host::fail;
using lbm::host::ok;

struct lbm2d {
    lbm2d_config cfg{};
    int storage = LBM_STORE_F32;
    std::size_t cells = 0;
    std::uint32_t stride = 0;
    lbm::host::Tables tables;
    hipStream_t stream = nullptr;
    std::uint64_t steps = 0;
    int block = 256;
    // LBM_SCHEDULE_FUSED: bulk cells and boundary nodes in one launch (default);
    // LBM_SCHEDULE_SERIAL: the bulk kernel, then the boundary kernel.
    int schedule = LBM_SCHEDULE_FUSED;
    void* f = nullptr;
    std::uint8_t* cls = nullptr;
    std::uint32_t* bcell = nullptr;
    std::uint32_t* bcode = nullptr;
    void* stash = nullptr;
    void* hist = nullptr;
    std::uint32_t* lcell = nullptr;
    std::uint8_t* ldir = nullptr;
    std::uint32_t* lslot = nullptr;
    double* partial = nullptr;
    double* result = nullptr;
    void* shift = nullptr;  // the 16-bit formats' storage offsets (Lattice2::shift)
    std::vector<void*> owned;
    // Replayed runs of steps of lbm2d_step.
    lbm::StepGraph graph;
};

namespace {

std::size_t store_bytes2(int storage) {
    return storage == LBM_STORE_F64 ? 8 : (storage == LBM_STORE_F32 ? 4 : 2);
}
std::size_t real_bytes2(int storage) { return storage == LBM_STORE_F64 ? 8 : 4; }

int hip_fail2(const char* what, hipError_t e) {
    return fail(LBM_EDEVICE, std::string("lbm2d gpu: ") + what + ": " + hipGetErrorString(e));
}

template <class T> lbm::Lattice2<typename T::R> descriptor2(const lbm2d& h) {
    using R = typename T::R;
    lbm::Lattice2<R> p;
    std::memset(static_cast<void*>(&p), 0, sizeof(p));  // padding too: the bytes key step graphs
    const lbm2d_config& c = h.cfg;
    p.f = h.f;
    p.cls = h.cls;
    p.bcell = h.bcell;
    p.bcode = h.bcode;
    p.stash = static_cast<R*>(h.stash);
    p.hist = static_cast<R*>(h.hist);
    p.nx = static_cast<std::uint32_t>(c.nx);
    p.ny = static_cast<std::uint32_t>(c.ny);
    p.cells = static_cast<std::uint32_t>(h.cells);
    p.stride = h.stride;
    p.nb = static_cast<std::uint32_t>(h.tables.bcell.size());
    p.bstride = h.tables.bstride;
    p.div_nx = lbm::FastDiv(p.nx);
    p.x_mode = c.x_boundary;
    p.convective = c.outlet == LBM_OUTLET_CONVECTIVE ? 1 : 0;
    p.side[0] = c.bottom.kind;
    p.side[1] = c.top.kind;
    p.wall_u[0] = static_cast<R>(c.bottom.velocity_x);
    p.wall_u[1] = static_cast<R>(c.top.velocity_x);
    p.inlet[0] = static_cast<R>(c.inlet_ux);
    p.inlet[1] = static_cast<R>(c.inlet_uy);
    p.tau = static_cast<R>(c.tau);
    p.omega = static_cast<R>(1) / static_cast<R>(c.tau);
    const double k = c.smagorinsky_enabled ? c.smagorinsky_constant : 0.0;
    p.smag_c = static_cast<R>(k);
    p.smag_c2 = static_cast<R>(k * k);
    p.les = T::shifted ? (p.smag_c2 > R(0) ? 1 : 0) : (c.smagorinsky_enabled ? 1 : 0);
    p.force[0] = static_cast<R>(c.body_force_x);
    p.force[1] = static_cast<R>(c.body_force_y);
    p.forced = (c.body_force_x != 0.0 || c.body_force_y != 0.0) ? 1 : 0;
    p.outlet_speed = static_cast<R>(c.inlet_ux);
    p.shift = static_cast<const R*>(h.shift);
    return p;
}

// The 16-bit formats' storage offsets: each direction's free-stream equilibrium less its weight,
// f_eq_i(1, u_ref) - w_i with u_ref the inlet velocity (zero without an inlet), then the same
// scaled as the format scales its stored values; empty for the other formats.
template <class T> std::vector<typename T::R> storage_offsets2(const lbm2d_config& c) {
    using R = typename T::R;
    std::vector<R> t;
    if constexpr (T::offset) {
        t.assign(18, R(0));
        const bool inflow = c.x_boundary == LBM_X_INFLOW_OUTFLOW;
        const double ux = inflow ? c.inlet_ux : 0.0;
        const double uy = inflow ? c.inlet_uy : 0.0;
        const double uu = ux * ux + uy * uy;
        for (int i = 0; i < 9; ++i) {
            const double cu = lbm::D2Q9::cx(i) * ux + lbm::D2Q9::cy(i) * uy;
            t[i] = static_cast<R>(lbm::D2Q9::w(i) * (3.0 * cu + 4.5 * cu * cu - 1.5 * uu));
            t[9 + i] = t[i] * T::kOffsetScale;  // exact: a power of two
        }
    }
    return t;
}

template <class Fn> int with_store2(const lbm2d& h, Fn&& fn) {
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

void release2(lbm2d* h) {
    if (h == nullptr) {
        return;
    }
    if (h->stream != nullptr) {
        (void)hipStreamSynchronize(h->stream);
    }
    for (void* p : h->owned) {
        (void)hipFree(p);
    }
    if (h->stream != nullptr) {
        (void)hipStreamDestroy(h->stream);
    }
    delete h;
}

template <typename P> int alloc2(lbm2d& h, P** out, std::size_t bytes, const char* what) {
    void* p = nullptr;
    const hipError_t e = hipMalloc(&p, bytes == 0 ? 8 : bytes);
    if (e != hipSuccess) {
        return fail(LBM_ENOMEM, std::string("lbm2d gpu: ") + what + ": " + hipGetErrorString(e));
    }
    h.owned.push_back(p);
    *out = static_cast<P*>(p);
    return ok();
}

template <typename V> int upload2(lbm2d& h, void* dst, const std::vector<V>& src, const char* what) {
    if (src.empty()) {
        return ok();
    }
    const hipError_t e = hipMemcpyAsync(dst, src.data(), src.size() * sizeof(V), hipMemcpyHostToDevice, h.stream);
    return e == hipSuccess ? ok() : hip_fail2(what, e);
}

int sync2(lbm2d& h, const char* what) {
    const hipError_t e = hipStreamSynchronize(h.stream);
    return e == hipSuccess ? ok() : hip_fail2(what, e);
}

int initialize2(lbm2d* h, int mode, const double* host_planes, std::uint64_t seed, double amplitude) {
    if (h == nullptr) {
        return fail(LBM_EINVAL, "lbm2d: null handle");
    }
    const bool inflow = h->cfg.x_boundary == LBM_X_INFLOW_OUTFLOW;
    const double u0[2] = {inflow ? h->cfg.inlet_ux : 0.0, inflow ? h->cfg.inlet_uy : 0.0};
    double* planes = nullptr;
    if (mode == 1) {
        if (hipMalloc(reinterpret_cast<void**>(&planes), 3U * h->cells * sizeof(double)) != hipSuccess) {
            return fail(LBM_ENOMEM, "lbm2d gpu: initial state buffer");
        }
        const hipError_t e =
            hipMemcpyAsync(planes, host_planes, 3U * h->cells * sizeof(double), hipMemcpyHostToDevice, h->stream);
        if (e != hipSuccess) {
            (void)hipFree(planes);
            return hip_fail2("initial state upload", e);
        }
    }
    const int rc = with_store2(*h, [&]<class T>() -> int {
        const hipError_t e = lbm::d2q9_initialize<T>(descriptor2<T>(*h), mode, planes, u0, seed, amplitude, h->stream);
        return e == hipSuccess ? ok() : hip_fail2("initialize", e);
    });
    const int s = sync2(*h, "initialize");
    if (planes != nullptr) {
        (void)hipFree(planes);
    }
    if (rc != LBM_OK) {
        return rc;
    }
    if (s != LBM_OK) {
        return s;
    }
    h->steps = 0;
    return ok();
}

} // namespace

extern "C" {

int lbm2d_create(const lbm2d_config* config, const uint8_t* solid, int32_t storage, lbm2d** out) {
    if (config == nullptr || out == nullptr) {
        return fail(LBM_EINVAL, "lbm2d: null argument");
    }
    *out = nullptr;
    if (storage < LBM_STORE_F32 || storage > LBM_STORE_F64) {
        return fail(LBM_EINVAL, "lbm2d gpu: unknown population storage");
    }
    const lbm2d_config& c = *config;
    if (std::string p = lbm::host::problem2(c, solid); !p.empty()) {
        return fail(LBM_EINVAL, p);
    }
    if (storage != LBM_STORE_F64) {
        if (std::string p = lbm::host::binary32_problem2(c); !p.empty()) {
            return fail(LBM_EINVAL, p);
        }
    }
    const std::size_t cells = c.nx * c.ny;
    const std::uint32_t stride = lbm::host::padded(cells, 128);
    auto* h = new (std::nothrow) lbm2d();
    if (h == nullptr) {
        return fail(LBM_ENOMEM, "lbm2d: host allocation failed");
    }
    h->cfg = c;
    h->storage = storage;
    h->cells = cells;
    h->stride = stride;
    h->tables = lbm::host::tables2(c, solid);
    const lbm::host::Tables& t = h->tables;
    const std::size_t sb = store_bytes2(storage);
    const std::size_t rb = real_bytes2(storage);
    const std::size_t nb = t.bcell.size();
    const std::size_t nl = t.links.size();
    const std::size_t red = std::max<std::size_t>(lbm::reduction_blocks(cells) + lbm::reduction_blocks(nb),
                                                  3U * lbm::reduction_blocks(nl));
    const std::size_t need = 9U * stride * sb + cells + nb * 4U + 9U * t.bstride * 4U + 2U * 9U * t.bstride * rb +
                             3U * t.bstride * rb + nl * 9U + (red + 8U) * 8U;
    std::size_t available = 0;
    if (int rc = lbm::host::device_memory(&available, nullptr); rc != LBM_OK) {
        delete h;
        return rc;
    }
    if (need > available) {
        delete h;
        return fail(LBM_ENOMEM, "lbm2d gpu: the lattice needs " + lbm::host::megabytes(need) + " of device memory, " +
                                    lbm::host::megabytes(available) + " are available; use fewer cells");
    }
    hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete h;
        return hip_fail2("stream", e);
    }
    int rc = LBM_OK;
    if ((rc = alloc2(*h, reinterpret_cast<char**>(&h->f), 9U * stride * sb, "population buffer")) ||
        (rc = alloc2(*h, &h->cls, cells, "node class buffer")) || (rc = alloc2(*h, &h->bcell, nb * 4U, "boundary list")) ||
        (rc = alloc2(*h, &h->bcode, 9U * t.bstride * 4U, "source codes")) ||
        (rc = alloc2(*h, reinterpret_cast<char**>(&h->stash), 2U * 9U * t.bstride * rb, "boundary stash")) ||
        (rc = alloc2(*h, reinterpret_cast<char**>(&h->hist), 3U * t.bstride * rb, "outlet buffer")) ||
        (rc = alloc2(*h, &h->lcell, nl * 4U, "link buffer")) || (rc = alloc2(*h, &h->ldir, nl, "link buffer")) ||
        (rc = alloc2(*h, &h->lslot, nl * 4U, "link buffer")) ||
        (rc = alloc2(*h, &h->partial, (red + 8U) * 8U, "reduction buffer"))) {
        release2(h);
        return rc;
    }
    rc = with_store2(*h, [&]<class T>() -> int {
        const std::vector<typename T::R> offsets = storage_offsets2<T>(*config);
        if (offsets.empty()) {
            return ok();
        }
        int r = alloc2(*h, reinterpret_cast<char**>(&h->shift), offsets.size() * sizeof(typename T::R), "storage offsets");
        return r != LBM_OK ? r : upload2(*h, h->shift, offsets, "offset upload");
    });
    if (rc != LBM_OK) {
        release2(h);
        return rc;
    }
    h->result = h->partial + red;
    std::vector<std::uint32_t> lcell(nl);
    std::vector<std::uint8_t> ldir(nl);
    for (std::size_t k = 0; k < nl; ++k) {
        lcell[k] = t.links[k].cell;
        ldir[k] = static_cast<std::uint8_t>(t.links[k].direction);
    }
    if ((rc = upload2(*h, h->cls, t.cls, "class upload")) || (rc = upload2(*h, h->bcell, t.bcell, "boundary upload")) ||
        (rc = upload2(*h, h->bcode, t.bcode, "code upload")) || (rc = upload2(*h, h->lcell, lcell, "link upload")) ||
        (rc = upload2(*h, h->ldir, ldir, "link upload")) || (rc = upload2(*h, h->lslot, t.lslot, "link upload"))) {
        release2(h);
        return rc;
    }
    if ((e = hipMemsetAsync(h->hist, 0, 3U * t.bstride * rb, h->stream)) != hipSuccess ||
        (e = hipMemsetAsync(h->stash, 0, 2U * 9U * t.bstride * rb, h->stream)) != hipSuccess) {
        release2(h);
        return hip_fail2("clear", e);
    }
    if ((rc = initialize2(h, 0, nullptr, 0, 0.0)) != LBM_OK) {
        release2(h);
        return rc;
    }
    *out = h;
    return ok();
}

int lbm2d_destroy(lbm2d* h) {
    release2(h);
    return ok();
}

int lbm2d_initialize(lbm2d* h, const double* moments) { return initialize2(h, moments ? 1 : 0, moments, 0, 0.0); }

int lbm2d_initialize_perturbed(lbm2d* h, uint64_t seed, double amplitude) {
    if (!std::isfinite(amplitude)) {
        return fail(LBM_EINVAL, "lbm2d: the perturbation amplitude must be finite");
    }
    return initialize2(h, 2, nullptr, seed, amplitude);
}

int lbm2d_step(lbm2d* h, uint64_t count) {
    if (h == nullptr) {
        return fail(LBM_EINVAL, "lbm2d: null handle");
    }
    return with_store2(*h, [&]<class T>() -> int {
        const auto p = descriptor2<T>(*h);
        std::vector<char> key;
        lbm::key_append(key, p);
        lbm::key_append(key, h->block);
        lbm::key_append(key, h->schedule);
        const hipError_t e = h->graph.run(h->stream, count, h->steps, key, [&](bool local) {
            return h->schedule == LBM_SCHEDULE_SERIAL ? lbm::d2q9_step_split<T>(p, local, h->block, h->stream)
                                                      : lbm::d2q9_step<T>(p, local, h->block, h->stream);
        });
        return e == hipSuccess ? ok() : hip_fail2("step", e);
    });
}

int lbm2d_set_schedule(lbm2d* h, int32_t schedule) {
    if (h == nullptr) {
        return fail(LBM_EINVAL, "lbm2d: null handle");
    }
    if (schedule != LBM_SCHEDULE_FUSED && schedule != LBM_SCHEDULE_SERIAL) {
        return fail(LBM_EINVAL, "lbm2d: the step schedule must be fused or serial");
    }
    h->schedule = schedule;
    return ok();
}

int lbm2d_set_graphs(lbm2d* h, int32_t enabled) {
    if (h == nullptr) {
        return fail(LBM_EINVAL, "lbm2d: null handle");
    }
    h->graph.enable(enabled != 0);
    return ok();
}

uint64_t lbm2d_steps(const lbm2d* h) { return h != nullptr ? h->steps : 0; }

int lbm2d_force(lbm2d* h, double* fx, double* fy, double* moment) {
    if (h == nullptr || fx == nullptr || fy == nullptr || moment == nullptr) {
        return fail(LBM_EINVAL, "lbm2d: null argument");
    }
    const std::size_t nl = h->tables.links.size();
    if (nl == 0) {
        *fx = *fy = *moment = 0.0;
        return ok();
    }
    const int phase = static_cast<int>(h->steps & 1U);
    const int rc = with_store2(*h, [&]<class T>() -> int {
        const hipError_t e = lbm::d2q9_force_sum<T>(descriptor2<T>(*h), phase, h->lcell, h->ldir, h->lslot,
                                                    static_cast<std::uint32_t>(nl), h->partial, h->result, h->stream);
        return e == hipSuccess ? ok() : hip_fail2("force", e);
    });
    if (rc != LBM_OK) {
        return rc;
    }
    double v[3];
    hipError_t e = hipMemcpyAsync(v, h->result, sizeof(v), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) {
        e = hipStreamSynchronize(h->stream);
    }
    if (e != hipSuccess) {
        return hip_fail2("force readback", e);
    }
    *fx = v[0];
    *fy = v[1];
    *moment = v[2];
    return ok();
}

int lbm2d_mass(lbm2d* h, double* mass) {
    if (h == nullptr || mass == nullptr) {
        return fail(LBM_EINVAL, "lbm2d: null argument");
    }
    const int phase = static_cast<int>(h->steps & 1U);
    const int rc = with_store2(*h, [&]<class T>() -> int {
        const hipError_t e = lbm::d2q9_mass<T>(descriptor2<T>(*h), phase, h->partial, h->result, h->stream);
        return e == hipSuccess ? ok() : hip_fail2("mass", e);
    });
    if (rc != LBM_OK) {
        return rc;
    }
    hipError_t e = hipMemcpyAsync(mass, h->result, sizeof(double), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) {
        e = hipStreamSynchronize(h->stream);
    }
    return e == hipSuccess ? ok() : hip_fail2("mass readback", e);
}

int lbm2d_sample(lbm2d* h, const uint32_t* cells, size_t n, double* rho, double* ux, double* uy) {
    if (h == nullptr || (n > 0 && cells == nullptr)) {
        return fail(LBM_EINVAL, "lbm2d: null argument");
    }
    if (n == 0) {
        return ok();
    }
    if (n > std::numeric_limits<std::int32_t>::max()) {
        return fail(LBM_EINVAL, "lbm2d: too many sample nodes");
    }
    // Boundary slot of each node (the boundary list is in cell order), or -1.
    const std::vector<std::uint32_t>& bcell = h->tables.bcell;
    std::vector<std::int32_t> slots(n);
    for (std::size_t k = 0; k < n; ++k) {
        if (cells[k] >= h->cells) {
            return fail(LBM_EINVAL, "lbm2d: sample node outside the lattice");
        }
        const auto it = std::lower_bound(bcell.begin(), bcell.end(), cells[k]);
        slots[k] = it != bcell.end() && *it == cells[k] ? static_cast<std::int32_t>(it - bcell.begin()) : -1;
    }
    char* buffer = nullptr;
    const std::size_t bytes = n * (4U + 4U + 3U * sizeof(double));
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&buffer), bytes);
    if (e != hipSuccess) {
        return fail(LBM_ENOMEM, std::string("lbm2d gpu: sample buffer: ") + hipGetErrorString(e));
    }
    auto* dcells = reinterpret_cast<std::uint32_t*>(buffer);
    auto* dslots = reinterpret_cast<std::int32_t*>(buffer + 4U * n);
    auto* dout = reinterpret_cast<double*>(buffer + 8U * n);
    e = hipMemcpyAsync(dcells, cells, 4U * n, hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) {
        e = hipMemcpyAsync(dslots, slots.data(), 4U * n, hipMemcpyHostToDevice, h->stream);
    }
    int rc = e == hipSuccess ? LBM_OK : hip_fail2("sample upload", e);
    if (rc == LBM_OK) {
        const int phase = static_cast<int>(h->steps & 1U);
        const double force[2] = {h->cfg.body_force_x, h->cfg.body_force_y};
        rc = with_store2(*h, [&]<class T>() -> int {
            const hipError_t err = lbm::d2q9_sample<T>(descriptor2<T>(*h), phase, force, dcells, dslots,
                                                       static_cast<std::uint32_t>(n), dout, h->stream);
            return err == hipSuccess ? ok() : hip_fail2("sample", err);
        });
    }
    double* outs[3] = {rho, ux, uy};
    for (int k = 0; k < 3 && rc == LBM_OK; ++k) {
        if (outs[k] != nullptr) {
            e = hipMemcpyAsync(outs[k], dout + k * n, n * sizeof(double), hipMemcpyDeviceToHost, h->stream);
            if (e != hipSuccess) {
                rc = hip_fail2("sample readback", e);
            }
        }
    }
    const int sy = sync2(*h, "sample readback");
    (void)hipFree(buffer);
    return rc != LBM_OK ? rc : sy;
}

int lbm2d_max_speed(lbm2d* h, double* speed) {
    if (h == nullptr || speed == nullptr) {
        return fail(LBM_EINVAL, "lbm2d: null argument");
    }
    const int phase = static_cast<int>(h->steps & 1U);
    const double force[2] = {h->cfg.body_force_x, h->cfg.body_force_y};
    const int rc = with_store2(*h, [&]<class T>() -> int {
        const hipError_t e = lbm::d2q9_max_speed<T>(descriptor2<T>(*h), phase, force, h->partial, h->result, h->stream);
        return e == hipSuccess ? ok() : hip_fail2("max speed", e);
    });
    if (rc != LBM_OK) {
        return rc;
    }
    hipError_t e = hipMemcpyAsync(speed, h->result, sizeof(double), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) {
        e = hipStreamSynchronize(h->stream);
    }
    return e == hipSuccess ? ok() : hip_fail2("max speed readback", e);
}

int lbm2d_macroscopic(lbm2d* h, double* rho, double* ux, double* uy) {
    if (h == nullptr) {
        return fail(LBM_EINVAL, "lbm2d: null handle");
    }
    double* planes = nullptr;
    if (hipMalloc(reinterpret_cast<void**>(&planes), 3U * h->cells * sizeof(double)) != hipSuccess) {
        return fail(LBM_ENOMEM, "lbm2d gpu: moment buffer");
    }
    const int phase = static_cast<int>(h->steps & 1U);
    const double force[2] = {h->cfg.body_force_x, h->cfg.body_force_y};
    int rc = with_store2(*h, [&]<class T>() -> int {
        const hipError_t e = lbm::d2q9_moments<T>(descriptor2<T>(*h), phase, force, planes, h->stream);
        return e == hipSuccess ? ok() : hip_fail2("moments", e);
    });
    double* outs[3] = {rho, ux, uy};
    for (int k = 0; k < 3 && rc == LBM_OK; ++k) {
        if (outs[k] != nullptr) {
            const hipError_t e = hipMemcpyAsync(outs[k], planes + k * h->cells, h->cells * sizeof(double),
                                                hipMemcpyDeviceToHost, h->stream);
            if (e != hipSuccess) {
                rc = hip_fail2("moment readback", e);
            }
        }
    }
    const int s = sync2(*h, "moment readback");
    (void)hipFree(planes);
    return rc != LBM_OK ? rc : s;
}

int lbm2d_populations(lbm2d* h, double* f) {
    if (h == nullptr || f == nullptr) {
        return fail(LBM_EINVAL, "lbm2d: null argument");
    }
    double* buffer = nullptr;
    const std::size_t bytes = 9U * h->cells * sizeof(double);
    if (hipMalloc(reinterpret_cast<void**>(&buffer), bytes) != hipSuccess) {
        return fail(LBM_ENOMEM, "lbm2d gpu: population readback buffer");
    }
    const int phase = static_cast<int>(h->steps & 1U);
    int rc = with_store2(*h, [&]<class T>() -> int {
        const hipError_t e = lbm::d2q9_populations<T>(descriptor2<T>(*h), phase, buffer, h->stream);
        return e == hipSuccess ? ok() : hip_fail2("populations", e);
    });
    if (rc == LBM_OK) {
        const hipError_t e = hipMemcpyAsync(f, buffer, bytes, hipMemcpyDeviceToHost, h->stream);
        if (e != hipSuccess) {
            rc = hip_fail2("population readback", e);
        }
    }
    const int s = sync2(*h, "population readback");
    (void)hipFree(buffer);
    return rc != LBM_OK ? rc : s;
}

int lbm2d_links(const lbm2d* h, uint32_t* cell, uint32_t* direction, size_t capacity, size_t* count) {
    if (h == nullptr) {
        return fail(LBM_EINVAL, "lbm2d: null handle");
    }
    const auto& links = h->tables.links;
    for (std::size_t k = 0; k < links.size() && k < capacity; ++k) {
        if (cell != nullptr) cell[k] = links[k].cell;
        if (direction != nullptr) direction[k] = links[k].direction;
    }
    if (count != nullptr) {
        *count = links.size();
    }
    return ok();
}

int lbm2d_synchronize(lbm2d* h) {
    if (h == nullptr) {
        return fail(LBM_EINVAL, "lbm2d: null handle");
    }
    return sync2(*h, "synchronize");
}

} // extern "C"
