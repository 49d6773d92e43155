// orrery_lbm_solver.hpp -- the C ABI of liborrery_lbm.so behind Orrery's own solver interface.
//
// HipLbm3D implements orrery::aero::lbm::Lbm3DSolver
// (src/aero/lbm3d/include/orrery/aero/lbm3d/lbm3d.h:196-232) and HipLbm2D implements
// orrery::aero::lbm::Lbm2DSolver (src/aero/lbm/include/orrery/aero/lbm/lbm2d.h:173-199) on one
// MI355X lattice; the make_hip_solver() overloads have the shape of make_gpu_solver
// (lbm3d.h:246-247, lbm2d.h:210-211) without the compute backend: a caller that holds
// `make_gpu_solver(backend, config, mask)` (LbmAeroSolver3D::setup,
// src/aero/lbm3d/lbm_aero_solver3d.cpp:312-313; LbmAeroSolver2D::setup,
// src/aero/lbm/lbm_aero_solver.cpp:145-146) switches by calling
// `orrery_amd::make_hip_solver(config, mask)` instead. Header-only; compile it together with
// the reference's headers on the include path and link liborrery_lbm.so.
#pragma once

#include <array>
#include <cstdint>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "orrery/aero/lbm/lbm2d.h"
#include "orrery/aero/lbm3d/lbm3d.h"
#include "orrery_lbm.h"

namespace orrery_amd {

namespace ref = orrery::aero::lbm;

inline lbm_wall to_c(const ref::Wall& w) {
    return lbm_wall{static_cast<int32_t>(w.kind), w.velocity_x};
}

// Lbm3DConfig -> lbm3d_config, field for field (the enum values are the reference's).
inline lbm3d_config to_c(const ref::Lbm3DConfig& c) {
    lbm3d_config o{};
    o.nx = c.nx;
    o.ny = c.ny;
    o.nz = c.nz;
    o.tau = c.tau;
    o.x_boundary = static_cast<int32_t>(c.x_boundary);
    o.outlet = static_cast<int32_t>(c.outlet);
    o.y_low = to_c(c.y_low);
    o.y_high = to_c(c.y_high);
    o.z_low = to_c(c.z_low);
    o.z_high = to_c(c.z_high);
    o.inlet_ux = c.inlet_ux;
    o.inlet_uy = c.inlet_uy;
    o.inlet_uz = c.inlet_uz;
    o.body_force_x = c.body_force_x;
    o.body_force_y = c.body_force_y;
    o.body_force_z = c.body_force_z;
    o.smagorinsky_enabled = c.smagorinsky.enabled ? 1 : 0;
    o.smagorinsky_constant = c.smagorinsky.constant;
    o.sponge_strength = c.sponge.strength;
    o.sponge_inlet = c.sponge.inlet;
    o.sponge_outlet = c.sponge.outlet;
    o.sponge_sides = c.sponge.sides;
    o.solid_wall = static_cast<int32_t>(c.solid_wall);
    return o;
}

inline std::string last_error() { return std::string(lbm_last_error()); }

class HipLbm3D final : public ref::Lbm3DSolver {
public:
    using Made = orrery::core::Result<std::unique_ptr<ref::Lbm3DSolver>, std::string>;

    // make_gpu_solver's contract: refusals before allocating, with the reference's reasons;
    // initialised with uniform_inlet_state(config). `storage` is LBM_STORE_*.
    static Made create(const ref::Lbm3DConfig& config, ref::SolidMask3D mask, int32_t storage = LBM_STORE_F32) {
        const lbm3d_config c = to_c(config);
        // The C ABI takes the mask's arrays by pointer, sized by the configuration: check the
        // shapes here, as problem_of does (lbm3d_common.cpp:82-84,127-129).
        if (mask.nx != config.nx || mask.ny != config.ny || mask.nz != config.nz ||
            mask.solid.size() != config.nx * config.ny * config.nz) {
            return orrery::core::failure(std::string("lbm3d: the solid mask does not match the lattice size"));
        }
        lbm3d* h = nullptr;
        const double* distance = mask.distance.size() == mask.solid.size() ? mask.distance.data() : nullptr;
        if (lbm3d_create(&c, mask.solid.data(), distance, storage, &h) != LBM_OK) {
            return orrery::core::failure(last_error());
        }
        std::unique_ptr<HipLbm3D> s(new HipLbm3D(config, std::move(mask), h));
        size_t n = 0;
        if (lbm3d_links(h, nullptr, nullptr, nullptr, 0, &n) != LBM_OK) {
            return orrery::core::failure(last_error());
        }
        std::vector<uint32_t> cell(n);
        std::vector<uint32_t> dir(n);
        std::vector<double> q(n);
        if (lbm3d_links(h, cell.data(), dir.data(), q.data(), n, &n) != LBM_OK) {
            return orrery::core::failure(last_error());
        }
        s->links_.resize(n);
        for (size_t k = 0; k < n; ++k) {
            s->links_[k] = ref::BoundaryLink3D{cell[k], dir[k], q[k]};
        }
        return std::unique_ptr<ref::Lbm3DSolver>(std::move(s));
    }

    ~HipLbm3D() override { lbm3d_destroy(h_); }

    const ref::Lbm3DConfig& config() const noexcept override { return config_; }
    const ref::SolidMask3D& mask() const noexcept override { return mask_; }
    std::uint64_t steps() const noexcept override { return lbm3d_steps(h_); }

    ref::Status initialize(const ref::InitialState3D& initial) override {
        const size_t cells = config_.nx * config_.ny * config_.nz;
        std::vector<double> planes(4 * cells);
        for (size_t z = 0; z < config_.nz; ++z) {
            for (size_t y = 0; y < config_.ny; ++y) {
                for (size_t x = 0; x < config_.nx; ++x) {
                    const size_t i = x + config_.nx * (y + config_.ny * z);
                    const auto m = initial(x, y, z); // d3q19::Moments (d3q19.h:40-45)
                    planes[i] = m.rho;
                    planes[cells + i] = m.ux;
                    planes[2 * cells + i] = m.uy;
                    planes[3 * cells + i] = m.uz;
                }
            }
        }
        return status(lbm3d_initialize(h_, planes.data()));
    }

    ref::Status step(std::size_t count) override { return status(lbm3d_step(h_, count)); }

    ref::Status set_inlet_velocity(const std::array<double, 3>& velocity) override {
        return status(lbm3d_set_inlet(h_, velocity.data()));
    }

    std::array<double, 3> inlet_velocity() const noexcept override {
        std::array<double, 3> v{};
        lbm3d_get_inlet(h_, v.data());
        return v;
    }

    orrery::core::Result<ref::MacroFields3D, std::string> macroscopic() override {
        ref::MacroFields3D m;
        m.nx = config_.nx;
        m.ny = config_.ny;
        m.nz = config_.nz;
        const size_t cells = m.nx * m.ny * m.nz;
        m.density.resize(cells);
        m.velocity_x.resize(cells);
        m.velocity_y.resize(cells);
        m.velocity_z.resize(cells);
        if (lbm3d_macroscopic(h_, m.density.data(), m.velocity_x.data(), m.velocity_y.data(),
                              m.velocity_z.data()) != LBM_OK) {
            return orrery::core::failure(last_error());
        }
        m.solid = mask_.solid;
        return m;
    }

    orrery::core::Result<std::vector<double>, std::string> populations() override {
        std::vector<double> f(19 * config_.nx * config_.ny * config_.nz);
        if (lbm3d_populations(h_, f.data()) != LBM_OK) {
            return orrery::core::failure(last_error());
        }
        return f;
    }

    orrery::core::Result<ref::Force3D, std::string> force() override {
        ref::Force3D f;
        if (lbm3d_force(h_, f.force.data(), f.moment.data()) != LBM_OK) {
            return orrery::core::failure(last_error());
        }
        return f;
    }

    orrery::core::Result<double, std::string> mass() override {
        double m = 0.0;
        if (lbm3d_mass(h_, &m) != LBM_OK) {
            return orrery::core::failure(last_error());
        }
        return m;
    }

    const std::vector<ref::BoundaryLink3D>& links() const noexcept override { return links_; }

    // The library handle, for the calls the interface does not carry (timing, launch shape).
    lbm3d* handle() const noexcept { return h_; }

private:
    HipLbm3D(const ref::Lbm3DConfig& config, ref::SolidMask3D mask, lbm3d* h)
        : config_(config), mask_(std::move(mask)), h_(h) {}

    static ref::Status status(int rc) {
        if (rc != LBM_OK) {
            return orrery::core::failure(last_error());
        }
        return std::monostate{};
    }

    ref::Lbm3DConfig config_;
    ref::SolidMask3D mask_;
    lbm3d* h_ = nullptr;
    std::vector<ref::BoundaryLink3D> links_;
};

// make_gpu_solver without the backend: the HIP lattice in binary32 storage.
inline HipLbm3D::Made make_hip_solver(const ref::Lbm3DConfig& config, ref::SolidMask3D mask,
                                      int32_t storage = LBM_STORE_F32) {
    return HipLbm3D::create(config, std::move(mask), storage);
}

// Lbm2DConfig -> lbm2d_config, field for field.
inline lbm2d_config to_c(const ref::Lbm2DConfig& c) {
    lbm2d_config o{};
    o.nx = c.nx;
    o.ny = c.ny;
    o.tau = c.tau;
    o.x_boundary = static_cast<int32_t>(c.x_boundary);
    o.outlet = static_cast<int32_t>(c.outlet);
    o.bottom = to_c(c.bottom);
    o.top = to_c(c.top);
    o.inlet_ux = c.inlet_ux;
    o.inlet_uy = c.inlet_uy;
    o.body_force_x = c.body_force_x;
    o.body_force_y = c.body_force_y;
    o.smagorinsky_enabled = c.smagorinsky.enabled ? 1 : 0;
    o.smagorinsky_constant = c.smagorinsky.constant;
    return o;
}

class HipLbm2D final : public ref::Lbm2DSolver {
public:
    using Made = orrery::core::Result<std::unique_ptr<ref::Lbm2DSolver>, std::string>;

    // make_gpu_solver's contract (lbm2d_gpu.cpp:68-378): refusals before allocating, with the
    // reference's reasons; initialised with uniform_inlet_state(config).
    static Made create(const ref::Lbm2DConfig& config, ref::SolidMask2D mask, int32_t storage = LBM_STORE_F32) {
        const lbm2d_config c = to_c(config);
        if (mask.nx != config.nx || mask.ny != config.ny || mask.solid.size() != config.nx * config.ny) {
            return orrery::core::failure(std::string("lbm2d: the solid mask does not match the lattice size"));
        }
        lbm2d* h = nullptr;
        if (lbm2d_create(&c, mask.solid.data(), storage, &h) != LBM_OK) {
            return orrery::core::failure(last_error());
        }
        return std::unique_ptr<ref::Lbm2DSolver>(new HipLbm2D(config, std::move(mask), h));
    }

    ~HipLbm2D() override { lbm2d_destroy(h_); }

    const ref::Lbm2DConfig& config() const noexcept override { return config_; }
    const ref::SolidMask2D& mask() const noexcept override { return mask_; }
    std::uint64_t steps() const noexcept override { return lbm2d_steps(h_); }

    ref::Status initialize(const ref::InitialState& initial) override {
        const size_t cells = config_.nx * config_.ny;
        std::vector<double> planes(3 * cells);
        for (size_t y = 0; y < config_.ny; ++y) {
            for (size_t x = 0; x < config_.nx; ++x) {
                const size_t i = x + config_.nx * y;
                const auto m = initial(x, y); // d2q9::Moments (d2q9.h:27-31)
                planes[i] = m.rho;
                planes[cells + i] = m.ux;
                planes[2 * cells + i] = m.uy;
            }
        }
        return status(lbm2d_initialize(h_, planes.data()));
    }

    ref::Status step(std::size_t count) override { return status(lbm2d_step(h_, count)); }

    orrery::core::Result<ref::MacroFields2D, std::string> macroscopic() override {
        ref::MacroFields2D m;
        m.nx = config_.nx;
        m.ny = config_.ny;
        const size_t cells = m.nx * m.ny;
        m.density.resize(cells);
        m.velocity_x.resize(cells);
        m.velocity_y.resize(cells);
        if (lbm2d_macroscopic(h_, m.density.data(), m.velocity_x.data(), m.velocity_y.data()) != LBM_OK) {
            return orrery::core::failure(last_error());
        }
        m.solid = mask_.solid;
        return m;
    }

    orrery::core::Result<std::vector<double>, std::string> populations() override {
        std::vector<double> f(9 * config_.nx * config_.ny);
        if (lbm2d_populations(h_, f.data()) != LBM_OK) {
            return orrery::core::failure(last_error());
        }
        return f;
    }

    orrery::core::Result<ref::Force2D, std::string> force() override {
        ref::Force2D f;
        if (lbm2d_force(h_, &f.x, &f.y, &f.moment) != LBM_OK) {
            return orrery::core::failure(last_error());
        }
        return f;
    }

    orrery::core::Result<double, std::string> mass() override {
        double m = 0.0;
        if (lbm2d_mass(h_, &m) != LBM_OK) {
            return orrery::core::failure(last_error());
        }
        return m;
    }

    // The library handle, for the calls the interface does not carry (graphs, schedule).
    lbm2d* handle() const noexcept { return h_; }

private:
    HipLbm2D(const ref::Lbm2DConfig& config, ref::SolidMask2D mask, lbm2d* h)
        : config_(config), mask_(std::move(mask)), h_(h) {}

    static ref::Status status(int rc) {
        if (rc != LBM_OK) {
            return orrery::core::failure(last_error());
        }
        return std::monostate{};
    }

    ref::Lbm2DConfig config_;
    ref::SolidMask2D mask_;
    lbm2d* h_ = nullptr;
};

// make_gpu_solver (lbm2d.h:210-211) without the backend.
inline HipLbm2D::Made make_hip_solver(const ref::Lbm2DConfig& config, ref::SolidMask2D mask,
                                      int32_t storage = LBM_STORE_F32) {
    return HipLbm2D::create(config, std::move(mask), storage);
}

} // namespace orrery_amd
