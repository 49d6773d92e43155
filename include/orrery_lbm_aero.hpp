// orrery_lbm_aero.hpp -- the library's aerodynamic adapters behind Orrery's IAeroSolver.
//
// HipLbmAeroSolver3D and HipLbmAeroSolver2D implement orrery::aero::IAeroSolver
// (src/aero/include/orrery/aero/aero_solver.h:36-56) with the options of LbmAeroSolver3D
// (LbmSolver3DOptions, src/aero/lbm3d/include/orrery/aero/lbm3d/lbm_aero_solver3d.h:45-81) and
// LbmAeroSolver2D (LbmSolverOptions, src/aero/lbm/include/orrery/aero/lbm/lbm_aero_solver.h:38-52)
// over lbm_aero3d_* and lbm_aero2d_* of orrery_lbm.h. A caller that constructs
// LbmAeroSolver3D(options) or LbmAeroSolver2D(options) with a GPU backend constructs the Hip
// class instead; options.backend is not used (the lattice runs on the current HIP device).
// Header-only; compile with Orrery's include directories on the path and link liborrery_lbm.so.
#pragma once

#include <cstdint>
#include <optional>
#include <stdexcept>
#include <string>
#include <vector>

#include "orrery/aero/aero_solver.h"
#include "orrery/aero/lbm/lbm_aero_solver.h"
#include "orrery/aero/lbm3d/lbm_aero_solver3d.h"
#include "orrery_lbm.h"

namespace orrery_amd {

class HipLbmAeroSolver3D final : public orrery::aero::IAeroSolver {
public:
    explicit HipLbmAeroSolver3D(const orrery::aero::lbm::LbmSolver3DOptions& options = {},
                                int32_t storage = LBM_STORE_F32) {
        lbm_aero3d_options o{};
        lbm_aero3d_default_options(&o);
        o.cells_per_length = options.resolution.cells_per_length;
        o.lattice_velocity = options.resolution.lattice_velocity;
        o.upstream = options.upstream;
        o.downstream = options.downstream;
        o.width = options.width;
        o.sides = static_cast<int32_t>(options.sides);
        o.wall = static_cast<int32_t>(options.wall);
        o.les = static_cast<int32_t>(options.les);
        o.smagorinsky_constant = options.smagorinsky_constant;
        o.duration = options.duration;
        o.average_window = options.average_window;
        o.start_ramp = options.start_ramp;
        o.sponge_strength = options.sponge_strength;
        o.sponge_inlet = options.sponge_inlet;
        o.sponge_outlet = options.sponge_outlet;
        o.sponge_sides = options.sponge_sides;
        o.steps_per_call = options.steps_per_call;
        o.sample_every = options.sample_every;
        o.max_cells = options.max_cells;
        o.storage = storage;
        if (lbm_aero3d_create(&o, &h_) != LBM_OK) {
            throw std::invalid_argument(lbm_aero3d_last_error());
        }
    }
    HipLbmAeroSolver3D(const HipLbmAeroSolver3D&) = delete;
    HipLbmAeroSolver3D& operator=(const HipLbmAeroSolver3D&) = delete;
    ~HipLbmAeroSolver3D() override { lbm_aero3d_destroy(h_); }

    orrery::aero::SolverKind kind() const noexcept override { return orrery::aero::SolverKind::lattice_boltzmann; }

    // Throws std::invalid_argument with the reference's reasons (setup refuses before the
    // lattice is allocated).
    void setup(const orrery::aero::AeroGeometry& geometry, const orrery::aero::FlowConditions& conditions) override {
        std::vector<double> v;
        v.reserve(3 * geometry.mesh.vertices.size());
        for (const auto& p : geometry.mesh.vertices) {
            v.insert(v.end(), {p.x, p.y, p.z});
        }
        std::vector<uint32_t> t;
        t.reserve(3 * geometry.mesh.triangles.size());
        for (const auto& tri : geometry.mesh.triangles) {
            t.insert(t.end(), {tri[0], tri[1], tri[2]});
        }
        lbm_mesh* mesh = nullptr;
        if (lbm_mesh_create(v.data(), v.size() / 3, t.data(), t.size() / 3, &mesh) != LBM_OK) {
            throw std::invalid_argument(std::string("lbm3d: invalid mesh: ") + lbm_geom_last_error());
        }
        const lbm_reference_values r{geometry.reference.area,
                                     geometry.reference.length,
                                     {geometry.reference.moment_centre.x, geometry.reference.moment_centre.y,
                                      geometry.reference.moment_centre.z}};
        const lbm_flow_conditions c{conditions.speed,           conditions.density, conditions.viscosity,
                                    conditions.temperature,     conditions.angle_of_attack,
                                    conditions.sideslip,        conditions.gamma,   conditions.gas_constant};
        const int rc = lbm_aero3d_setup(h_, mesh, &r, &c);
        lbm_mesh_destroy(mesh);
        if (rc != LBM_OK) {
            throw std::invalid_argument(lbm_aero3d_last_error());
        }
        result_ = orrery::aero::AeroResult{};
        read_warnings();
    }

    bool step() override {
        int32_t finished = 0;
        if (lbm_aero3d_step(h_, &finished) != LBM_OK) {
            throw std::runtime_error(lbm_aero3d_last_error());
        }
        read_result();
        return finished != 0;
    }

    const orrery::aero::AeroResult& result() const noexcept override { return result_; }

    lbm_aero3d* handle() const noexcept { return h_; }

private:
    void read_warnings() {
        size_t n = 0;
        lbm_aero3d_warning_count(h_, &n);
        result_.warnings.clear();
        for (size_t k = 0; k < n; ++k) {
            char code[64];
            char message[1024];
            lbm_aero3d_warning(h_, k, code, sizeof(code), message, sizeof(message));
            result_.warnings.push_back({code, message});
        }
    }

    void read_result() {
        lbm_aero3d_result r{};
        lbm_aero3d_get_result(h_, &r);
        orrery::aero::AeroResult next;
        next.cd = r.cd;
        next.cl = r.cl;
        next.cy = r.cy;
        next.cm = r.cm;
        next.force = {r.force[0], r.force[1], r.force[2]};
        next.moment = {r.moment[0], r.moment[1], r.moment[2]};
        next.iterations = r.iterations;
        next.converged = r.converged != 0;
        size_t n = 0;
        lbm_aero3d_residuals(h_, nullptr, 0, &n);
        next.residuals.resize(n);
        lbm_aero3d_residuals(h_, next.residuals.data(), n, &n);
        lbm_aero3d_surface(h_, nullptr, nullptr, nullptr, nullptr, 0, &n);
        std::vector<double> points(3 * n);
        std::vector<double> normals(3 * n);
        next.surface.areas.resize(n);
        next.surface.cp.resize(n);
        lbm_aero3d_surface(h_, points.data(), normals.data(), next.surface.areas.data(), next.surface.cp.data(), n, &n);
        for (size_t k = 0; k < n; ++k) {
            next.surface.points.push_back({points[3 * k], points[3 * k + 1], points[3 * k + 2]});
            next.surface.normals.push_back({normals[3 * k], normals[3 * k + 1], normals[3 * k + 2]});
        }
        result_ = std::move(next);
        read_warnings();
    }

    lbm_aero3d* h_ = nullptr;
    orrery::aero::AeroResult result_;
};

class HipLbmAeroSolver2D final : public orrery::aero::IAeroSolver {
public:
    explicit HipLbmAeroSolver2D(const orrery::aero::lbm::LbmSolverOptions& options = {},
                                int32_t storage = LBM_STORE_F32) {
        lbm_aero2d_options o{};
        lbm_aero2d_default_options(&o);
        o.cells_per_length = options.resolution.cells_per_length;
        o.lattice_velocity = options.resolution.lattice_velocity;
        o.upstream = options.upstream;
        o.downstream = options.downstream;
        o.height = options.height;
        o.les = static_cast<int32_t>(options.les);
        o.smagorinsky_constant = options.smagorinsky_constant;
        o.duration = options.duration;
        o.average_window = options.average_window;
        o.steps_per_call = options.steps_per_call;
        o.sample_every = options.sample_every;
        o.storage = storage;
        if (lbm_aero2d_create(&o, &h_) != LBM_OK) {
            throw std::invalid_argument(lbm_aero2d_last_error());
        }
    }
    HipLbmAeroSolver2D(const HipLbmAeroSolver2D&) = delete;
    HipLbmAeroSolver2D& operator=(const HipLbmAeroSolver2D&) = delete;
    ~HipLbmAeroSolver2D() override { lbm_aero2d_destroy(h_); }

    orrery::aero::SolverKind kind() const noexcept override { return orrery::aero::SolverKind::lattice_boltzmann; }

    // Throws std::invalid_argument with the reference's reasons (setup refuses before the
    // lattice is allocated).
    void setup(const orrery::aero::AeroGeometry& geometry, const orrery::aero::FlowConditions& conditions) override {
        std::vector<double> section;
        section.reserve(2 * geometry.section.size());
        for (const auto& p : geometry.section) {
            section.insert(section.end(), {p.x, p.z});
        }
        const lbm_reference_values r{geometry.reference.area,
                                     geometry.reference.length,
                                     {geometry.reference.moment_centre.x, geometry.reference.moment_centre.y,
                                      geometry.reference.moment_centre.z}};
        const lbm_flow_conditions c{conditions.speed,           conditions.density, conditions.viscosity,
                                    conditions.temperature,     conditions.angle_of_attack,
                                    conditions.sideslip,        conditions.gamma,   conditions.gas_constant};
        if (lbm_aero2d_setup(h_, section.data(), section.size() / 2, &r, &c) != LBM_OK) {
            throw std::invalid_argument(lbm_aero2d_last_error());
        }
        result_ = orrery::aero::AeroResult{};
        read_warnings();
    }

    bool step() override {
        int32_t finished = 0;
        if (lbm_aero2d_step(h_, &finished) != LBM_OK) {
            throw std::runtime_error(lbm_aero2d_last_error());
        }
        read_result();
        return finished != 0;
    }

    const orrery::aero::AeroResult& result() const noexcept override { return result_; }

    // LbmAeroSolver2D::strouhal (lbm_aero_solver.cpp:249-263).
    std::optional<double> strouhal() const {
        double st = 0.0;
        int32_t found = 0;
        if (lbm_aero2d_strouhal(h_, &st, &found) != LBM_OK || found == 0) {
            return std::nullopt;
        }
        return st;
    }

    lbm_aero2d* handle() const noexcept { return h_; }

private:
    void read_warnings() {
        size_t n = 0;
        lbm_aero2d_warning_count(h_, &n);
        result_.warnings.clear();
        for (size_t k = 0; k < n; ++k) {
            char code[64];
            char message[1024];
            lbm_aero2d_warning(h_, k, code, sizeof(code), message, sizeof(message));
            result_.warnings.push_back({code, message});
        }
    }

    void read_result() {
        lbm_aero3d_result r{};
        lbm_aero2d_get_result(h_, &r);
        orrery::aero::AeroResult next;
        next.cd = r.cd;
        next.cl = r.cl;
        next.cy = r.cy;
        next.cm = r.cm;
        next.force = {r.force[0], r.force[1], r.force[2]};
        next.moment = {r.moment[0], r.moment[1], r.moment[2]};
        next.iterations = r.iterations;
        next.converged = r.converged != 0;
        size_t n = 0;
        lbm_aero2d_residuals(h_, nullptr, 0, &n);
        next.residuals.resize(n);
        lbm_aero2d_residuals(h_, next.residuals.data(), n, &n);
        lbm_aero2d_surface(h_, nullptr, nullptr, nullptr, nullptr, 0, &n);
        std::vector<double> points(3 * n);
        std::vector<double> normals(3 * n);
        next.surface.areas.resize(n);
        next.surface.cp.resize(n);
        lbm_aero2d_surface(h_, points.data(), normals.data(), next.surface.areas.data(), next.surface.cp.data(), n, &n);
        for (size_t k = 0; k < n; ++k) {
            next.surface.points.push_back({points[3 * k], points[3 * k + 1], points[3 * k + 2]});
            next.surface.normals.push_back({normals[3 * k], normals[3 * k + 1], normals[3 * k + 2]});
        }
        result_ = std::move(next);
        read_warnings();
    }

    lbm_aero2d* h_ = nullptr;
    orrery::aero::AeroResult result_;
};

} // namespace orrery_amd
