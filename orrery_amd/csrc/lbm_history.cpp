// Force-coefficient history analysis on the host: Cd/Cl, window means, the dominant (shedding)
// oscillation and the convergence tests (steady, settled periodic shedding, and the 3D adapter's
// per-component combination of the two).
//
// The results equal those of src/aero/lbm/force_history.cpp:9-193 and
// src/aero/lbm3d/lbm_aero_solver3d.cpp:60-83 bit for bit. Two arithmetic choices fix that:
//   - a mean is one ascending-index sum over the selected samples followed by one division by
//     their count;
//   - an upward mean crossing between samples (t0, a) and (t1, b) (a, b relative to the mean) lies
//     at t0 + (a / (a - b)) (t1 - t0).
// Minima, maxima and comparisons are exact, so their order is free. Two window tests appear:
// the open tail window keeps a sample when t >= from; the closed span [lo, hi] drops it when
// t < lo or t > hi.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstddef>

#include "../../include/orrery_lbm.h"

namespace {

// Running sum, least and greatest value of the samples fed to it.
struct Tally {
    std::size_t n = 0;
    double total = 0.0;
    double low = 0.0;
    double high = 0.0;

    void take(double v) {
        low = n == 0 ? v : std::min(low, v);
        high = n == 0 ? v : std::max(high, v);
        total += v;
        ++n;
    }
    double average() const { return n == 0 ? 0.0 : total / static_cast<double>(n); }
    double width() const { return high - low; }
};

Tally tail_tally(const double* t, const double* v, std::size_t n, double from) {
    Tally s;
    for (std::size_t k = 0; k < n; ++k) {
        if (t[k] >= from) {
            s.take(v[k]);
        }
    }
    return s;
}

bool in_span(double t, double lo, double hi) { return !(t < lo || t > hi); }

// Upward crossings of `level`, each counted only once the signal has been below level - band
// (hysteresis against noise about the mean). Keeps the first and last crossing and the shortest
// and longest interval between consecutive ones.
class UpCrossings {
public:
    UpCrossings(double level, double band) : level_(level), band_(band) {}

    void segment(double t0, double v0, double t1, double v1) {
        const double a = v0 - level_;
        const double b = v1 - level_;
        if (a < -band_) {
            primed_ = true;
        }
        if (primed_ && a < 0.0 && b >= 0.0) {
            mark(t0 + a / (a - b) * (t1 - t0));
            primed_ = false;
        }
    }

    std::size_t count() const { return count_; }

    // Valid for count() >= 3.
    lbm_oscillation summary(double amplitude) const {
        lbm_oscillation o{};
        o.cycles = count_ - 1;
        o.period = (last_ - first_) / static_cast<double>(o.cycles);
        o.frequency = 1.0 / o.period;
        o.amplitude = amplitude;
        o.period_spread = (widest_ - narrowest_) / o.period;
        return o;
    }

private:
    void mark(double when) {
        if (count_ == 0) {
            first_ = when;
        } else {
            const double gap = when - last_;
            narrowest_ = count_ == 1 ? gap : std::min(narrowest_, gap);
            widest_ = count_ == 1 ? gap : std::max(widest_, gap);
        }
        last_ = when;
        ++count_;
    }

    double level_;
    double band_;
    bool primed_ = false;
    std::size_t count_ = 0;
    double first_ = 0.0;
    double last_ = 0.0;
    double narrowest_ = 0.0;
    double widest_ = 0.0;
};

// force_history.cpp:49-100: period from the mean crossings of the samples at t >= from; needs at
// least three crossings (two full cycles).
bool oscillation_of(const double* t, const double* v, std::size_t n, double from, lbm_oscillation* out) {
    const Tally s = tail_tally(t, v, n, from);
    UpCrossings clock(s.average(), 0.25 * s.width());
    for (std::size_t k = 1; k < n; ++k) {
        if (!(t[k - 1] < from)) {
            clock.segment(t[k - 1], v[k - 1], t[k], v[k]);
        }
    }
    if (clock.count() < 3) {
        return false;
    }
    *out = clock.summary(0.5 * s.width());
    return true;
}

bool close_to(double a, double b, double tolerance) {
    return std::abs(a - b) <= tolerance * std::max(std::abs(a), std::abs(b));
}

// Mean drag and lift width over the closed span [lo, hi].
void span_summary(const double* t, const double* drag, const double* lift, std::size_t n, double lo, double hi,
                  Tally* d, Tally* l) {
    for (std::size_t k = 0; k < n; ++k) {
        if (in_span(t[k], lo, hi)) {
            d->take(drag[k]);
            l->take(lift[k]);
        }
    }
}

// force_history.cpp:159-193: lift shedding with consistent periods, at least two periods of
// samples, and the first and last period agreeing in mean drag and lift width.
bool settled(const double* t, const double* drag, const double* lift, std::size_t n, double from,
             double max_period_spread, double stationarity, lbm_oscillation* out) {
    lbm_oscillation o{};
    if (!oscillation_of(t, lift, n, from, &o) || !(o.period_spread <= max_period_spread)) {
        return false;
    }
    std::size_t first = n;
    std::size_t last = n;
    for (std::size_t k = 0; k < n; ++k) {
        if (t[k] >= from) {
            first = first == n ? k : first;
            last = k;
        }
    }
    if (first == n || !(t[last] - t[first] >= 2.0 * o.period)) {
        return false;
    }
    Tally head_drag, head_lift, tail_drag, tail_lift;
    span_summary(t, drag, lift, n, t[first], t[first] + o.period, &head_drag, &head_lift);
    span_summary(t, drag, lift, n, t[last] - o.period, t[last], &tail_drag, &tail_lift);
    if (head_drag.n == 0 || tail_drag.n == 0 || !close_to(head_drag.average(), tail_drag.average(), stationarity) ||
        !close_to(head_lift.width(), tail_lift.width(), stationarity)) {
        return false;
    }
    *out = o;
    return true;
}

} // namespace

extern "C" {

// force_history.cpp:9-13
void lbm_force_coefficients(double force_x, double force_y, double speed, double length, double* drag, double* lift) {
    const double q = 0.5 * speed * speed * length;
    *drag = force_x / q;
    *lift = force_y / q;
}

// force_history.cpp:36-47
double lbm_mean_since(const double* time, const double* values, size_t n, double from) {
    return tail_tally(time, values, n, from).average();
}

int lbm_dominant_oscillation(const double* time, const double* values, size_t n, double from, lbm_oscillation* out) {
    return oscillation_of(time, values, n, from, out) ? 1 : 0;
}

// force_history.cpp:149-157: the span runs to the largest finite time.
int lbm_steady_since(const double* time, const double* drag, const double* lift, size_t n, double from,
                     double steady_tolerance) {
    Tally d, l;
    span_summary(time, drag, lift, n, from, DBL_MAX, &d, &l);
    if (d.n < 2) {
        return 0;
    }
    const double limit = steady_tolerance * std::abs(d.average());
    return d.width() <= limit && l.width() <= limit ? 1 : 0;
}

int lbm_settled_oscillation(const double* time, const double* drag, const double* lift, size_t n, double from,
                            double max_period_spread, double stationarity, lbm_oscillation* out) {
    return settled(time, drag, lift, n, from, max_period_spread, stationarity, out) ? 1 : 0;
}

// lbm_aero_solver3d.cpp:60-83: each of drag, lift and side force is either steady (two or more
// samples whose width is within steady_tolerance x |mean drag|) or settled shedding paired with
// the drag.
int lbm_forces_settled(const double* time, const double* drag, const double* lift, const double* side, size_t n,
                       double from, double steady_tolerance, double max_period_spread, double stationarity) {
    const double limit = steady_tolerance * std::abs(tail_tally(time, drag, n, from).average());
    const double* components[3] = {drag, lift, side};
    for (const double* values : components) {
        const Tally s = tail_tally(time, values, n, from);
        if (s.n >= 2 && s.width() <= limit) {
            continue;
        }
        lbm_oscillation o{};
        if (!settled(time, drag, values, n, from, max_period_spread, stationarity, &o)) {
            return 0;
        }
    }
    return 1;
}

} // extern "C"
