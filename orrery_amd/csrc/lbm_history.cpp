// Force-coefficient history analysis on the host: Cd/Cl, window means, the dominant (shedding)
// oscillation from hysteresis mean crossings, and the steady / settled-periodic tests.
// Restates src/aero/lbm/force_history.cpp:9-193 (same loops and expression order, FP64).
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <limits>
#include <vector>

#include "../../include/orrery_lbm.h"

namespace {

double mean_since(const double* time, const double* values, std::size_t n, double from) {
    double sum = 0.0;
    std::size_t count = 0;
    for (std::size_t k = 0; k < n; ++k) {
        if (time[k] >= from) {
            sum += values[k];
            ++count;
        }
    }
    return count == 0 ? 0.0 : sum / static_cast<double>(count);
}

// force_history.cpp:49-100
bool dominant(const double* time, const double* values, std::size_t n, double from, lbm_oscillation* out) {
    const double mean = mean_since(time, values, n, from);
    double lo = 0.0;
    double hi = 0.0;
    bool any = false;
    for (std::size_t k = 0; k < n; ++k) {
        if (time[k] >= from) {
            lo = any ? std::min(lo, values[k]) : values[k];
            hi = any ? std::max(hi, values[k]) : values[k];
            any = true;
        }
    }
    const double band = 0.25 * (hi - lo);
    std::vector<double> crossings;
    bool armed = false;
    for (std::size_t k = 1; k < n; ++k) {
        if (time[k - 1] < from) {
            continue;
        }
        const double a = values[k - 1] - mean;
        const double b = values[k] - mean;
        if (a < -band) {
            armed = true;
        }
        if (armed && a < 0.0 && b >= 0.0) {
            const double fraction = a / (a - b);
            crossings.push_back(time[k - 1] + fraction * (time[k] - time[k - 1]));
            armed = false;
        }
    }
    if (crossings.size() < 3) {
        return false;
    }
    double shortest = crossings[1] - crossings[0];
    double longest = shortest;
    for (std::size_t k = 2; k < crossings.size(); ++k) {
        const double p = crossings[k] - crossings[k - 1];
        shortest = std::min(shortest, p);
        longest = std::max(longest, p);
    }
    out->cycles = crossings.size() - 1;
    out->period = (crossings.back() - crossings.front()) / static_cast<double>(out->cycles);
    out->frequency = 1.0 / out->period;
    out->amplitude = 0.5 * (hi - lo);
    out->period_spread = (longest - shortest) / out->period;
    return true;
}

struct SpanStats {
    std::size_t count = 0;
    double drag_mean = 0.0;
    double lift_range = 0.0;
    double drag_range = 0.0;
};

// force_history.cpp:113-140
SpanStats span_stats(const double* time, const double* drag, const double* lift, std::size_t n, double from,
                     double to) {
    SpanStats stats;
    double sum = 0.0;
    double lift_lo = 0.0;
    double lift_hi = 0.0;
    double drag_lo = 0.0;
    double drag_hi = 0.0;
    for (std::size_t k = 0; k < n; ++k) {
        if (time[k] < from || time[k] > to) {
            continue;
        }
        const bool first = stats.count == 0;
        lift_lo = first ? lift[k] : std::min(lift_lo, lift[k]);
        lift_hi = first ? lift[k] : std::max(lift_hi, lift[k]);
        drag_lo = first ? drag[k] : std::min(drag_lo, drag[k]);
        drag_hi = first ? drag[k] : std::max(drag_hi, drag[k]);
        sum += drag[k];
        ++stats.count;
    }
    if (stats.count > 0) {
        stats.drag_mean = sum / static_cast<double>(stats.count);
        stats.lift_range = lift_hi - lift_lo;
        stats.drag_range = drag_hi - drag_lo;
    }
    return stats;
}

bool agree(double a, double b, double tolerance) {
    return std::abs(a - b) <= tolerance * std::max(std::abs(a), std::abs(b));
}

} // namespace

extern "C" {

void lbm_force_coefficients(double force_x, double force_y, double speed, double length, double* drag, double* lift) {
    // force_history.cpp:9-13
    const double dynamic = 0.5 * speed * speed * length;
    *drag = force_x / dynamic;
    *lift = force_y / dynamic;
}

double lbm_mean_since(const double* time, const double* values, size_t n, double from) {
    return mean_since(time, values, n, from);
}

int lbm_dominant_oscillation(const double* time, const double* values, size_t n, double from, lbm_oscillation* out) {
    return dominant(time, values, n, from, out) ? 1 : 0;
}

// force_history.cpp:149-157
int lbm_steady_since(const double* time, const double* drag, const double* lift, size_t n, double from,
                     double steady_tolerance) {
    const SpanStats stats = span_stats(time, drag, lift, n, from, std::numeric_limits<double>::max());
    if (stats.count < 2) {
        return 0;
    }
    const double bound = steady_tolerance * std::abs(stats.drag_mean);
    return stats.drag_range <= bound && stats.lift_range <= bound ? 1 : 0;
}

// force_history.cpp:159-193
int lbm_settled_oscillation(const double* time, const double* drag, const double* lift, size_t n, double from,
                            double max_period_spread, double stationarity, lbm_oscillation* out) {
    lbm_oscillation shedding{};
    if (!dominant(time, lift, n, from, &shedding) || !(shedding.period_spread <= max_period_spread)) {
        return 0;
    }
    double start = 0.0;
    double end = 0.0;
    bool any = false;
    for (std::size_t k = 0; k < n; ++k) {
        if (time[k] >= from) {
            start = any ? start : time[k];
            end = time[k];
            any = true;
        }
    }
    const double period = shedding.period;
    if (!any || !(end - start >= 2.0 * period)) {
        return 0;
    }
    const SpanStats head = span_stats(time, drag, lift, n, start, start + period);
    const SpanStats tail = span_stats(time, drag, lift, n, end - period, end);
    if (head.count == 0 || tail.count == 0 || !agree(head.drag_mean, tail.drag_mean, stationarity) ||
        !agree(head.lift_range, tail.lift_range, stationarity)) {
        return 0;
    }
    *out = shedding;
    return 1;
}

} // extern "C"
