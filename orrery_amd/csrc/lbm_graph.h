// Captured runs of steps for the step calls of a lattice handle. A call of `count` steps
// replays a HIP graph of G steps (G even, so the A/B parity at its start is also the parity at
// its end) count / G times and launches the rest one by one. The graph is captured again when
// anything it bakes in changes: the launch parameters (the kernel descriptor and the launch
// shape, compared byte for byte), G, or the starting parity. To avoid capturing for parameters
// that change on every call (a ramped inlet), a graph is captured only for a call long enough to
// replay it kReplays times, or when the previous call asked for the same parameters.
//
// ORRERY_LBM_GRAPHS=0 in the environment turns the replays off for every handle created after
// it is set (each step is then launched on its own), e.g. for tools that trace launches one by one.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace lbm {

class StepGraph {
public:
    StepGraph() {
        const char* env = std::getenv("ORRERY_LBM_GRAPHS");
        enabled_ = !(env != nullptr && env[0] == '0');
    }
    StepGraph(const StepGraph&) = delete;
    StepGraph& operator=(const StepGraph&) = delete;
    ~StepGraph() { reset(); }

    void reset() {
        if (exec_ != nullptr) {
            (void)hipGraphExecDestroy(exec_);
            exec_ = nullptr;
        }
        if (graph_ != nullptr) {
            (void)hipGraphDestroy(graph_);
            graph_ = nullptr;
        }
        key_.clear();
    }

    // With graphs off every step is launched on its own.
    void enable(bool on) {
        enabled_ = on;
        if (!on) {
            reset();
        }
    }
    bool enabled() const { return enabled_; }

    // Steps per graph for a call of `count` steps: the largest even number up to kMax, or 0
    // when the call is too short to gain from a graph.
    static std::uint64_t steps_for(std::uint64_t count) {
        if (count < kMin) {
            return 0;
        }
        const std::uint64_t g = count < kMax ? count : kMax;
        return g & ~std::uint64_t{1};
    }

    // Runs `count` steps: `launch(local)` enqueues one step of parity `local` on `stream`;
    // `key` holds the bytes of everything the launches bake in (with the starting parity).
    // Returns the first failing HIP status.
    template <class Launch>
    hipError_t run(hipStream_t stream, std::uint64_t count, std::uint64_t& steps, const std::vector<char>& key,
                   Launch&& launch) {
        const std::uint64_t g = steps_for(count);
        std::vector<char> full = key;
        const std::uint64_t parity = steps & 1U;
        append(full, g);
        append(full, parity);
        hipError_t e = hipSuccess;
        std::uint64_t done = 0;
        if (enabled_ && g > 0) {
            const bool ready = exec_ != nullptr && full == key_;
            if (!ready && (full == last_ || count / g >= kReplays)) {
                // G steps launched one by one first keep the device busy while the host captures
                // and instantiates (a few ms for 64 steps, all of it idle device time otherwise:
                // 128^3 D3Q19, 1000 steps after 50, 71 ms against 67 ms once captured); G is even,
                // so the parity, and with it the key, is the same after them.
                for (std::uint64_t k = 0; k < g && e == hipSuccess; ++k) {
                    e = launch((steps & 1U) != 0);
                    if (e == hipSuccess) {
                        ++steps;
                        ++done;
                    }
                }
                if (e == hipSuccess) {
                    e = capture(stream, g, parity, full, launch);
                }
            }
            if (e == hipSuccess && exec_ != nullptr && full == key_) {
                while (done + g <= count && e == hipSuccess) {
                    e = hipGraphLaunch(exec_, stream);
                    if (e == hipSuccess) {
                        steps += g;
                        done += g;
                    }
                }
            }
        }
        last_ = full;
        for (; done < count && e == hipSuccess; ++done) {
            e = launch((steps & 1U) != 0);
            if (e == hipSuccess) {
                ++steps;
            }
        }
        return e;
    }

private:
    static constexpr std::uint64_t kMin = 8;
    static constexpr std::uint64_t kMax = 64;
    static constexpr std::uint64_t kReplays = 4;

    static void append(std::vector<char>& v, std::uint64_t x) {
        const char* b = reinterpret_cast<const char*>(&x);
        v.insert(v.end(), b, b + sizeof(x));
    }

    template <class Launch>
    hipError_t capture(hipStream_t stream, std::uint64_t g, std::uint64_t parity, const std::vector<char>& full,
                       Launch& launch) {
        reset();
        hipError_t e = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) {
            return e;
        }
        hipError_t le = hipSuccess;
        for (std::uint64_t k = 0; k < g && le == hipSuccess; ++k) {
            le = launch(((parity + k) & 1U) != 0);
        }
        hipGraph_t graph = nullptr;
        e = hipStreamEndCapture(stream, &graph);
        if (le != hipSuccess || e != hipSuccess) {
            if (graph != nullptr) {
                (void)hipGraphDestroy(graph);
            }
            return le != hipSuccess ? le : e;
        }
        hipGraphExec_t exec = nullptr;
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            (void)hipGraphDestroy(graph);
            return e;
        }
        graph_ = graph;
        exec_ = exec;
        key_ = full;
        return hipSuccess;
    }

    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
    std::vector<char> key_;
    std::vector<char> last_;
    bool enabled_ = true;
};

// The bytes of a trivially copyable value, appended to a graph key.
template <class V> void key_append(std::vector<char>& key, const V& value) {
    const char* b = reinterpret_cast<const char*>(&value);
    key.insert(key.end(), b, b + sizeof(V));
}

} // namespace lbm
