// C++ host side of the contact-tracing path over the C ABI of glint.h.
//
// Mirrors the reference's stage interface so that its tick loop reads the
// same: class workbench with filter() / reachability() / update_meetings() /
// update_schema() (reference src/tracing/workbench.h:301-313, 339) and class
// tracer with process() (src/tracing/trace.h:24-54, process.cpp:378-634).
// Errors surface as glint::error exceptions carrying the glint_status, where
// the reference asserts or calls exit() (cuda_util.cuh:19-27).
#pragma once

#include <algorithm>
#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "glint.h"

namespace glint {

class error : public std::runtime_error {
public:
    error(glint_status s, const std::string &what) : std::runtime_error(what), status(s) {}
    glint_status status;
};

inline void check(glint_status s, const glint_ctx *ctx, const char *where)
{
    if (s != GLINT_OK) {
        std::string msg = std::string(where) + ": " + glint_status_string(s);
        const char *detail = glint_last_error_message(ctx);
        if (detail && *detail) msg += std::string(" (") + detail + ")";
        throw error(s, msg);
    }
}

// One device context: the reference's workbench plus its device twin
// (cuda_create_device_bench, geometry.cu:1395).
class workbench {
public:
    explicit workbench(const glint_config &cfg) : cfg_(cfg)
    {
        check(glint_create(&cfg_, &ctx_), nullptr, "glint_create");
    }
    ~workbench()
    {
        if (ctx_) glint_destroy(ctx_);
    }
    workbench(const workbench &) = delete;
    workbench &operator=(const workbench &) = delete;

    // bench->points = trace + t*N; bench->cur_time = st + t (process.cpp:434-435)
    void set_points(const double *xy, uint32_t cur_time)
    {
        check(glint_set_positions(ctx_, xy, cfg_.num_objects, cur_time), ctx_, "glint_set_positions");
    }
    void filter() { check(glint_partition(ctx_), ctx_, "glint_partition"); }          // filter.cpp:290
    void reachability() { check(glint_refine(ctx_), ctx_, "glint_refine"); }         // refinement.cpp:66
    // meeting.cpp:166; the tick's meetings, sorted by (pid_a, pid_b)
    const std::vector<glint_meeting> &update_meetings()
    {
        uint64_t n = 0;
        glint_status s = glint_evaluate(ctx_, meetings_.data(), meetings_.size(), &n);
        if (s == GLINT_ERR_OUTPUT_CAPACITY && n > meetings_.size()) {
            meetings_.resize(n);
            const glint_meeting *d = nullptr;
            check(glint_device_meetings(ctx_, &d, &n), ctx_, "glint_device_meetings");
            check(glint_memcpy_to_host(meetings_.data(), d, n * sizeof(glint_meeting)), ctx_,
                  "glint_memcpy_to_host");
            std::sort(meetings_.begin(), meetings_.begin() + n, [](const glint_meeting &a, const glint_meeting &b) {
                return a.pid_a != b.pid_a ? a.pid_a < b.pid_a : a.pid_b < b.pid_b;
            });
        } else {
            check(s, ctx_, "glint_evaluate");
        }
        last_.assign(meetings_.begin(), meetings_.begin() + n);
        return last_;
    }
    void update_schema() { check(glint_update_schema(ctx_), ctx_, "glint_update_schema"); }   // filter.cpp:428
    glint_tick_stats stats()
    {
        glint_tick_stats st{};
        check(glint_stats(ctx_, &st), ctx_, "glint_stats");
        return st;
    }
    glint_ctx *handle() { return ctx_; }
    const glint_config &config() const { return cfg_; }

private:
    glint_config cfg_;
    glint_ctx *ctx_ = nullptr;
    std::vector<glint_meeting> meetings_ = std::vector<glint_meeting>(1 << 16);
    std::vector<glint_meeting> last_;
};

// The tick loop of tracer::process (process.cpp:378-634): the first tick's
// positions build the quadtree, then every tick runs the four stages.  The
// positions come from a callback (the reference's generator or .tr loader).
class tracer {
public:
    using source = std::function<const double *(uint32_t tick)>;
    using sink = std::function<void(uint32_t tick, const std::vector<glint_meeting> &)>;

    tracer(const glint_config &cfg, source positions) : bench_(cfg), positions_(std::move(positions)) {}

    void process(uint32_t duration, const sink &on_meetings)
    {
        const uint32_t st = bench_.config().start_time;
        for (uint32_t t = 0; t < duration; t++) {
            bench_.set_points(positions_(st + t), st + t);
            bench_.filter();
            bench_.reachability();
            const auto &m = bench_.update_meetings();
            if (on_meetings) on_meetings(st + t, m);
            if (bench_.config().dynamic_schema) bench_.update_schema();
        }
    }
    workbench &bench() { return bench_; }

private:
    workbench bench_;
    source positions_;
};

}  // namespace glint
