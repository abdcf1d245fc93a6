// Command-line driver in the manner of the reference's `pipeline` tool
// (src/tools/pipeline.cpp:7-20): generate a seeded trace and run the tick
// loop on the GPU, printing per-tick meeting counts and stage times.
//
//   glint_pipeline [objects] [ticks] [side_m] [hotspots] [grid_capacity] [min_meet] [reach_m]
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "glint_gen.h"
#include "glint_tracer.hpp"

int main(int argc, char **argv)
{
    const uint32_t n = argc > 1 ? (uint32_t)std::atol(argv[1]) : 100000;
    const uint32_t ticks = argc > 2 ? (uint32_t)std::atol(argv[2]) : 100;
    const double side = argc > 3 ? std::atof(argv[3]) : 10000.0;
    const uint32_t spots = argc > 4 ? (uint32_t)std::atol(argv[4]) : 0;
    const uint32_t cap = argc > 5 ? (uint32_t)std::atol(argv[5]) : 200;
    const uint32_t min_meet = argc > 6 ? (uint32_t)std::atol(argv[6]) : 5;
    const double reach = argc > 7 ? std::atof(argv[7]) : 2.0;

    glint_gen_config g{};
    g.seed = 1;
    g.num_objects = n;
    g.num_hotspots = spots;
    g.side = side;
    g.street_gap = 50.0;
    g.speed_min = 1.0;
    g.speed_max = 15.0;
    g.hotspot_fraction = spots ? 0.8 : 0.0;
    g.hotspot_sigma = 200.0;
    g.zipf_s = 1.1;
    g.move_fraction = 1.0;
    std::vector<double> xy(2 * (size_t)n), speed(n);
    std::vector<int32_t> tx(n), ty(n);
    std::vector<uint8_t> dir(n);
    glint_walkers w{xy.data(), speed.data(), tx.data(), ty.data(), dir.data()};
    if (glint_gen_init(&g, &w, 0, n) != 0) {
        std::fprintf(stderr, "invalid generator configuration\n");
        return 2;
    }

    glint_config cfg;
    glint_config_default(&cfg);
    cfg.num_objects = n;
    cfg.grid_capacity = cap;
    cfg.min_meet_time = min_meet;
    cfg.reach_distance = reach;
    cfg.low_x = cfg.low_y = 0.0;
    cfg.high_x = cfg.high_y = side;
    try {
        glint::tracer tr(cfg, [&](uint32_t tick) -> const double * {
            if (tick > 0) glint_gen_step(&g, &w, tick, 0, n);
            return xy.data();
        });
        uint64_t total = 0;
        auto t0 = std::chrono::steady_clock::now();
        tr.process(ticks, [&](uint32_t tick, const std::vector<glint_meeting> &m) {
            total += m.size();
            glint_tick_stats st = tr.bench().stats();
            // digest of the tick's meetings in their sorted order: h = (h ^ v) * FNV prime
            // over every field of every meeting
            uint64_t h = 1469598103934665603ull;
            for (const glint_meeting &x : m)
                for (uint32_t v : {x.pid_a, x.pid_b, x.start, x.end}) h = (h ^ v) * 1099511628211ull;
            std::printf("tick %u meetings %zu digest %016llx contacts %llu leaves %u partition %.3f refine %.3f "
                        "evaluate %.3f ms\n",
                        tick, m.size(), (unsigned long long)h, (unsigned long long)st.num_contacts, st.num_leaves,
                        st.ms_partition, st.ms_refine, st.ms_evaluate);
        });
        double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::printf("%u objects x %u ticks: %llu meetings, %.3f s (%.1f object-updates/s incl. host generator)\n", n,
                    ticks, (unsigned long long)total, s, (double)n * ticks / s);
    } catch (const glint::error &e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
