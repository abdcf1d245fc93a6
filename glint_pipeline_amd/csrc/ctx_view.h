// ctx_view.h - what the serving layer reads from a context: its stream, the
// table generation holding the pairs active after the last evaluated tick,
// and the meeting log ring.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "glint.h"

struct CtxView {
    hipStream_t stream;
    int device;
    uint32_t num_objects;
    uint32_t tick;              // last evaluated tick
    uint32_t min_meet;
    bool evaluated;
    const unsigned long long *keys;   // the tick's pairs (bit 63: a tag, masked off by readers)
    const uint32_t *starts;           // first tick of each pair's run (~0: a second copy, skipped)
    const uint32_t *count;            // device word holding the number of entries
    uint64_t cap;                     // entries the list can hold
    const glint_meeting *mlog;  // meeting log ring (null when off)
    uint64_t mlog_cap;
};

extern "C" glint_status glint_ctx_view(glint_ctx *c, CtxView *v);
