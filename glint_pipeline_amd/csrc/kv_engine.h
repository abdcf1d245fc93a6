// kv_engine.h - sorted key-value meeting output: the device-side engine that
// turns the meetings logged by the table update into sorted directed keys,
// STR file ids, per-file limits, bitmaps and variable-width keys.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "glint.h"
#include "glint_device.h"

struct KvEngine;

struct KvParams {
    uint32_t num_objects;
    uint32_t split_num;            // split_num^2 files per flush
    uint32_t max_meet_time;        // longer meetings mark both objects oversize
    uint32_t first_end_time_min;   // end offsets of the first run count from here
    double area_limit;             // object box area above which the object is oversize
    double map[4];                 // root box: low_x, low_y, high_x, high_y
    uint64_t key_capacity;         // directed keys per flush (two per logged meeting)
};

glint_status kv_create(const KvParams &p, hipStream_t stream, KvEngine **out, std::string &err);
void kv_destroy(KvEngine *e);
glint::KvLogEntry *kv_log(KvEngine *e);
uint64_t kv_log_capacity(const KvEngine *e);
// Organises the first nlog logged meetings after the tick cur_time; the
// next run's end offsets count from cur_time.
glint_status kv_flush(KvEngine *e, uint64_t nlog, uint32_t cur_time, glint_kv_summary *out, std::string &err);
uint64_t kv_part_bytes(const KvEngine *e, uint32_t part);
// The tick end offsets of the next flush count from (snapshot and restore).
uint32_t kv_end_time_min(const KvEngine *e);
void kv_set_end_time_min(KvEngine *e, uint32_t t);
glint_status kv_copy(KvEngine *e, uint32_t part, void *dst, uint64_t bytes, std::string &err);
