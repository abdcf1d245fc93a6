/*
 * oracle.h - CPU restatement of GLINT's per-tick contact-tracing path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * or as the reported CPU baseline.  The product (libglint.so) never links,
 * loads or calls it.
 *
 * What it restates (reference paths relative to the reference repository):
 *   tree build      src/index/QTree.h:98-214 (QTNode ctor, split, insert)
 *   filter          src/tracing/filter.cpp:210-288 (filter_unit, phased lookup)
 *   lookup          src/tracing/filter.cpp:130-152 (lookup_rec, top/right rule)
 *   refine          src/tracing/refinement.cpp:10-64 (reachability_unit)
 *   evaluate        src/tracing/meeting.cpp:32-76 (update_meetings_unit)
 *   schema update   src/cuda/geometry.cu:1114-1234 (collect, split, merge)
 *   grid rebuild    src/cuda/geometry.cu:1270-1393,1629-1645 (build_qtree,
 *                   merge_qtree: dynamic_schema off on the GPU)
 *   predicates      src/geometry/geometry.h:40-48,140-145,174-185
 *   tick order      src/tracing/process.cpp:431-612
 * with coordinates in planar metres (geography=false predicates), or lon/lat
 * degrees with the geography=true predicates (oracle_config.geography), and the
 * deterministic semantics the reference leaves undefined: no leaf-overflow
 * drops, one table entry per unordered pair, unbounded probing, schema
 * decisions taken from the counts of the tick before any split or merge.
 */
#ifndef GLINT_ORACLE_H
#define GLINT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Tree maintenance between ticks (config.h:35 dynamic_schema):
 *   STATIC       the tree built from the first tick is kept (the reference's
 *                CPU path with dynamic_schema off, process.cpp:608)
 *   DYNAMIC      split and merge after every tick (geometry.cu:1114-1234)
 *   GRID_REBUILD the tree rebuilt from a 16384^2 cell histogram every tick
 *                (the reference's GPU path with dynamic_schema off,
 *                geometry.cu:1629-1645) */
enum { ORACLE_SCHEMA_STATIC = 0, ORACLE_SCHEMA_DYNAMIC = 1, ORACLE_SCHEMA_GRID_REBUILD = 2 };

typedef struct oracle_config {
    uint32_t num_objects;
    uint32_t grid_capacity;        /* config.h:29 */
    uint32_t min_meet_time;        /* config.h:39 */
    uint32_t schema_update_delay;  /* config.h:38 */
    uint32_t dynamic_schema;       /* config.h:35: ORACLE_SCHEMA_* */
    uint32_t complete_mode;        /* 1: every pair within r, no top/right rule */
    uint32_t num_threads;          /* 0: OpenMP default */
    uint32_t start_time;           /* first tick number, config.h:28 */
    double low_x, low_y, high_x, high_y;   /* root box */
    double reach_distance;         /* config.h:41 */
    uint32_t geography;            /* 1: lon/lat degrees with the geography=true predicates */
    uint32_t pad_;
} oracle_config;

typedef struct oracle_stats {
    uint64_t num_leaves;
    uint64_t num_nodes;
    uint64_t num_border;          /* objects with last_valid != leaf */
    uint64_t num_units;           /* check units (one per object + border units) */
    uint64_t num_tests;           /* distance tests performed */
    uint64_t num_contacts;        /* unique contact pairs this tick */
    uint64_t num_active;          /* active pairs after this tick */
    uint64_t num_meetings;        /* meetings emitted this tick */
    uint64_t max_leaf_count;      /* largest leaf population this tick */
    uint64_t ref_drop_leaves;     /* leaves over 2*grid_capacity (reference would drop) */
    uint64_t splits, merges;      /* schema changes applied after this tick */
    uint64_t ref_drop_objects;    /* objects beyond 2*grid_capacity in those leaves */
    double t_filter, t_refine, t_evaluate, t_schema;  /* seconds */
} oracle_stats;

/* Geography (geometry.h:40-48,159-185; util.h:38,57-66): degrees of
 * longitude per metre at a latitude, the reference's Point::distance with
 * geography (p1's latitude), and box::distance with geography (the point's
 * latitude). */
double oracle_degree_per_meter_longitude(double latitude);
int oracle_point_within_geo(double x1, double y1, double x2, double y2, double r);
double oracle_box_distance_geo(double lx, double ly, double hx, double hy, double px, double py);

void *oracle_create(const oracle_config *cfg);
void oracle_destroy(void *h);

/* One tick: xy is AoS (x0,y0,x1,y1,...) for all objects, tick is cur_time.
 * Builds the tree on the first call (QTree.h build rule), then filter,
 * refine, evaluate and (if dynamic) the schema update.  Returns 0 or <0. */
int oracle_tick(void *h, const double *xy, uint32_t tick);

/* Results of the last tick.  Contacts are (a<<32)|b with a<b, sorted,
 * unique.  Meetings are 4 x u32 (pid_a, pid_b, start, end), sorted. */
uint64_t oracle_num_contacts(void *h);
const uint64_t *oracle_contacts(void *h);
uint64_t oracle_num_meetings(void *h);
const uint32_t *oracle_meetings(void *h);
void oracle_get_stats(void *h, oracle_stats *out);

/* Active pairs (key, start) after the last tick, sorted by key. */
uint64_t oracle_num_active(void *h);
void oracle_get_active(void *h, uint64_t *keys, uint32_t *starts);

/* Emit every still-active pair as if it ended now (not a reference
 * behaviour: the reference never emits meetings active at the last tick). */
uint64_t oracle_flush(void *h, uint32_t *out4, uint64_t cap);

/* Leaves of the current tree, in depth-first (BL, BR, TL, TR) order:
 * boxes 4 doubles each (lx, ly, hx, hy), level, and the leaf population of
 * the last tick (0 before the first tick). */
uint64_t oracle_num_leaves(void *h);
uint64_t oracle_get_leaves(void *h, double *boxes, uint32_t *levels, uint32_t *counts,
                           uint64_t cap);

/* Build the tree from xy without running a tick (tests). */
int oracle_build(void *h, const double *xy);

/* Cell x + 16384*y of every point for the grid rebuild (geometry.cu:1276-
 * 1278); returns -2 if a point has no cell (on the high y edge). */
int oracle_grid_cells(const oracle_config *cfg, const double *xy, uint32_t *cell);

/* Descend one point: returns the depth-first index of its leaf, writes the
 * range [lo, hi) of depth-first leaf indices under last_valid and the
 * candidate leaves (depth-first leaf indices) of the lookup. */
int64_t oracle_locate(void *h, double x, double y, int64_t *lv_span_lo, int64_t *lv_span_hi,
                      uint32_t *cands, uint32_t cap, uint32_t *ncand);

/* Feed a contact set directly to the evaluate stage (tests of the meeting
 * rule): keys sorted or not, duplicates allowed. */
int oracle_evaluate_contacts(void *h, const uint64_t *keys, uint64_t n, uint32_t tick);

/* Apply one schema update from explicit per-leaf counts given in the
 * depth-first leaf order (tests of split/merge). */
int oracle_update_schema_counts(void *h, const uint32_t *counts, uint64_t nleaves);

/* ------------------------------------------------ sorted key-value output
 * The flush branch of the reference's GPU path (src/cuda/geometry.cu:665-806,
 * 1752-2138; src/tracing/step_merge.cpp:140-310): every meeting carries the
 * float box of both objects over its contact ticks; ended meetings become
 * two directed 128-bit keys; a flush sorts them, assigns STR file ids and
 * encodes per-file bitmaps and variable-width keys.  oracle_kv.c states the
 * rules line by line. */
typedef struct oracle_kv_config {
    uint32_t split_num;        /* config.h:60: split_num^2 files per flush */
    uint32_t max_meet_time;    /* config.h:40 */
    double area_limit;         /* dProbFalse, process.cpp:408-411 (squared input units) */
} oracle_kv_config;

typedef struct oracle_kv_summary {
    uint64_t kept_keys;        /* keys in files (sid >= 2 or unassigned) */
    uint64_t oversize_keys;    /* keys of sid-1 objects */
    uint64_t packed_bytes;
    uint64_t logged_meetings;  /* meetings whose box lies in the map box */
    uint32_t ctf_count;
    uint32_t sid_count;        /* objects with at least one key */
    uint32_t oversize_oid_count;
    uint32_t zero_one_count;
    uint32_t end_time_min, end_time_max;
    uint32_t o_start_time_min, o_start_time_max;
} oracle_kv_summary;

typedef struct oracle_ctf {    /* CTF (step_merge.h:64-84), one per file */
    float mbr[4];              /* low_x, low_y, high_x, high_y */
    uint64_t key_offset;       /* first key in the sorted kept keys */
    uint64_t packed_offset;    /* first byte in the packed key stream */
    uint32_t kv_count, start_time_min, start_time_max, end_time_min, end_time_max, bitmap_bytes;
    uint16_t key_bit, id_bit, duration_bit, end_bit, low_x_bit, low_y_bit, edge_bit, mbr_bit, x_grid, y_grid;
    uint32_t reserved;
} oracle_ctf;

/* Before the first tick: start tracking meeting boxes and collecting keys. */
int oracle_kv_enable(void *h, const oracle_kv_config *kc);
/* Directed keys collected since the last flush. */
uint64_t oracle_kv_pending(void *h);
/* Organise the collected keys after the last tick (cur_time = that tick). */
int oracle_kv_flush(void *h, oracle_kv_summary *out);
/* Results of the last flush: part 0 kept keys, 1 oversize keys; keys are
 * (lo, hi) u64 pairs, boxes 4 floats; sids one per object; bitmaps 8192
 * bytes per file; the oversize bitmap 8192 bytes. */
const uint64_t *oracle_kv_keys(void *h, int part);
const float *oracle_kv_boxes(void *h, int part);
const uint16_t *oracle_kv_sids(void *h);
const oracle_ctf *oracle_kv_ctfs(void *h);
const uint8_t *oracle_kv_bitmaps(void *h);
const uint8_t *oracle_kv_obitmap(void *h);
const uint8_t *oracle_kv_packed(void *h);

/* Predicates, exactly as the reference evaluates them in planar metres. */
int oracle_point_within(double x1, double y1, double x2, double y2, double r);
double oracle_box_distance(double lx, double ly, double hx, double hy, double px, double py);

#ifdef __cplusplus
}
#endif

#endif
