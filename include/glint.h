/*
 * glint.h - C ABI of the MI355X (gfx950) contact-tracing path:
 * partition -> refine -> evaluate -> schema update, one tick at a time.
 *
 * The reference has no plugin or FFI API for this path.  Its seam is two
 * C++-linkage functions declared at src/tracing/process.cpp:373-376 and the
 * CPU stage methods of class workbench (src/tracing/workbench.h:301-313):
 *
 *   workbench *cuda_create_device_bench(workbench*, gpu_info*)   geometry.cu:1395
 *   void process_with_gpu(workbench*, workbench* d_bench, gpu_info*)  geometry.cu:1573
 *   void workbench::filter()          filter.cpp:290   (partition + lookup)
 *   void workbench::reachability()    refinement.cpp:66 (refine)
 *   void workbench::update_meetings() meeting.cpp:166  (evaluate)
 *   void workbench::update_schema()   filter.cpp:428   (schema update)
 *
 * Each entry point below names the reference function it replaces.  All
 * functions return a glint_status instead of exit()/assert (the reference's
 * CUDA_SAFE_CALL exits, cuda_util.cuh:19-27).  Pointers are plain host or
 * device pointers; no C++ or framework types cross the boundary.
 *
 * Threading: one context owns one HIP stream; calls on one context must not
 * overlap; different contexts may be used from different threads.
 *
 * Coordinates are planar metres (FP64), one (x, y) pair per object id, the
 * layout of the reference's Point array (geometry.h:23-56, 16 bytes each).
 */
#ifndef GLINT_H
#define GLINT_H

#include <stddef.h>
#include <stdint.h>

#include "glint_gen.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum glint_status {
    GLINT_OK = 0,
    GLINT_ERR_INVALID_ARGUMENT = 1,
    GLINT_ERR_OUT_OF_MEMORY = 2,
    GLINT_ERR_HIP = 3,             /* HIP runtime error (no device, launch failure) */
    GLINT_ERR_BAD_INPUT = 4,       /* NaN or a position outside the root box */
    GLINT_ERR_TABLE_FULL = 5,      /* meeting table probe sequence exhausted */
    GLINT_ERR_NODE_CAPACITY = 6,   /* quadtree node pool exhausted */
    GLINT_ERR_UNIT_CAPACITY = 7,   /* border check-unit buffer overflow */
    GLINT_ERR_OUTPUT_CAPACITY = 8, /* meeting or contact buffer overflow */
    GLINT_ERR_STATE = 9,           /* call out of order (e.g. refine before partition) */
    GLINT_ERR_INJECTED = 10        /* failure forced by glint_debug_inject_failure */
} glint_status;

enum {
    GLINT_CANDIDATES_REFERENCE = 0, /* the reference's top/right lookup rule (filter.cpp:139-141) */
    GLINT_CANDIDATES_COMPLETE = 1   /* every pair within reach; a superset, for diagnostics */
};

enum {
    GLINT_PRECISION_FP64 = 0,       /* FP64 distance test for every candidate pair */
    GLINT_PRECISION_FP32_GUARD = 1  /* FP32 test with a guard band, FP64 recheck inside it */
};

/* glint_config.dynamic_schema (config.h:35): how the tree follows the objects.
 * A binding that replaces the reference's GPU path maps its flag to
 * DYNAMIC when set and GRID_REBUILD when not (geometry.cu:1629,2160). */
enum {
    GLINT_SCHEMA_STATIC = 0,        /* the tree built at the first tick is kept (the reference's
                                       CPU path with dynamic_schema off, process.cpp:608) */
    GLINT_SCHEMA_DYNAMIC = 1,       /* leaves split and merge after every tick (geometry.cu:1114-1234) */
    GLINT_SCHEMA_GRID_REBUILD = 2   /* the tree rebuilt every tick from a 16384^2 cell grid
                                       (geometry.cu:1270-1393,1629-1645); positions on the high
                                       y edge of the box have no cell there and are refused */
};

/* Mirrors the hot-path fields of the reference's configuration
 * (src/util/config.h:28-48, defaults at config.h:173-187). */
typedef struct glint_config {
    uint32_t num_objects;          /* config.h:21; below 2^30 */
    uint32_t grid_capacity;        /* config.h:29: leaf split threshold */
    uint32_t min_meet_time;        /* config.h:39 */
    uint32_t max_meet_time;        /* config.h:40: longer meetings are flagged oversize */
    uint32_t schema_update_delay;  /* config.h:38 */
    uint32_t dynamic_schema;       /* config.h:35: GLINT_SCHEMA_* */
    uint32_t start_time;           /* config.h:28 */
    uint32_t candidate_mode;       /* GLINT_CANDIDATES_* */
    uint32_t precision;            /* GLINT_PRECISION_* */
    uint32_t device;               /* HIP device ordinal */
    double low_x, low_y, high_x, high_y;  /* root box (workbench::mbr) */
    double reach_distance;         /* config.h:41, metres */
    double guard_epsilon;          /* FP32 guard band half-width, metres */
    uint64_t num_meeting_buckets;  /* slots per table generation; 0: automatic (power of 2) */
    uint64_t unit_capacity;        /* border check units per tick; 0: automatic */
    uint64_t meeting_capacity;     /* meetings per tick; 0: automatic */
    uint64_t contact_capacity;     /* contacts the join may store per tick (all of them with
                                      record_contacts, geography or min_meet_time 0, else those
                                      the meeting table needs); 0: 16*N + 2^20 */
    uint32_t node_capacity;        /* quadtree nodes; 0: automatic */
    uint32_t record_contacts;      /* keep each tick's contact pairs for glint_contacts() */
    uint32_t graph_ticks;          /* glint_run_ticks: ticks per captured hipGraph (0: plain launches) */
    uint32_t kv_split_num;         /* config.h:60: sorted key-value output in kv_split_num^2 files */
    uint64_t kv_capacity;          /* directed keys held between flushes (2 per meeting); 0: output off */
    double kv_area_limit;          /* dProbFalse (process.cpp:408-411): object box area, squared input units */
    uint64_t meeting_log_capacity; /* meetings kept in the device meeting log ring; 0: log off */
    uint32_t join_tile;            /* points per LDS tile of the distance join: 256 or 512 (0: 256) */
    uint32_t table_threads;        /* workgroup size of the table update: 256, 512 or 1024 (0: 512) */
    uint32_t geography;            /* 1: coordinates are lon/lat degrees and the predicates are the
                                      reference's geography=true forms (geometry.h:40-48,159-185,
                                      filter.cpp:30-31): the pair distance scales dx at the latitude
                                      of the unit's object, box distances at the point's latitude,
                                      splits use widths and heights in metres.  FP64 only, not with
                                      GLINT_SCHEMA_GRID_REBUILD; every contact is kept */
    uint32_t reserved;
} glint_config;

/* One emitted meeting: pid_a < pid_b, first and last tick of contact. */
typedef struct glint_meeting {
    uint32_t pid_a, pid_b, start, end;
} glint_meeting;

typedef struct glint_tick_stats {
    uint32_t tick;
    uint32_t num_leaves;           /* leaves after the partition of this tick */
    uint32_t num_nodes;            /* allocated quadtree nodes */
    uint32_t num_border;           /* objects with last_valid != leaf */
    uint32_t max_leaf_count;
    uint32_t splits, merges;       /* applied by the last schema update */
    uint32_t error_flags;          /* device-side error bits (0 when healthy) */
    uint64_t num_units;            /* border check units */
    uint64_t num_contacts;         /* contact pairs found (with duplicates) */
    uint64_t num_active;           /* distinct active pairs after this tick */
    uint64_t num_meetings;         /* meetings emitted this tick */
    uint64_t num_rechecks;         /* FP64 rechecks (FP32 guard mode) */
    uint64_t table_slots;
    float ms_partition, ms_lookup, ms_refine, ms_evaluate, ms_schema;  /* HIP event times */
    uint64_t num_stored;           /* contacts kept in the meeting table's generation */
    uint64_t ref_drop_objects;     /* objects beyond 2*grid_capacity in their leaf: the reference
                                      stores at most that many per leaf and drops the rest in
                                      thread order (workbench.cpp:36, filter.cpp:100-104) */
    uint32_t ref_drop_leaves;      /* leaves holding more than 2*grid_capacity objects */
    uint32_t reserved;
    uint64_t num_tests;            /* candidate pairs tested: every pair of a leaf once and every
                                      border unit against its leaf's objects (refinement.cpp:28-33) */
    uint64_t num_evaluated;        /* pair slots the join computed a distance for: 64 per column of a
                                      64-row task, after skipping column groups whose bounding box is
                                      out of reach of the rows' box (those pairs are misses); a work
                                      count that varies with the order objects land in the join's
                                      tiles, unlike every result */
} glint_tick_stats;

typedef struct glint_ctx glint_ctx;

const char *glint_status_string(glint_status s);
/* Detail of the last failure on this context (or the last create failure). */
const char *glint_last_error_message(const glint_ctx *ctx);

/* Squared-distance threshold of the device predicates: the largest double
 * s with sqrt(s) <= r under IEEE round-to-nearest, so that for every s,
 * s <= glint_reach_threshold(r) exactly when sqrt(s) <= r (the reference's
 * test, geometry.h:47 with refinement.cpp:33).  Host-only. */
double glint_reach_threshold(double r);

/* Degrees per metre of longitude at each of n latitudes (device memory in
 * and out), computed on `device` by the same function the geography
 * predicates use: util.h:57-66's expression, PI = 3.14159265, 90 degrees
 * taken as 89.  The predicates are exact only where these equal the host's
 * libm results bit for bit, which this lets a caller check on its own data. */
glint_status glint_longitude_scale_device(const double *d_latitudes, double *d_scale, uint64_t n,
                                          uint32_t device);

/* Defaults of the reference (config.h:29-41): capacity 50, min 5, r = 2 m,
 * delay 1, dynamic schema on; the caller sets num_objects and the box. */
void glint_config_default(glint_config *cfg);

/* Replaces cuda_create_device_bench (geometry.cu:1395-1565): allocates the
 * device state (quadtree pool, sorted object arrays, two meeting-table
 * generations, output buffers) on cfg->device. */
glint_status glint_create(const glint_config *cfg, glint_ctx **out);
glint_status glint_destroy(glint_ctx *ctx);

/* Positions of every object for tick `tick` (the reference's
 * bench->points = trace + t*N and bench->cur_time = st + t,
 * process.cpp:434-435).  The host variant copies N*16 bytes to the device
 * (the per-tick H2D copy of geometry.cu:1625); the device variant uses the
 * caller's device buffer in place, which must stay valid until the tick's
 * glint_update_schema has returned. */
glint_status glint_set_positions(glint_ctx *ctx, const double *xy, uint32_t n, uint32_t tick);
glint_status glint_set_positions_device(glint_ctx *ctx, const double *d_xy, uint32_t n, uint32_t tick);

/* Replaces the partition and filter half of workbench::filter()
 * (filter.cpp:210-313) and cuda_partition + cuda_filtering
 * (geometry.cu:221-392): on the first call builds the quadtree from these
 * positions (QTree.h:157-214, partitioner::build_schema filter.cpp:27-93);
 * then assigns every object to its leaf, sorts objects by leaf and generates
 * the border check units with the reference's lookup rule. */
glint_status glint_partition(glint_ctx *ctx);

/* Replaces workbench::reachability() (refinement.cpp:66-85) and
 * cuda_unroll + cuda_refinement (geometry.cu:401-497): the distance join of
 * every check unit and the meeting-table upsert of every contact. */
glint_status glint_refine(glint_ctx *ctx);

/* Replaces workbench::update_meetings() (meeting.cpp:166-189) and
 * cuda_identify_meetings (geometry.cu:595-737): emits the meetings that
 * ended before this tick and lasted long enough, and retires every pair not
 * seen this tick.  If `out` is not NULL the tick's meetings are copied to it,
 * sorted by (pid_a, pid_b); *n_out receives the count (also when it exceeds
 * cap, in which case GLINT_ERR_OUTPUT_CAPACITY is returned). */
glint_status glint_evaluate(glint_ctx *ctx, glint_meeting *out, uint64_t cap, uint64_t *n_out);

/* Replaces workbench::update_schema() (filter.cpp:428-449) and
 * cuda_update_schema_collect/split/merge (geometry.cu:1114-1234,2160-2182):
 * one level of split/merge decided from this tick's leaf counts. */
glint_status glint_update_schema(glint_ctx *ctx);

/* The four stages above for one tick (process_with_gpu's whole body). */
glint_status glint_process_tick(glint_ctx *ctx, const double *xy, uint32_t tick);

/* Runs `nticks` consecutive ticks from a device-resident trace laid out
 * tick-major (tick k at d_trace + 2*N*k), with no host synchronisation
 * between ticks; meetings stay on the device.  With cfg.graph_ticks = G the
 * ticks run as replays of one captured hipGraph of G ticks.  If tick_ms is
 * not NULL it receives each tick's latency in milliseconds from the device
 * wall clock (from the tick's first launch to the end of its last).  The
 * reference drives the same loop from the host, one tick
 * per call (process.cpp:431). */
glint_status glint_run_ticks(glint_ctx *ctx, const double *d_trace, uint32_t first_tick,
                             uint32_t nticks, float *tick_ms);

/* glint_run_ticks for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * allocation): it has no counterpart in the reference.  The same ticks and
 * results as glint_run_ticks, with plain launches (no graph is captured
 * or instantiated), and its per-tick stamps kept in a fixed device array
 * of the library, so at most 4096 ticks per call (GLINT_ERR_INVALID_ARGUMENT
 * beyond).  That array is shared by every call on one GPU: calls that
 * overlap on one GPU race.  GLINT_ERR_STATE while kernel profiling is on. */
glint_status glint_run_ticks_noalloc(glint_ctx *ctx, const double *d_trace, uint32_t first_tick,
                                     uint32_t nticks, float *tick_ms);

/* Waits for the context's stream and reports device-side errors. */
glint_status glint_synchronize(glint_ctx *ctx);

/* Statistics of the last completed tick (synchronises). */
glint_status glint_stats(glint_ctx *ctx, glint_tick_stats *out);

/* Contact pairs of the last tick ((a<<32)|b, a<b, sorted, unique); requires
 * cfg.record_contacts. */
glint_status glint_contacts(glint_ctx *ctx, uint64_t *out, uint64_t cap, uint64_t *n_out);

/* Leaves of the current quadtree in depth-first (BL, BR, TL, TR) order:
 * boxes as (lx, ly, hx, hy), depth, and the population of the last tick. */
glint_status glint_leaves(glint_ctx *ctx, double *boxes, uint32_t *levels, uint32_t *counts,
                          uint64_t cap, uint64_t *n_out);

/* Active pairs after the last tick: keys (a<<32)|b and start ticks, sorted.
 * With record_contacts or min_meet_time 0 every active pair is listed;
 * otherwise a pair first seen at the last tick is not kept (it can only
 * continue, when it is found again, or end unemitted) and only the pairs
 * whose run began earlier are listed.  glint_tick_stats.num_active counts
 * all of them. */
glint_status glint_active_pairs(glint_ctx *ctx, uint64_t *keys, uint32_t *starts, uint64_t cap,
                                uint64_t *n_out);

/* Running totals since creation or the last reset (consumed by benchmarks
 * to price each kernel's algorithmic bytes). */
typedef struct glint_totals {
    uint64_t ticks;
    uint64_t contacts;            /* contact pairs found, duplicates included */
    uint64_t contacts_border;     /* of which from border check units */
    uint64_t distinct_pairs;      /* distinct active pairs, summed over ticks */
    uint64_t meetings;
    uint64_t units;               /* border check units */
    uint64_t border_objects;
    uint64_t rechecks;            /* FP64 rechecks in FP32 guard mode */
    uint64_t max_contacts_tick;
    uint64_t max_distinct_tick;
    uint64_t max_meetings_tick;
    uint64_t stored;              /* contacts kept in the meeting table, summed over ticks */
    uint64_t max_stored_tick;
    uint64_t ref_drop_leaves;     /* leaves over 2*grid_capacity objects, summed over ticks */
    uint64_t ref_drop_objects;    /* objects beyond 2*grid_capacity in those leaves, summed */
    uint64_t max_ref_drop_objects_tick;
    uint64_t max_leaf_count;      /* largest leaf population of any tick */
    uint64_t tests;               /* candidate pairs tested, summed over ticks */
    uint64_t evaluated;           /* pair slots computed, summed over ticks */
} glint_totals;
glint_status glint_totals_get(glint_ctx *ctx, glint_totals *out, int reset);

/* Per-kernel HIP-event timing on the context's stream.  glint_kernel_times
 * reports, for each kernel of the path, the summed milliseconds and launch
 * count since the last reset (names are static strings). */
glint_status glint_set_kernel_profiling(glint_ctx *ctx, int enable);
glint_status glint_kernel_times(glint_ctx *ctx, const char **names, double *total_ms, uint64_t *launches,
                                uint32_t cap, uint32_t *n_out, int reset);

/* Device pointer to the last tick's meetings (unsorted) and their count;
 * valid until the next glint_evaluate. */
glint_status glint_device_meetings(glint_ctx *ctx, const glint_meeting **d_out, uint64_t *n_out);

/* Forces the next call of the named stage ("partition", "refine",
 * "evaluate", "update_schema", "alloc") to fail with GLINT_ERR_INJECTED, for
 * failure-path tests. */
glint_status glint_debug_inject_failure(glint_ctx *ctx, const char *stage);

/* ----------------------------------------------------------------- batches
 * Many independent scenarios (cities of uneven size, each with its own
 * configuration) ticked together on one device: each scenario is a context
 * of its own, with its own state, so its results are those of the same
 * context run alone.  Every step of the tick sequence is one launch over
 * all scenarios: each scenario gets the workgroups its own launch would
 * have, and every workgroup finds its scenario from its index.  Scenarios
 * that differ in a setting that picks a kernel variant (join tile, table
 * workgroup, key-value output, geography, schema mode) go to separate
 * groups, one launch per step and group; a group holds at most 64.  With
 * graph_ticks > 0, graph_ticks ticks of every group are captured once in
 * one hipGraph and replayed.  The reference runs one scenario per process
 * (process.cpp:378-634). */
typedef struct glint_batch glint_batch;

glint_status glint_batch_create(const glint_config *cfgs, uint32_t n, uint32_t graph_ticks, glint_batch **out);
glint_status glint_batch_destroy(glint_batch *batch);
/* The context of scenario i (owned by the batch): every per-context query
 * (glint_stats, glint_active_pairs, glint_meeting_log_read, ...) works on it. */
glint_status glint_batch_context(glint_batch *batch, uint32_t i, glint_ctx **out);
const char *glint_batch_last_error(const glint_batch *batch);
/* Number of groups of scenarios sharing fused launches. */
glint_status glint_batch_groups(const glint_batch *batch, uint32_t *groups);
/* nticks ticks of every scenario: scenario i reads tick k at
 * d_traces[i] + 2*N_i*k (device memory, FP64 AoS).  tick_ms (optional,
 * n*nticks floats): tick_ms[i*nticks + k] is the device wall time of
 * scenario i's tick k inside the batch.  GLINT_ERR_STATE if kernel
 * profiling is on for a scenario (its events would serialise the batch). */
glint_status glint_batch_run_ticks(glint_batch *batch, const double *const *d_traces, uint32_t first_tick,
                                   uint32_t nticks, float *tick_ms);

/* ------------------------------------------------------------------ serving
 * Location updates ingested per tick into device-resident state, and tracing
 * queries answered from the meeting log and the active table while ticks
 * continue (the reference's interactive searches: search_single over active
 * meetings, geometry.cu:605-662, whose targets seed the next search,
 * process.cpp:443-495; id_search_in_disk over flushed meetings by time
 * window, workbench.cpp:397-444). */

/* Scatter k location updates (pid, x, y; device memory) into the resident
 * positions and start tick `tick` on them (like glint_set_positions); the
 * first tick of a context needs every position. */
glint_status glint_ingest_updates(glint_ctx *ctx, const uint32_t *d_pids, const double *d_xy, uint32_t k,
                                  uint32_t tick);

typedef struct glint_server glint_server;

typedef struct glint_server_config {
    uint64_t arena_bytes;      /* device memory shared by admitted queries (the HBM admission budget) */
    uint32_t max_running;      /* queries on the device at once (one stream each) */
    uint32_t estimate_percent; /* scale of the hit-buffer estimates (0: 100); an overflow re-queues */
    uint64_t min_hits;         /* smallest hit buffer per hop (entries; 0: 4096) */
    uint32_t index_block_ticks;/* ticks of meetings per block of the per-object meeting index (0: 64;
                                  0xFFFFFFFF: no index, queries scan the log) */
    uint32_t reserved;
} glint_server_config;

/* One contact found by a query: source (a seed at hop 1, a hop-1 contact at
 * hop 2), target, first and last tick of the meeting. */
typedef struct glint_hit {
    uint32_t source, target, start, end;
} glint_hit;

enum { GLINT_QUERY_PENDING = 0, GLINT_QUERY_RUNNING = 1, GLINT_QUERY_DONE = 2, GLINT_QUERY_FAILED = 3 };

typedef struct glint_query_info {
    uint64_t id;
    uint32_t state;            /* GLINT_QUERY_* */
    uint32_t hops;
    uint64_t hits1, hits2;
    uint64_t footprint_bytes;
    uint32_t submit_tick;      /* the state the answer describes */
    uint32_t admitted_tick;
    uint32_t retries;          /* re-admissions after a hit buffer overflow */
    uint32_t reserved;
    double submit_ms, admit_ms, done_ms;   /* host clock since the server was created */
} glint_query_info;

/* The context must have a meeting log (meeting_log_capacity > 0) sized for the
 * longest window with an eighth of the ring to spare. */
glint_status glint_server_create(glint_ctx *ctx, const glint_server_config *cfg, glint_server **out);
glint_status glint_server_destroy(glint_server *server);
const char *glint_server_last_error(const glint_server *server);
/* One tick: every position (d_full_xy, device) or k location updates, then
 * partition, refine, evaluate and schema update; then pending queries are
 * admitted on the new state.  Running queries advance (hop 2 starts, hits
 * are copied to the host) in glint_server_poll, so that a tick never waits
 * for a query's results. */
glint_status glint_server_tick(glint_server *server, const double *d_full_xy, const uint32_t *d_pids,
                               const double *d_xy, uint32_t k, uint32_t tick);
/* Queue a query: contacts of the seeds (host array) whose meetings overlap
 * [t_start, t_end], with hops 1 or 2, answered on the state after the last
 * tick.  It runs when a slot and arena memory are free. */
glint_status glint_server_submit(glint_server *server, const uint32_t *seeds, uint32_t nseeds, uint32_t t_start,
                                 uint32_t t_end, uint32_t hops, uint64_t *id);
/* The per-object meeting index (the role of the reference's id search over
 * meetings sorted by object, workbench.cpp:397-444, step_merge.cpp:501-568):
 * blocks of directed meetings sorted by (object, target, start), sealed from
 * the meeting log every index_block_ticks ticks; queries binary-search them
 * for each seed and scan only the log since the last block.  Reports the
 * blocks held, their entries (two per meeting) and the log entries below
 * which every meeting is in a block. */
glint_status glint_server_index_info(glint_server *server, uint64_t *blocks, uint64_t *entries, uint64_t *sealed);
/* Advance finished work and admit pending queries; counts are optional. */
glint_status glint_server_poll(glint_server *server, uint32_t *running, uint32_t *pending);
/* Poll until no query is running or pending. */
glint_status glint_server_drain(glint_server *server);
glint_status glint_server_query_info(glint_server *server, uint64_t id, glint_query_info *out);
/* Hits of a finished query (hop 1 then hop 2, each sorted by source, target,
 * start); releases it.  With out = NULL only *n_out is set. */
glint_status glint_server_result(glint_server *server, uint64_t id, glint_hit *out, uint64_t cap,
                                 uint64_t *n_out);

/* ------------------------------------------------------------ meeting log
 * With meeting_log_capacity > 0 every emitted meeting is appended to a
 * device-resident ring (the role of the reference's per-tick meeting buffer
 * h_meetings_ps, meeting.cpp:20-22, kept across ticks); the newest
 * meeting_log_capacity meetings are retained.  Entries are in emission order
 * (ticks in order; within a tick unordered). */
glint_status glint_meeting_log_info(glint_ctx *ctx, uint64_t *total, uint64_t *first_retained);
/* Copy entries [from, total) (at most cap) to host memory; from must be retained. */
glint_status glint_meeting_log_read(glint_ctx *ctx, uint64_t from, glint_meeting *out, uint64_t cap,
                                    uint64_t *n_out);

/* --------------------------------------------------------------- snapshots
 * The reference dumps its meeting buckets and metadata and loads them back
 * into the table (dump_meetings / load_meetings / dump_meta,
 * workbench.cpp:446-470,574-595; insert_meetings_to_buckets,
 * geometry.cu:543-565).  Here a snapshot is everything the next tick reads,
 * taken at a tick boundary (after an evaluated tick): the quadtree, the last
 * tick's objects in leaf order with their candidate lists, the position
 * histories, the last generation of the meeting table, the counters, the
 * meeting log and the key-value log.  A context created with the same
 * configuration that restores it continues the run exactly; its resident
 * positions (for glint_ingest_updates) are the last tick's. */
glint_status glint_snapshot_size(glint_ctx *ctx, uint64_t *bytes);
/* Copies the snapshot into host memory (bytes >= glint_snapshot_size). */
glint_status glint_snapshot(glint_ctx *ctx, void *buf, uint64_t bytes);
/* GLINT_ERR_INVALID_ARGUMENT if the snapshot belongs to another configuration. */
glint_status glint_restore(glint_ctx *ctx, const void *buf, uint64_t bytes);

/* ---------------------------------------------------- sorted key-value output
 * The flush branch of the reference's GPU path (geometry.cu:1752-2138).  With
 * kv_capacity > 0 every active pair carries the float box of both objects
 * over its run (geometry.cu:469-494), and every ended meeting whose box lies
 * in the root box is logged (geometry.cu:665-735).  A flush turns the log into
 * two directed keys per meeting
 *     [sid:12 | oid:26 | target:26 | duration:32 | end - end_time_min:32]
 * (util.h:47-52, geometry.cu:724-726), marks objects oversize (sid 1) by
 * meeting duration and object box area, assigns the other objects STR file
 * ids 2..split^2+1 (geometry.cu:740-796,1832-1895), sorts the keys of each
 * part, and computes per-file limits, widths (step_merge.cpp:145-186),
 * bitmaps and variable-width keys (geometry.cu:820-1107). */
typedef struct glint_kv_summary {
    uint64_t kept_keys;            /* keys in the files (h_keys) */
    uint64_t oversize_keys;        /* keys of sid-1 objects (oversize buffer) */
    uint64_t packed_bytes;         /* bytes of the variable-width key stream */
    uint64_t logged_meetings;      /* meetings organised by this flush */
    uint32_t ctf_count;            /* files: kv_split_num^2 */
    uint32_t sid_count;            /* objects with at least one key */
    uint32_t oversize_oid_count;   /* the reference's count: oversize meetings + oversize objects */
    uint32_t zero_one_count;       /* STR position of the first object given a file */
    uint32_t end_time_min, end_time_max;
    uint32_t o_start_time_min, o_start_time_max;   /* oversize buffer start times */
} glint_kv_summary;

typedef struct glint_ctf {         /* CTF (step_merge.h:64-84): one file */
    float mbr[4];                  /* low_x, low_y, high_x, high_y of its keys' boxes */
    uint64_t key_offset;           /* first key in the sorted kept keys */
    uint64_t packed_offset;        /* first byte in the packed key stream */
    uint32_t kv_count, start_time_min, start_time_max, end_time_min, end_time_max, bitmap_bytes;
    uint16_t key_bit, id_bit, duration_bit, end_bit, low_x_bit, low_y_bit, edge_bit, mbr_bit, x_grid, y_grid;
    uint32_t reserved;
} glint_ctf;

typedef enum glint_kv_part {
    GLINT_KV_KEYS = 0,             /* kept keys, sorted: 16 B each, little-endian 128-bit */
    GLINT_KV_BOXES = 1,            /* their boxes: 4 floats each */
    GLINT_KV_OVERSIZE_KEYS = 2,    /* oversize keys, sorted */
    GLINT_KV_OVERSIZE_BOXES = 3,
    GLINT_KV_SIDS = 4,             /* uint16 file id per object (0: no key, 1: oversize) */
    GLINT_KV_CTFS = 5,             /* glint_ctf per file */
    GLINT_KV_BITMAPS = 6,          /* 8192 bytes per file: bit x + y*x_grid */
    GLINT_KV_OVERSIZE_BITMAP = 7,  /* 8192 bytes: 256x256 cells of the root box */
    GLINT_KV_PACKED = 8            /* files' keys at key_bit/8 bytes each, little-endian */
} glint_kv_part;

/* Directed keys waiting for the next flush (2 per logged meeting). */
glint_status glint_kv_pending(glint_ctx *ctx, uint64_t *keys);
/* Organise everything logged since the last flush, after the last evaluated
 * tick (the reference flushes when the count passes kv_restriction,
 * geometry.cu:1752); the next run's end offsets count from that tick
 * (process.cpp:549,271). */
glint_status glint_kv_flush(glint_ctx *ctx, glint_kv_summary *out);
/* Size in bytes of one output part of the last flush. */
glint_status glint_kv_part_size(glint_ctx *ctx, uint32_t part, uint64_t *bytes);
/* Copy one output part of the last flush to host memory (bytes = its size). */
glint_status glint_kv_copy(glint_ctx *ctx, uint32_t part, void *dst, uint64_t bytes);

/* Device memory helpers for callers without a GPU framework. */
glint_status glint_device_malloc(void **ptr, uint64_t bytes, uint32_t device);
glint_status glint_device_free(void *ptr);
/* Device-to-device copy by a plain streaming kernel (16-byte aligned); used
 * to calibrate HBM byte counters against a copy of known size. */
glint_status glint_device_copy(void *dst, const void *src, uint64_t bytes);
glint_status glint_memcpy_to_device(void *dst, const void *src, uint64_t bytes);
glint_status glint_memcpy_to_host(void *dst, const void *src, uint64_t bytes);
glint_status glint_device_count(int *count);

/* Device stepper of the seeded generator (glint_gen.h): advances walkers
 * held in device memory from tick first_tick-1 and writes nticks snapshots of
 * the positions, tick-major, to d_trace (which may be NULL to only advance).
 * Bit-identical to glint_gen_step on the host. */
glint_status glint_gen_walk_device(const glint_gen_config *cfg, glint_walkers *d_walkers,
                                   uint32_t first_tick, uint32_t nticks, double *d_trace,
                                   uint32_t device);
/* One tick of the device walk that also writes the objects that moved as a
 * location-update list (d_pids, d_xy; capacity num_objects) and its length
 * (*d_count, device memory). */
glint_status glint_gen_walk_updates_device(const glint_gen_config *cfg, glint_walkers *d_walkers, uint32_t tick,
                                           uint32_t *d_pids, double *d_xy, uint32_t *d_count, uint32_t device);

/* ------------------------------------------------------------------ reference traces
 * The reference's trace files hold FP64 (lon, lat) points tick-major, one
 * Point (16 B) per object and tick:
 *  - a chunk file "<prefix><st>_<N>.tr" holds ticks st..st+99 of N objects
 *    and nothing else (datagen.cpp:18-26, read by process.cpp:128-144);
 *  - a trace file starts with the number of objects (u32), the number of
 *    ticks (u32) and the bounding box (4 doubles: low x, low y, high x,
 *    high y), then every tick's points (process.cpp:146-187). */
typedef struct glint_tr_header {
    uint32_t num_objects;
    uint32_t duration;
    double low_x, low_y, high_x, high_y;
} glint_tr_header;

/* Header of a trace file. */
glint_status glint_tr_header_read(const char *path, glint_tr_header *out);
/* Ticks st..st+duration-1 of num_objects objects of a trace file into out
 * (duration*num_objects points, tick-major), as the reference loads them
 * (process.cpp:146-187): with more objects than the file holds, the next
 * num_objects-loaded objects are read from the file's later ticks.  Errors:
 * the file cannot be opened or is short (GLINT_ERR_BAD_INPUT), or the range
 * exceeds the file (GLINT_ERR_INVALID_ARGUMENT; the reference asserts). */
glint_status glint_tr_read(const char *path, uint32_t num_objects, uint32_t st, uint32_t duration, double *out);
/* The 100 ticks of a chunk file: prefix + st + "_" + num_objects + ".tr"
 * into out (100*num_objects points). */
glint_status glint_tr_read_chunk(const char *prefix, uint32_t st, uint32_t num_objects, double *out);

/* Planar metres from lon/lat degrees with the reference's conversion
 * (util.h:38-39,57-66): metres = degrees / degree_per_meter, the longitude
 * factor fixed at the box's low latitude, as the filter's x_buffer is
 * (filter.cpp:30-31); origin at the box's low corner. */
typedef struct glint_projection {
    double lon0, lat0;        /* degrees mapped to (0, 0) */
    double deg_per_m_lon;     /* 360 / (sin((90 - |lat0|) * PI / 180) * 40076 * 1000), PI = 3.14159265 */
    double deg_per_m_lat;     /* 360 / (40076 * 1000) */
} glint_projection;
glint_status glint_projection_for_box(double lon_lo, double lat_lo, glint_projection *out);
/* x = (lon - lon0) / deg_per_m_lon, y = (lat - lat0) / deg_per_m_lat for n
 * points (FP64 AoS) in device memory, on the GPU; d_xy may equal
 * d_lonlat.  Synchronous. */
glint_status glint_project_device(const glint_projection *proj, const double *d_lonlat, double *d_xy, uint64_t n,
                                  uint32_t device);

#ifdef __cplusplus
}
#endif

#endif /* GLINT_H */
