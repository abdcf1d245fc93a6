// Device-side data layout of the contact-tracing path on gfx950.
//
// HBM layout (one context):
//   nodes[node_cap]      64 B quadtree nodes: descent midpoints, box, children
//   meta[node_cap]       16 B: type, depth, split/merge hysteresis counters
//   count/offset         u32 per node: leaf population and first slot of the
//                        leaf in the sorted object arrays
//   sx, sy, spid         objects sorted by leaf (FP64 x, FP64 y, object id)
//   fx, fy               the same positions as FP32 offsets from the leaf's
//                        low corner (guard-band mode)
//   units                border check units (object id, candidate leaf)
//   gcell, g10, gpyr,    grid rebuild only: object cells, the coarse 1024^2
//   gheavy, gfine        histogram with its pyramid, and 16x16 cell histograms
//                        of the coarse blocks over capacity
//   gen[2]               the pairs of the last two ticks: each tick's contact
//                        list grouped by table region (8 B keys) with a start
//                        tick per entry (4 B) and the first entry per region
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "glint.h"

namespace glint {

constexpr uint32_t kLeafTag = 0xFFFFFFFFu;      // child[0] of a leaf
constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr unsigned long long kEmptyKey = ~0ull; // (a<<32)|b with a<b is never all ones
constexpr uint32_t kDupStart = 0xFFFFFFFFu;     // start of the second copy of a pair found twice
constexpr uint32_t kCandOverflow = 0xFFFFFFFEu; // third candidate slot of a lookup with more than three
// Flags carried in bits of a contact's key that pids (below 2^30) leave free:
constexpr unsigned long long kTagPrev = 1ull << 63;     // the pair was within reach at the tick before
constexpr unsigned long long kTagBorder = 1ull << 62;   // found through a border unit (may be found twice)
constexpr unsigned long long kTagTested = 1ull << 31;   // with kTagPrev: the join of the tick before tested it
constexpr unsigned long long kKeyBits = 0x3FFFFFFF3FFFFFFFull;   // the pair itself

enum NodeType : int32_t { kInvalid = 0, kBranch = 1, kLeaf = 2 };

enum ErrorBits : uint32_t {
    kErrBadInput = 1u << 0,
    kErrTableFull = 1u << 1,
    kErrNodeCap = 1u << 2,
    kErrUnitCap = 1u << 3,
    kErrMeetCap = 1u << 4,
    kErrContactCap = 1u << 5,
    kErrStack = 1u << 6,
    kErrKvCap = 1u << 7,
    kErrGridEdge = 1u << 8,   // grid rebuild: a position with no cell (on the high y edge)
};

struct alignas(64) Node {
    double mx, my;          // midpoints used by the descent: child = (y>my)*2 + (x>mx)
    double lx, ly, hx, hy;  // box
    uint32_t child[4];      // child[0] == kLeafTag for a leaf
};
static_assert(sizeof(Node) == 64, "one node per 64-byte line");

// An object whose lookup runs this tick: id, last_valid, leaf and position.
struct alignas(8) BorderObj {
    uint32_t pid, lv, leaf, pleaf;   // pleaf: its leaf at the tick before
    uint32_t prevk;                  // its sorted position at the tick before (kNone: none known)
    uint32_t rank;                   // its rank in its leaf (sorted position offset[leaf] + rank)
    double x, y;
};

struct alignas(16) NodeMeta {
    int32_t type;
    int32_t level;
    uint32_t overflow;
    uint32_t underflow;
};

struct Counters {
    unsigned long long n_units;
    unsigned long long n_contacts;
    unsigned long long n_contacts_border;
    unsigned long long n_dup;       // second copies of pairs found twice this tick
    unsigned long long n_stored;    // contacts kept in the tick's generation
    unsigned long long n_meet;
    unsigned long long n_rechecks;
    unsigned long long n_clist;     // contact-list entries reserved this tick (chunks)
    unsigned long long work_border;
    int32_t free_top;
    uint32_t n_split;
    uint32_t n_merge;
    uint32_t n_border;
    uint32_t n_leaf_list;   // join work list: light leaves from the front
    uint32_t n_heavy;       // and leaves of more than kHeavyLeaf objects from the back
    uint32_t n_leaves;
    uint32_t max_leaf;
    uint32_t error;
    uint32_t work_inside;
    uint32_t n_nodes;
    uint32_t last_splits;
    uint32_t last_merges;
    uint32_t n_drop_leaves;         // leaves over 2*cap this tick (the reference's leaf storage, filter.cpp:100-104)
    uint32_t sorted_valid;          // spid holds a complete leaf order (after the first scatter)
    uint32_t tab_bits;              // log2(regions) of the last tick's table
    uint32_t pad0_;
    unsigned long long n_drop_objs; // objects beyond 2*cap in those leaves
    unsigned long long n_tests;     // candidate pairs the join tested (the reference's distance tests)
    unsigned long long n_evals;     // pair slots the join evaluated (64 per column of a wave task)
    unsigned long long n_kvlog;     // meetings logged for the key-value output since its last flush
    unsigned long long n_mlog;      // meetings ever appended to the meeting log
    // grid rebuild (GLINT_SCHEMA_GRID_REBUILD)
    uint32_t g_nodes;               // nodes handed out this tick
    uint32_t g_heavy;               // coarse blocks over capacity this tick
    uint32_t g_list[16];            // branch blocks per depth
    // running totals over ticks (glint_totals); glint_totals_get resets from
    // acc_ticks to the end of the struct, so nothing else may follow them
    unsigned long long acc_ticks;
    unsigned long long acc_contacts;
    unsigned long long acc_contacts_border;
    unsigned long long acc_distinct;
    unsigned long long acc_meet;
    unsigned long long acc_units;
    unsigned long long acc_border;
    unsigned long long acc_rechecks;
    unsigned long long max_contacts;
    unsigned long long max_distinct;
    unsigned long long max_meet;
    unsigned long long acc_stored;
    unsigned long long max_stored;
    unsigned long long acc_drop_leaves;
    unsigned long long acc_drop_objs;
    unsigned long long max_drop_objs;
    unsigned long long acc_tests;
    unsigned long long acc_evals;
    uint32_t max_leaf_run;          // largest leaf over the ticks
    uint32_t pad_;
};

// Per-tick values, kept in device memory so a tick's launches can be
// replayed from a graph; advanced on the device by k_advance_tick.
// One ended meeting kept for the sorted key-value output: the pair, its
// first and last contact tick and the float box of both objects over the
// run (low_x, low_y, high_x, high_y).
struct KvLogEntry {
    uint32_t pid_a, pid_b, start, end;
    float4 box;
};

// One generation of the meeting table: the contacts of one tick grouped by
// table region, keys[i] = (pid_a << 32) | pid_b with the kTag flags;
// starts[i] = first tick of the pair's run (kDupStart: a second copy,
// ignored); with 2^b regions at that tick, region r holds entries
// rstart[r] .. rstart[r+1]; rstart[rs_total] is the entry count and
// rstart[rs_total + 1] is b.
struct TableGen {
    unsigned long long *keys;
    uint32_t *starts;
    uint32_t *rstart;
};

struct TickParams {
    const double *xy;
    const double *trace_base;   // run_ticks: tick k of the trace at trace_base + 2*N*k
    TableGen tab_old;           // pairs seen at prev_tick
    TableGen tab_cur;           // pairs seen at tick (written by the table update)
    const double *xy_prev;      // positions at prev_tick (pid order), valid if has_prev
    float4 *mbr_old;            // boxes of the pairs in tab_old, by entry (key-value output only)
    float4 *mbr_cur;            // boxes of the pairs in tab_cur
    double *xy_save;            // receives this tick's positions for the next tick (null: not needed)
    double *hist_prev;          // the two position-history buffers, swapped every tick: the one
    double *hist_save;          // holding prev_tick's positions and the one for this tick's
    uint32_t *sleaf;            // leaf of each sorted object at tick (written by the scatter)
    uint32_t *sleaf_old;        // the same at prev_tick
    double *sx, *sy;            // objects sorted by leaf at tick (written by the scatter)
    double *sx_old, *sy_old;    // the same at prev_tick (read by the scatter: position k of the
                                // tick's assignment is the object sorted at k then)
    uint4 *cand;                // candidate lists written at tick (see DevState::cand)
    uint4 *cand_prev;           // those written at prev_tick
    unsigned long long *stamps; // run_ticks: device wall clock at the end of every tick
    uint32_t nstamps;
    uint32_t tick;
    uint32_t prev_tick;
    uint32_t trace_index;
    uint32_t has_prev;
    uint32_t next_tick;         // tick number given by the next k_advance_tick
    uint32_t has_prev_next;
    uint32_t nticks;            // run_ticks: ticks of the run
};

struct DevState {
    uint32_t n;
    uint32_t cap;
    uint32_t min_meet;
    uint32_t node_cap;
    uint32_t complete;
    uint32_t precision;
    uint32_t store_all;     // keep every contact in the generation (else only those flagged kTagPrev/kTagBorder)
    uint32_t delay;
    double lx, ly, hx, hy;
    double r;
    double t64;         // largest s with sqrt(s) <= r (host-computed)
    float t32_in;       // FP32 squared-distance bounds of the guard band
    float t32_out;
    uint32_t geography; // lon/lat degrees with the reference's scaled predicates (geometry.h:40-48,174-185)
    double xbuf, ybuf;  // last_valid margins (filter.cpp:30-31): r, or r in degrees at the root's low latitude
    double *sgeo;       // geography: degrees of longitude per metre at each sorted object's latitude
    double w32_max;     // widest leaf decided in FP32 (error bound inside the band)
    double pf_lim;      // largest FP32 offset with the same error bound (w32_max + r)
    unsigned long long unit_cap;
    unsigned long long meet_cap;
    unsigned long long clist_cap;
    uint32_t chunk;     // contact-list entries a wave of the join reserves at a time (64..4096)
    uint32_t pad_chunk_;
    TickParams *tp;
    Node *nodes;
    NodeMeta *meta;
    int32_t *free_stack;
    uint32_t *count;
    uint32_t *offset;
    uint32_t *block_sums;
    uint32_t *split_list;
    uint32_t *merge_list;
    uint32_t *leaf_list;
    uint32_t *parent;       // per node: its parent (kNone for the root), for stackless tree walks
    double4 *region;        // per node: the points the descent sends into it, (x.lo, x.hi] x (y.lo, y.hi]
    uint32_t incremental;   // objects are placed from their leaf of the tick before (not with the grid rebuild)
    uint32_t pad_inc_;
    uint32_t *sleafb[2];    // leaf of every sorted object, two ticks (TickParams::sleaf, sleaf_old)
    uint4 *uaux;            // per unit (lookup order): the object's leaf (bit 31: its candidates
                            // overflowed), its leaf at the tick before, its sorted position at
                            // the tick before (kNone: none known) and at the tick
    uint4 *unit_lf;         // the same grouped like unit_pid
    // Candidate leaves of the lookups, at the object's sorted position of the
    // tick that wrote them, in two buffers taken in turn (TickParams::cand,
    // cand_prev): up to three leaves (the third is kCandOverflow when there
    // were more) and that tick, so an entry not written at the tick read is
    // empty and nothing is reset.
    uint4 *cand;
    uint4 *prev_cand;
    BorderObj *border;      // objects whose last_valid is not their leaf, for the lookup
    uint2 *units;           // (object, candidate leaf) in lookup order
    uint32_t *urank;        // rank of the unit among the units of its leaf
    uint32_t *unit_pid;     // units grouped by candidate leaf
    uint32_t *ucount;       // units per leaf
    uint32_t *ustart;       // first unit of each leaf in unit_pid
    double *sxb[2], *syb[2]; // sorted positions of two ticks (TickParams::sx, sx_old)
    uint32_t *sprev;        // sorted position at prev_tick of each sorted object (kNone: none known)
    uint32_t *spid;
    uint32_t *order;        // object handled by position k of the tick's assignment (spid of the tick before)
    double2 *ord_xy;        // its position, read once by the assignment
    uint32_t *ord_leaf;     // its leaf and rank in that leaf, by k
    uint32_t *ord_rank;
    uint32_t *ord_pleaf;    // its leaf at the tick before, by k
    uint32_t *splf;         // leaf at the tick before, sorted like sx/sy
    float *fx;
    float *fy;
    float *pfx;             // positions at the tick before as FP32 offsets from the same corner
    float *pfy;             // (NaN beyond pf_lim: left to the FP64 test)
    glint_meeting *meet;
    unsigned long long *clist2;    // bucketing buffer (the join's output when there is one pass)
    uint32_t *bcount;              // coarse-bin counts per bucketing block
    uint32_t *boff;                // their exclusive offsets
    uint32_t rbits_min;            // log2(regions) of the configured table: the fewest a tick uses
    uint32_t rbits_max;            // the most a tick can need (its stored contacts are below clist_cap)
    uint32_t coarse_max;           // first-pass bins at rbits_max
    uint32_t rs_total;             // 2^rbits_max: rstart index of a generation's entry count (its
                                   // region bits follow at rs_total + 1)
    uint32_t bucket_blocks;        // blocks of the first pass
    KvLogEntry *kvlog;             // key-value output log (null when off)
    float4 *pos32;                 // float positions of the tick and of the tick before (x, y, px, py),
                                   // pid order (key-value output)
    glint_meeting *mlog;           // meeting log ring (null when off)
    unsigned long long mlog_cap;
    unsigned long long kvlog_cap;
    // grid rebuild (null unless GLINT_SCHEMA_GRID_REBUILD)
    uint32_t *gcell;               // cell x + 16384*y of every object (pid order)
    uint32_t *g10;                 // points per block of the 1024^2 coarse grid (depth 10)
    uint32_t *gpyr;                // points per block at depths 0..9, depth d from (4^d-1)/3
    uint32_t *gheavy;              // per coarse block: its fine histogram, or kNone
    uint32_t *gfine;               // 16x16 cell counts of every coarse block over capacity
    uint2 *glist[2];               // branch blocks of one depth (node, x | y << 16), ping-pong
    uint32_t heavy_cap;            // fine histograms held
    uint32_t glist_cap;            // branch blocks held per depth
    Counters *ctr;
};

// SplitMix64 finaliser (Steele, Lea, Flood 2014): slot hash of a pair key.
__device__ __forceinline__ unsigned long long mix64(unsigned long long z)
{
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Wave-aggregated append: one atomic per wave, lanes get consecutive slots.
// Must be called by every active lane of the wave with its own predicate.
__device__ __forceinline__ unsigned long long wave_append_u64(unsigned long long *counter, bool pred)
{
    unsigned long long mask = __ballot(pred);
    if (mask == 0) return ~0ull;
    uint32_t lane = lane_id();
    uint32_t leader = __ffsll((long long)mask) - 1;
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(counter, (unsigned long long)__popcll(mask));
    base = __shfl(base, leader);
    return base + __popcll(mask & ((1ull << lane) - 1ull));
}

// Sum over the wave's lanes (every lane gets it).
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v)
{
    for (int d = 32; d >= 1; d >>= 1) {
        const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, d);
        const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), d);
        v += ((unsigned long long)hi << 32) | lo;
    }
    return v;
}

__device__ __forceinline__ uint32_t wave_append_u32(uint32_t *counter, bool pred)
{
    unsigned long long mask = __ballot(pred);
    if (mask == 0) return kNone;
    uint32_t lane = lane_id();
    uint32_t leader = __ffsll((long long)mask) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, leader);
    return base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
}

}  // namespace glint
