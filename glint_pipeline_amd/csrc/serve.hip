// serve.hip - the serving loop: location updates ingested every tick into
// device-resident state, and tracing queries admitted by HBM footprint and
// run on their own streams while ticks continue.
//
// A query (seed pids, window [t_start, t_end], 1 or 2 hops) is answered on
// the state of its submission tick T, whenever it is admitted:
//  - hop 1, ended meetings: every meeting in the log up to T (ended by T)
//    with a seed on either side and start <= t_end, end >= t_start;
//  - hop 1, meetings in progress at T (the reference's search_single rule,
//    geometry.cu:605-662: still active and T - start >= min + 1), clipped to
//    end = T: pairs of the table at admission with start <= T, plus pairs
//    logged after T with start <= T (they were in progress at T and have
//    ended since);
//  - hop 2 (the targets of hop 1 searched again, process.cpp:469-495):
//    meetings ended by T of a hop-1 contact c that is not a seed, with end >=
//    the start of c's first hop-1 contact.
// The answer therefore depends only on T, not on the admission delay, so a
// query gives the same result alone and in any mix.
//
// Meeting index: every index_block_ticks ticks the meetings logged since the
// last block are sealed into a block of directed entries (object, target,
// start, end), two per meeting, sorted by (object, target, start) with a
// stable radix sort, as the reference keeps its flushed meetings sorted by
// object for id searches (workbench.cpp:397-444, step_merge.cpp:501-568).  A
// query binary-searches every block that can hold meetings of its window for
// each seed (hop 1) or hop-1 contact (hop 2), keeping the meetings that ended
// by T (end < T) and overlap the window; only the log after the last block
// sealed is scanned.  Blocks and the scanned tail split the log at that
// block's end, so every meeting is read once.
//
// Admission (continuous batching): queries wait in FIFO order; one is
// admitted when a stream slot is free and its footprint (seed bitmap, seeds,
// hit buffers) fits a first-fit allocator over a preallocated device arena.
// A query whose hits overflow its estimate is re-queued with the exact size.

#include <hip/hip_runtime.h>

#include <cstring>   // before rocPRIM, whose iterators use memset

#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "ctx_view.h"
#include "glint.h"
#include "glint_device.h"

namespace {

constexpr uint32_t kB = 256;
constexpr uint64_t kAlign = 256;
constexpr unsigned long long kEmpty = ~0ull;
constexpr uint32_t kDup = 0xFFFFFFFFu;      // start of a second copy of a pair in a generation
constexpr unsigned long long kPairBits = 0x3FFFFFFF3FFFFFFFull;   // a list key without its flag bits

struct QCounters {
    unsigned long long hits1, hits2;
};

__global__ void k_fill_words(uint32_t *p, uint64_t n)
{
    for (uint64_t i = blockIdx.x * (uint64_t)kB + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kB) p[i] = 0;
}

__global__ void k_set_bits(const uint32_t *ids, uint32_t n, uint32_t *bitmap)
{
    for (uint32_t i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB)
        atomicOr(&bitmap[ids[i] >> 5], 1u << (ids[i] & 31));
}

__device__ __forceinline__ bool has_bit(const uint32_t *bm, uint32_t id) { return (bm[id >> 5] >> (id & 31)) & 1u; }

// Wave-aggregated append of one hit per active lane.
__device__ __forceinline__ void append_hit(bool put, glint_hit h, glint_hit *out, unsigned long long *count,
                                           uint64_t cap)
{
    const unsigned long long m = __ballot(put);
    if (!m) return;
    unsigned long long base = 0;
    if (__lane_id() == 0) base = atomicAdd(count, (unsigned long long)__popcll(m));
    base = __shfl(base, 0);
    if (put) {
        const unsigned long long j = base + __popcll(m & ((1ull << __lane_id()) - 1ull));
        if (j < cap) out[j] = h;
    }
}

// Pairs of the table with a seed on either side that were in progress at
// tick T (start <= T, T - start >= min + 1), clipped to end = T.
__global__ __launch_bounds__(kB) void k_active_hits(const unsigned long long *keys, const uint32_t *starts,
                                                    const uint32_t *count, uint32_t num_objects, uint32_t T,
                                                    uint32_t min_meet, uint32_t t_start, uint32_t t_end,
                                                    const uint32_t *seeds_bm, glint_hit *out,
                                                    unsigned long long *hits, uint64_t cap)
{
    const uint64_t n = *count;
    for (uint64_t i0 = blockIdx.x * (uint64_t)kB; i0 < n; i0 += (uint64_t)gridDim.x * kB) {
        const uint64_t i = i0 + threadIdx.x;
        unsigned long long key = kEmpty;
        uint32_t start = kDup;
        if (i < n) {
            key = keys[i] & kPairBits;
            start = starts[i];
        }
        const uint32_t a = (uint32_t)(key >> 32), b = (uint32_t)key;
        const bool valid = start != kDup && a < num_objects && b < num_objects && start <= T &&
                           T - start >= min_meet + 1 && start <= t_end && T >= t_start;
        const bool ha = valid && has_bit(seeds_bm, a), hb = valid && has_bit(seeds_bm, b);
        append_hit(ha, glint_hit{a, b, start, T}, out, hits, cap);
        append_hit(hb, glint_hit{b, a, start, T}, out, hits, cap);
    }
}

// Hop 1 over log entries [lo, hi_adm): entries before hi_sub ended by T;
// later ones count only if they were in progress at T (clipped to end = T).
__global__ __launch_bounds__(kB) void k_log_hop1(const glint_meeting *log, uint64_t log_cap, uint64_t lo,
                                                 uint64_t hi_sub, uint64_t hi_adm, uint32_t T, uint32_t min_meet,
                                                 uint32_t t_start, uint32_t t_end, const uint32_t *seeds_bm,
                                                 glint_hit *out, unsigned long long *count, uint64_t cap)
{
    for (uint64_t i0 = lo + blockIdx.x * (uint64_t)kB; i0 < hi_adm; i0 += (uint64_t)gridDim.x * kB) {
        const uint64_t i = i0 + threadIdx.x;
        bool valid = false;
        glint_meeting m{};
        if (i < hi_adm) {
            m = log[i % log_cap];
            if (i < hi_sub)
                valid = m.start <= t_end && m.end >= t_start;
            else if (m.start <= T && T - m.start >= min_meet + 1 && m.start <= t_end && T >= t_start) {
                valid = true;
                m.end = T;
            }
        }
        const bool ha = valid && has_bit(seeds_bm, m.pid_a), hb = valid && has_bit(seeds_bm, m.pid_b);
        append_hit(ha, glint_hit{m.pid_a, m.pid_b, m.start, m.end}, out, count, cap);
        append_hit(hb, glint_hit{m.pid_b, m.pid_a, m.start, m.end}, out, count, cap);
    }
}

__device__ __forceinline__ bool exposure_start(const uint32_t *ids, const uint32_t *starts, uint32_t n, uint32_t id,
                                               uint32_t &start)
{
    uint32_t a = 0, b = n;
    while (a < b) {
        const uint32_t m = (a + b) / 2;
        if (ids[m] < id)
            a = m + 1;
        else
            b = m;
    }
    if (a < n && ids[a] == id) {
        start = starts[a];
        return true;
    }
    return false;
}

// Hop 2 over the meetings ended by T: a hop-1 contact c (sorted ids, with the
// start of its first exposure) as source, end >= that start.
__global__ __launch_bounds__(kB) void k_log_hop2(const glint_meeting *log, uint64_t log_cap, uint64_t lo,
                                                 uint64_t hi_sub, uint32_t t_start, uint32_t t_end,
                                                 const uint32_t *hop_bm, const uint32_t *ids, const uint32_t *starts,
                                                 uint32_t nids, glint_hit *out, unsigned long long *count,
                                                 uint64_t cap)
{
    for (uint64_t i0 = lo + blockIdx.x * (uint64_t)kB; i0 < hi_sub; i0 += (uint64_t)gridDim.x * kB) {
        const uint64_t i = i0 + threadIdx.x;
        glint_meeting m{};
        bool valid = false;
        if (i < hi_sub) {
            m = log[i % log_cap];
            valid = m.start <= t_end && m.end >= t_start;
        }
        bool ha = valid && has_bit(hop_bm, m.pid_a), hb = valid && has_bit(hop_bm, m.pid_b);
        uint32_t s;
        if (ha) ha = exposure_start(ids, starts, nids, m.pid_a, s) && m.end >= s;
        if (hb) hb = exposure_start(ids, starts, nids, m.pid_b, s) && m.end >= s;
        append_hit(ha, glint_hit{m.pid_a, m.pid_b, m.start, m.end}, out, count, cap);
        append_hit(hb, glint_hit{m.pid_b, m.pid_a, m.start, m.end}, out, count, cap);
    }
}

// Directed index entries of log entries [l0, l1) (ring positions): two per
// meeting, key (object << 32) | target, value (start << 32) | end.
__global__ __launch_bounds__(kB) void k_index_expand(const glint_meeting *log, uint64_t log_cap, uint64_t l0,
                                                     uint64_t l1, unsigned long long *keys, unsigned long long *vals)
{
    for (uint64_t i = l0 + blockIdx.x * (uint64_t)kB + threadIdx.x; i < l1; i += (uint64_t)gridDim.x * kB) {
        const glint_meeting m = log[i % log_cap];
        const uint64_t j = 2 * (i - l0);
        const unsigned long long v = ((unsigned long long)m.start << 32) | m.end;
        keys[j] = ((unsigned long long)m.pid_a << 32) | m.pid_b;
        keys[j + 1] = ((unsigned long long)m.pid_b << 32) | m.pid_a;
        vals[j] = v;
        vals[j + 1] = v;
    }
}

// One (object, block) pair per thread: the block's entries of the object
// (a binary search for its first), kept if they ended by T, overlap the
// window and (hop 2) end at or after the object's first exposure.
__global__ __launch_bounds__(kB) void k_index_hits(const unsigned long long *keys, const unsigned long long *vals,
                                                   const ulonglong2 *blocks, uint32_t nblocks, const uint32_t *ids,
                                                   const uint32_t *exposure, uint32_t nids, uint32_t T,
                                                   uint32_t t_start, uint32_t t_end, glint_hit *out,
                                                   unsigned long long *count, uint64_t cap)
{
    const uint64_t total = (uint64_t)nids * nblocks;
    for (uint64_t w = blockIdx.x * (uint64_t)kB + threadIdx.x; w < total; w += (uint64_t)gridDim.x * kB) {
        const uint32_t id = ids[w / nblocks];
        const ulonglong2 b = blocks[w % nblocks];   // (first entry, entries)
        const unsigned long long want = (unsigned long long)id << 32;
        uint64_t lo = b.x, hi = b.x + b.y;
        while (lo < hi) {
            const uint64_t mid = (lo + hi) / 2;
            if (keys[mid] < want)
                lo = mid + 1;
            else
                hi = mid;
        }
        const uint32_t from = exposure ? exposure[w / nblocks] : 0u;
        for (uint64_t j = lo; j < b.x + b.y; j++) {
            const unsigned long long k = keys[j];
            if ((uint32_t)(k >> 32) != id) break;
            const unsigned long long v = vals[j];
            const uint32_t start = (uint32_t)(v >> 32), end = (uint32_t)v;
            if (end < T && start <= t_end && end >= t_start && end >= from) {
                const unsigned long long at = atomicAdd(count, 1ull);
                if (at < cap) out[at] = glint_hit{id, (uint32_t)k, start, end};
            }
        }
    }
}

inline uint32_t grid_of(uint64_t n, uint32_t cap = 4096)
{
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + kB - 1) / kB, cap));
}

inline uint64_t up(uint64_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

// First-fit allocator over [0, size) of the arena.
struct Arena {
    uint64_t size = 0;
    std::map<uint64_t, uint64_t> free_;   // offset -> bytes

    void init(uint64_t s)
    {
        size = s;
        free_.clear();
        free_[0] = s;
    }
    bool alloc(uint64_t bytes, uint64_t &off)
    {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                off = it->first;
                const uint64_t rest = it->second - bytes;
                free_.erase(it);
                if (rest) free_[off + bytes] = rest;
                return true;
            }
        }
        return false;
    }
    void release(uint64_t off, uint64_t bytes)
    {
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    uint64_t free_bytes() const
    {
        uint64_t s = 0;
        for (auto &kv : free_) s += kv.second;
        return s;
    }
};

enum QState : uint32_t { kPending = GLINT_QUERY_PENDING, kRunning = GLINT_QUERY_RUNNING, kDone = GLINT_QUERY_DONE,
                         kFailed = GLINT_QUERY_FAILED };

struct Query {
    uint64_t id = 0;
    std::vector<uint32_t> seeds;
    uint32_t t_start = 0, t_end = 0, hops = 1;
    uint32_t T = 0;              // submission tick
    uint64_t hi_sub = 0, lo = 0; // log range fixed at submission
    uint64_t cap1 = 0, cap2 = 0;
    QState state = kPending;
    int phase = 0;
    int slot = -1;
    uint64_t off = 0, bytes = 0;
    uint32_t retries = 0;
    uint32_t admitted_tick = 0;
    double t_submit = 0, t_admit = 0, t_done = 0;
    std::vector<glint_hit> hits1, hits2;
    std::vector<uint32_t> ids_h, starts_h;   // hop-1 contacts and their first exposure (hop 2)
    // device pointers inside the arena
    uint32_t *bm1 = nullptr, *bm2 = nullptr, *seeds_d = nullptr, *ids_d = nullptr, *starts_d = nullptr;
    glint_hit *h1 = nullptr, *h2 = nullptr;
    QCounters *ctr = nullptr;
    uint32_t nids = 0;
    ulonglong2 *blocks_d = nullptr;   // index blocks searched (first entry, entries), in the arena
    std::vector<ulonglong2> blocks_h;
    uint32_t nblocks = 0, room = 0;
    uint64_t tail = 0;                // log entries from here on are scanned
};

// One sealed block of the meeting index.
struct IndexBlock {
    uint64_t off, n;       // its entries in the index storage
    uint64_t l0, l1;       // the log entries it holds
    uint32_t t0, t1;       // the ticks that logged them
};

struct MeetingIndex {
    unsigned long long *keys = nullptr, *vals = nullptr;     // block storage (a ring of blocks)
    uint64_t cap = 0;
    unsigned long long *skeys = nullptr, *svals = nullptr;   // a block before its sort
    uint64_t block_max = 0;                                  // entries a block holds at most
    void *tmp = nullptr;
    size_t tmp_bytes = 0;
    std::deque<IndexBlock> blocks;
    uint64_t next = 0;          // where the next block goes
    uint64_t sealed = 0;        // log entries below this are in blocks
    uint32_t first_tick = 0;    // the first tick not in a block
    uint32_t block_ticks = 64;
    uint64_t entries = 0;       // entries in the blocks held
};

struct Slot_ {
    hipStream_t stream = nullptr;
    hipEvent_t ready = nullptr, snap = nullptr, done = nullptr;
    uint64_t query = 0;   // 0: free
};

}  // namespace

struct glint_server {
    glint_ctx *ctx = nullptr;
    glint_server_config cfg{};
    CtxView view{};
    uint8_t *arena_d = nullptr;
    Arena arena;
    std::vector<Slot_> slots;
    std::deque<uint64_t> pending;
    std::map<uint64_t, Query> queries;
    std::map<uint32_t, uint64_t> log_after;   // tick -> meetings logged after it
    MeetingIndex index;
    // the index is sealed on a stream of its own after each tick, so that a
    // sort does not hold up the next tick; readers wait for `sealed`
    hipStream_t index_stream = nullptr;
    hipEvent_t tick_done = nullptr, sealed = nullptr;
    bool any_sealed = false;
    uint64_t next_id = 1;
    bool have_tick = false;
    double meetings_per_tick = 0;             // running average, for hit estimates
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    std::string err;

    double now_ms() const
    {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace {

#define SV_TRY(s, call)                                                     \
    do {                                                                    \
        hipError_t e_ = (call);                                             \
        if (e_ != hipSuccess) {                                             \
            (s)->err = std::string(#call) + ": " + hipGetErrorString(e_);   \
            return GLINT_ERR_HIP;                                           \
        }                                                                   \
    } while (0)

uint64_t bitmap_bytes(uint32_t n) { return up(((uint64_t)n + 31) / 32 * 4); }

constexpr uint32_t kBlockSlack = 64;   // index blocks sealed between admission and hop 2 that a query may use

// The index blocks a query searches for meetings ended by T within the
// window (at most `room`, oldest first), and where the scan of the log
// takes over: the end of the last block used, within [lo, hi_sub].
void index_blocks(const glint_server *s, const Query &q, uint32_t room, std::vector<ulonglong2> &out,
                  uint64_t &tail)
{
    out.clear();
    uint64_t covered = q.lo;   // log entries [lo, covered) are in the blocks taken
    for (const IndexBlock &b : s->index.blocks) {
        if (b.l1 <= q.lo) continue;                          // ended before the window
        if (out.size() >= room || b.l0 >= q.hi_sub || b.l0 > covered) break;   // full, after T, or a gap
        covered = b.l1;
        if (b.t1 > q.t_start) out.push_back(make_ulonglong2(b.off, b.n));   // else all ended before the window
    }
    tail = std::min(covered, q.hi_sub);
}

uint64_t footprint(const glint_server *s, const Query &q)
{
    uint64_t b = up(sizeof(QCounters)) + bitmap_bytes(s->view.num_objects) + up(4ull * q.seeds.size()) +
                 up(q.cap1 * sizeof(glint_hit));
    if (q.hops >= 2)
        b += bitmap_bytes(s->view.num_objects) + 2 * up(4ull * q.cap1) + up(q.cap2 * sizeof(glint_hit));
    b += up(sizeof(ulonglong2) * (s->index.blocks.size() + kBlockSlack));   // the index blocks searched
    return b;
}

void estimate_caps(const glint_server *s, Query &q)
{
    // meetings touching one object per tick, times the window, both ends
    const double per_obj = s->meetings_per_tick * 2.0 / std::max<uint32_t>(1, s->view.num_objects);
    const double window = (double)(q.t_end - q.t_start + 1) + 1.0;
    const double scale = (s->cfg.estimate_percent ? s->cfg.estimate_percent : 100) / 100.0;
    const double est = (double)q.seeds.size() * (per_obj * window + 1.0) * 2.0 * scale;
    q.cap1 = std::max<uint64_t>(s->cfg.min_hits, (uint64_t)est);
    q.cap2 = q.hops >= 2 ? std::max<uint64_t>(s->cfg.min_hits, (uint64_t)((double)q.cap1 * (per_obj * window + 1.0)))
                         : 0;
}

void release(glint_server *s, Query &q)
{
    if (q.bytes) s->arena.release(q.off, q.bytes);
    q.bytes = 0;
    if (q.slot >= 0) s->slots[q.slot].query = 0;
    q.slot = -1;
}

glint_status fail_query(glint_server *s, Query &q, const std::string &why)
{
    release(s, q);
    q.state = kFailed;
    q.t_done = s->now_ms();
    s->err = "query " + std::to_string(q.id) + ": " + why;
    return GLINT_OK;
}

// Admission of the query: arena space and a stream slot, then its phase-1
// work: seed bitmap, table hits on the tick stream, log hits on its own.
glint_status start_query(glint_server *s, Query &q, int slot)
{
    uint8_t *base = s->arena_d + q.off;
    auto take = [&](uint64_t bytes) {
        uint8_t *p = base;
        base += up(bytes);
        return p;
    };
    const uint32_t N = s->view.num_objects;
    const uint64_t bmw = ((uint64_t)N + 31) / 32;
    q.ctr = (QCounters *)take(sizeof(QCounters));
    q.bm1 = (uint32_t *)take(bmw * 4);
    q.seeds_d = (uint32_t *)take(4ull * q.seeds.size());
    q.h1 = (glint_hit *)take(q.cap1 * sizeof(glint_hit));
    if (q.hops >= 2) {
        q.bm2 = (uint32_t *)take(bmw * 4);
        q.ids_d = (uint32_t *)take(4ull * q.cap1);
        q.starts_d = (uint32_t *)take(4ull * q.cap1);
        q.h2 = (glint_hit *)take(q.cap2 * sizeof(glint_hit));
    }
    q.room = (uint32_t)(s->index.blocks.size() + kBlockSlack);
    q.blocks_d = (ulonglong2 *)take(sizeof(ulonglong2) * q.room);
    Slot_ &sl = s->slots[slot];
    sl.query = q.id;
    q.slot = slot;
    hipStream_t st = sl.stream;
    SV_TRY(s, hipMemsetAsync(q.ctr, 0, sizeof(QCounters), st));
    k_fill_words<<<grid_of(bmw), kB, 0, st>>>(q.bm1, bmw);
    SV_TRY(s, hipMemcpyAsync(q.seeds_d, q.seeds.data(), 4ull * q.seeds.size(), hipMemcpyHostToDevice, st));
    k_set_bits<<<grid_of(q.seeds.size()), kB, 0, st>>>(q.seeds_d, (uint32_t)q.seeds.size(), q.bm1);
    SV_TRY(s, hipEventRecord(sl.ready, st));
    // table hits, ordered with the ticks on the context's stream
    CtxView v{};
    glint_ctx_view(s->ctx, &v);
    SV_TRY(s, hipStreamWaitEvent(v.stream, sl.ready, 0));
    k_active_hits<<<grid_of(v.cap, 8192), kB, 0, v.stream>>>(v.keys, v.starts, v.count, v.num_objects, q.T,
                                                             v.min_meet, q.t_start, q.t_end, q.bm1, q.h1,
                                                             &q.ctr->hits1, q.cap1);
    SV_TRY(s, hipEventRecord(sl.snap, v.stream));
    // log hits on the query's stream
    uint64_t total = 0, first = 0;
    glint_status gs = glint_meeting_log_info(s->ctx, &total, &first);
    if (gs) return gs;
    // an eighth of the ring stays as guard band for the ticks that run
    // while the scan does
    if (q.lo < first + v.mlog_cap / 8 && first > 0)
        return fail_query(s, q, "window no longer retained by the meeting log");
    index_blocks(s, q, q.room, q.blocks_h, q.tail);
    if (total > q.tail)
        k_log_hop1<<<grid_of(total - q.tail, 8192), kB, 0, st>>>(v.mlog, v.mlog_cap, q.tail, q.hi_sub, total, q.T,
                                                                 v.min_meet, q.t_start, q.t_end, q.bm1, q.h1,
                                                                 &q.ctr->hits1, q.cap1);
    // the table hits, and the sorts of the blocks listed
    SV_TRY(s, hipStreamWaitEvent(st, sl.snap, 0));
    if (s->any_sealed) SV_TRY(s, hipStreamWaitEvent(st, s->sealed, 0));
    q.nblocks = (uint32_t)q.blocks_h.size();
    if (q.nblocks) {
        SV_TRY(s, hipMemcpyAsync(q.blocks_d, q.blocks_h.data(), sizeof(ulonglong2) * q.nblocks,
                                 hipMemcpyHostToDevice, st));
        const uint64_t work = (uint64_t)q.seeds.size() * q.nblocks;
        k_index_hits<<<grid_of(work, 8192), kB, 0, st>>>(s->index.keys, s->index.vals, q.blocks_d, q.nblocks,
                                                         q.seeds_d, nullptr, (uint32_t)q.seeds.size(), q.T,
                                                         q.t_start, q.t_end, q.h1, &q.ctr->hits1, q.cap1);
    }
    SV_TRY(s, hipEventRecord(sl.done, st));
    SV_TRY(s, hipGetLastError());
    q.state = kRunning;
    q.phase = 1;
    q.admitted_tick = v.tick;
    q.t_admit = s->now_ms();
    return GLINT_OK;
}

glint_status admit(glint_server *s)
{
    while (!s->pending.empty()) {
        int slot = -1;
        for (size_t i = 0; i < s->slots.size(); i++)
            if (!s->slots[i].query) {
                slot = (int)i;
                break;
            }
        if (slot < 0) return GLINT_OK;
        Query &q = s->queries[s->pending.front()];
        const uint64_t need = footprint(s, q);
        if (need > s->cfg.arena_bytes) {
            s->pending.pop_front();
            fail_query(s, q, "footprint larger than the whole arena");
            continue;
        }
        if (!s->arena.alloc(need, q.off)) return GLINT_OK;   // FIFO: wait until memory frees up
        q.bytes = need;
        s->pending.pop_front();
        glint_status st = start_query(s, q, slot);
        if (st) return st;
    }
    return GLINT_OK;
}

// Phase transitions of running queries whose work has finished.
glint_status advance(glint_server *s)
{
    for (Slot_ &sl : s->slots) {
        if (!sl.query) continue;
        hipError_t e = hipEventQuery(sl.done);
        if (e == hipErrorNotReady) continue;
        if (e != hipSuccess) SV_TRY(s, e);
        Query &q = s->queries[sl.query];
        QCounters c{};
        SV_TRY(s, hipMemcpy(&c, q.ctr, sizeof(c), hipMemcpyDeviceToHost));
        if (q.phase == 1) {
            if (c.hits1 > q.cap1) {   // estimate too small: again, with the exact size
                release(s, q);
                q.cap1 = c.hits1 + 1;
                if (q.hops >= 2) q.cap2 = std::max(q.cap2, q.cap1);
                q.state = kPending;
                q.retries++;
                s->pending.push_front(q.id);
                continue;
            }
            q.hits1.resize(c.hits1);
            if (c.hits1) SV_TRY(s, hipMemcpy(q.hits1.data(), q.h1, c.hits1 * sizeof(glint_hit), hipMemcpyDeviceToHost));
            if (q.hops < 2) {
                release(s, q);
                q.state = kDone;
                q.t_done = s->now_ms();
                continue;
            }
            // hop-1 contacts that are not seeds, with their first exposure
            std::vector<uint32_t> seeds = q.seeds;
            std::sort(seeds.begin(), seeds.end());
            std::map<uint32_t, uint32_t> first;
            for (const glint_hit &h : q.hits1) {
                if (std::binary_search(seeds.begin(), seeds.end(), h.target)) continue;
                auto it = first.find(h.target);
                if (it == first.end() || h.start < it->second) first[h.target] = h.start;
            }
            std::vector<uint32_t> &ids = q.ids_h, &starts = q.starts_h;
            ids.clear();
            starts.clear();
            for (auto &kv : first) {
                ids.push_back(kv.first);
                starts.push_back(kv.second);
            }
            q.nids = (uint32_t)ids.size();
            const uint32_t N = s->view.num_objects;
            const uint64_t bmw = ((uint64_t)N + 31) / 32;
            hipStream_t st = sl.stream;
            k_fill_words<<<grid_of(bmw), kB, 0, st>>>(q.bm2, bmw);
            if (q.nids) {
                SV_TRY(s, hipMemcpyAsync(q.ids_d, ids.data(), 4ull * q.nids, hipMemcpyHostToDevice, st));
                SV_TRY(s, hipMemcpyAsync(q.starts_d, starts.data(), 4ull * q.nids, hipMemcpyHostToDevice, st));
                k_set_bits<<<grid_of(q.nids), kB, 0, st>>>(q.ids_d, q.nids, q.bm2);
                CtxView v{};
                glint_ctx_view(s->ctx, &v);
                uint64_t tail = 0;
                index_blocks(s, q, q.room, q.blocks_h, tail);
                if (q.hi_sub > tail)
                    k_log_hop2<<<grid_of(q.hi_sub - tail, 8192), kB, 0, st>>>(
                        v.mlog, v.mlog_cap, tail, q.hi_sub, q.t_start, q.t_end, q.bm2, q.ids_d, q.starts_d, q.nids,
                        q.h2, &q.ctr->hits2, q.cap2);
                q.nblocks = (uint32_t)q.blocks_h.size();
                if (q.nblocks) {
                    // blocks sealed since admission: their sorts are ordered
                    // before this point of the index stream
                    if (s->any_sealed) SV_TRY(s, hipStreamWaitEvent(st, s->sealed, 0));
                    SV_TRY(s, hipMemcpyAsync(q.blocks_d, q.blocks_h.data(), sizeof(ulonglong2) * q.nblocks,
                                             hipMemcpyHostToDevice, st));
                    const uint64_t work = (uint64_t)q.nids * q.nblocks;
                    k_index_hits<<<grid_of(work, 8192), kB, 0, st>>>(s->index.keys, s->index.vals, q.blocks_d,
                                                                     q.nblocks, q.ids_d, q.starts_d, q.nids, q.T,
                                                                     q.t_start, q.t_end, q.h2, &q.ctr->hits2,
                                                                     q.cap2);
                }
            }
            SV_TRY(s, hipEventRecord(sl.done, st));
            q.phase = 2;
            continue;
        }
        if (c.hits2 > q.cap2) {
            release(s, q);
            q.cap2 = c.hits2 + 1;
            q.state = kPending;
            q.retries++;
            q.hits1.clear();
            s->pending.push_front(q.id);
            continue;
        }
        q.hits2.resize(c.hits2);
        if (c.hits2) SV_TRY(s, hipMemcpy(q.hits2.data(), q.h2, c.hits2 * sizeof(glint_hit), hipMemcpyDeviceToHost));
        release(s, q);
        q.state = kDone;
        q.t_done = s->now_ms();
    }
    return GLINT_OK;
}

// Seals log entries [l0, l1) into one index block, on the index stream
// (after the tick that logged them, which index_tick orders).
glint_status seal_block(glint_server *s, uint64_t l0, uint64_t l1, uint32_t tick)
{
    MeetingIndex &ix = s->index;
    const uint64_t n = 2 * (l1 - l0);
    if (ix.next + n > ix.cap) ix.next = 0;   // blocks form a ring in the storage
    const uint64_t a = ix.next, b = a + n;
    for (auto it = ix.blocks.begin(); it != ix.blocks.end();) {
        if (it->off < b && a < it->off + it->n) {   // overwritten by the new block
            ix.entries -= it->n;
            it = ix.blocks.erase(it);
        } else {
            ++it;
        }
    }
    hipStream_t st = s->index_stream;
    k_index_expand<<<grid_of(l1 - l0, 4096), kB, 0, st>>>(s->view.mlog, s->view.mlog_cap, l0, l1, ix.skeys,
                                                           ix.svals);
    size_t bytes = ix.tmp_bytes;
    // stable: a pair's meetings stay in log order, which is the order of their starts
    SV_TRY(s, rocprim::radix_sort_pairs(ix.tmp, bytes, ix.skeys, ix.keys + a, ix.svals, ix.vals + a, (size_t)n, 0,
                                        62, st));
    SV_TRY(s, hipGetLastError());
    SV_TRY(s, hipEventRecord(s->sealed, st));
    s->any_sealed = true;
    ix.blocks.push_back(IndexBlock{a, n, l0, l1, ix.first_tick, tick});
    ix.entries += n;
    ix.next = b;
    ix.sealed = l1;
    ix.first_tick = tick + 1;
    return GLINT_OK;
}

// After a tick: seal the log since the last block when index_block_ticks
// ticks have passed, or in pieces of block_max entries when it grows faster.
glint_status index_tick(glint_server *s, uint64_t total, uint64_t first, uint32_t tick)
{
    MeetingIndex &ix = s->index;
    if (!ix.keys) return GLINT_OK;
    if (ix.sealed < first) ix.sealed = first;   // entries the ring lost before they were sealed
    const uint64_t most = ix.block_max / 2;     // meetings per block
    if (total - ix.sealed >= most || (total > ix.sealed && tick + 1 - ix.first_tick >= ix.block_ticks)) {
        // the sorts read what this tick logged
        SV_TRY(s, hipEventRecord(s->tick_done, s->view.stream));
        SV_TRY(s, hipStreamWaitEvent(s->index_stream, s->tick_done, 0));
    }
    while (total - ix.sealed >= most) {
        glint_status st = seal_block(s, ix.sealed, ix.sealed + most, tick);
        if (st) return st;
    }
    if (total > ix.sealed && tick + 1 - ix.first_tick >= ix.block_ticks) return seal_block(s, ix.sealed, total, tick);
    return GLINT_OK;
}

bool hit_less(const glint_hit &a, const glint_hit &b)
{
    if (a.source != b.source) return a.source < b.source;
    if (a.target != b.target) return a.target < b.target;
    if (a.start != b.start) return a.start < b.start;
    return a.end < b.end;
}

}  // namespace

extern "C" {

glint_status glint_server_destroy(glint_server *s)
{
    if (!s) return GLINT_ERR_INVALID_ARGUMENT;
    for (Slot_ &sl : s->slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        if (sl.ready) (void)hipEventDestroy(sl.ready);
        if (sl.snap) (void)hipEventDestroy(sl.snap);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
    }
    if (s->view.stream) (void)hipStreamSynchronize(s->view.stream);
    if (s->index_stream) (void)hipStreamSynchronize(s->index_stream);
    if (s->tick_done) (void)hipEventDestroy(s->tick_done);
    if (s->sealed) (void)hipEventDestroy(s->sealed);
    if (s->index_stream) (void)hipStreamDestroy(s->index_stream);
    if (s->arena_d) (void)hipFree(s->arena_d);
    for (void *p : {(void *)s->index.keys, (void *)s->index.vals, (void *)s->index.skeys, (void *)s->index.svals,
                    s->index.tmp})
        if (p) (void)hipFree(p);
    delete s;
    return GLINT_OK;
}

glint_status glint_server_create(glint_ctx *ctx, const glint_server_config *cfg, glint_server **out)
{
    if (!ctx || !cfg || !out || cfg->max_running == 0 || cfg->arena_bytes < 4096) return GLINT_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    glint_server *s = new (std::nothrow) glint_server();
    if (!s) return GLINT_ERR_OUT_OF_MEMORY;
    s->ctx = ctx;
    s->cfg = *cfg;
    if (!s->cfg.min_hits) s->cfg.min_hits = 4096;
    glint_ctx_view(ctx, &s->view);
    if (!s->view.mlog) {
        delete s;
        return GLINT_ERR_STATE;   // the context needs a meeting log
    }
    if (hipSetDevice(s->view.device) != hipSuccess || hipMalloc((void **)&s->arena_d, cfg->arena_bytes) != hipSuccess) {
        (void)hipGetLastError();
        delete s;
        return GLINT_ERR_OUT_OF_MEMORY;
    }
    s->arena.init(cfg->arena_bytes);
    if (hipStreamCreateWithFlags(&s->index_stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&s->tick_done, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&s->sealed, hipEventDisableTiming) != hipSuccess) {
        glint_server_destroy(s);
        return GLINT_ERR_HIP;
    }
    MeetingIndex &ix = s->index;
    ix.block_ticks = cfg->index_block_ticks ? cfg->index_block_ticks : 64;
    if (ix.block_ticks != 0xFFFFFFFFu) {
        // blocks of at most 2^22 meetings; storage for twice the log's meetings
        // (two entries each) and two blocks more, so blocks outlive the log
        const uint64_t L = s->view.mlog_cap;
        ix.block_max = 2 * std::min<uint64_t>(L, 1ull << 22);
        ix.cap = 2 * L + 2 * ix.block_max;
        bool ok = hipMalloc((void **)&ix.keys, 8 * ix.cap) == hipSuccess &&
                  hipMalloc((void **)&ix.vals, 8 * ix.cap) == hipSuccess &&
                  hipMalloc((void **)&ix.skeys, 8 * ix.block_max) == hipSuccess &&
                  hipMalloc((void **)&ix.svals, 8 * ix.block_max) == hipSuccess;
        if (ok) {
            ok = rocprim::radix_sort_pairs(nullptr, ix.tmp_bytes, ix.skeys, ix.keys, ix.svals, ix.vals,
                                           (size_t)ix.block_max, 0, 62, s->index_stream) == hipSuccess &&
                 hipMalloc(&ix.tmp, std::max<size_t>(ix.tmp_bytes, 256)) == hipSuccess;
        }
        if (!ok) {
            (void)hipGetLastError();
            glint_server_destroy(s);
            return GLINT_ERR_OUT_OF_MEMORY;
        }
    }
    s->slots.resize(cfg->max_running);
    for (Slot_ &sl : s->slots) {
        if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&sl.ready, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&sl.snap, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess) {
            glint_server_destroy(s);
            return GLINT_ERR_HIP;
        }
    }
    *out = s;
    return GLINT_OK;
}

const char *glint_server_last_error(const glint_server *s) { return s ? s->err.c_str() : ""; }

glint_status glint_server_tick(glint_server *s, const double *d_full_xy, const uint32_t *d_pids, const double *d_xy,
                               uint32_t k, uint32_t tick)
{
    if (!s) return GLINT_ERR_INVALID_ARGUMENT;
    glint_ctx *c = s->ctx;
    glint_status st;
    if (d_full_xy)
        st = glint_set_positions_device(c, d_full_xy, s->view.num_objects, tick);
    else
        st = glint_ingest_updates(c, d_pids, d_xy, k, tick);
    if (st) return st;
    if ((st = glint_partition(c)) || (st = glint_refine(c)) || (st = glint_evaluate(c, nullptr, 0, nullptr)) ||
        (st = glint_update_schema(c)))
        return st;
    uint64_t total = 0, first = 0;
    if ((st = glint_meeting_log_info(c, &total, &first))) return st;   // waits for the tick
    const uint64_t before = s->log_after.empty() ? 0 : s->log_after.rbegin()->second;
    const double m = (double)(total - before);
    s->meetings_per_tick = s->have_tick ? 0.9 * s->meetings_per_tick + 0.1 * m : m;
    s->have_tick = true;
    s->log_after[tick] = total;
    while (s->log_after.size() > 1 && s->log_after.begin()->second < first) s->log_after.erase(s->log_after.begin());
    glint_ctx_view(c, &s->view);
    if ((st = index_tick(s, total, first, tick))) return st;
    // pending queries start on this tick's state; running ones advance (and
    // copy their hits to the host) in glint_server_poll, between ticks
    return admit(s);
}

glint_status glint_server_submit(glint_server *s, const uint32_t *seeds, uint32_t nseeds, uint32_t t_start,
                                 uint32_t t_end, uint32_t hops, uint64_t *id)
{
    if (!s || !seeds || !nseeds || !id || hops < 1 || hops > 2 || t_start > t_end) return GLINT_ERR_INVALID_ARGUMENT;
    if (!s->have_tick) {
        s->err = "no tick yet";
        return GLINT_ERR_STATE;
    }
    for (uint32_t i = 0; i < nseeds; i++)
        if (seeds[i] >= s->view.num_objects) return GLINT_ERR_INVALID_ARGUMENT;
    Query q;
    q.id = s->next_id++;
    q.seeds.assign(seeds, seeds + nseeds);
    q.t_start = t_start;
    q.t_end = t_end;
    q.hops = hops;
    q.T = s->log_after.rbegin()->first;
    q.hi_sub = s->log_after.rbegin()->second;
    // meetings ending at or after t_start are logged after tick t_start: the
    // scan starts there (or at the log's start); every entry is still checked
    // against the window, so an earlier start only costs time
    auto it = s->log_after.upper_bound(t_start);
    q.lo = it == s->log_after.begin() ? 0 : std::prev(it)->second;
    if (q.lo > q.hi_sub) q.lo = q.hi_sub;
    estimate_caps(s, q);
    q.t_submit = s->now_ms();
    s->queries[q.id] = q;
    s->pending.push_back(q.id);
    *id = q.id;
    return admit(s);
}

glint_status glint_server_index_info(glint_server *s, uint64_t *blocks, uint64_t *entries, uint64_t *sealed)
{
    if (!s) return GLINT_ERR_INVALID_ARGUMENT;
    if (blocks) *blocks = s->index.blocks.size();
    if (entries) *entries = s->index.entries;
    if (sealed) *sealed = s->index.sealed;
    return GLINT_OK;
}

glint_status glint_server_poll(glint_server *s, uint32_t *running, uint32_t *pending)
{
    if (!s) return GLINT_ERR_INVALID_ARGUMENT;
    glint_status st = advance(s);
    if (st) return st;
    if ((st = admit(s))) return st;
    uint32_t r = 0;
    for (Slot_ &sl : s->slots) r += sl.query != 0;
    if (running) *running = r;
    if (pending) *pending = (uint32_t)s->pending.size();
    return GLINT_OK;
}

glint_status glint_server_drain(glint_server *s)
{
    if (!s) return GLINT_ERR_INVALID_ARGUMENT;
    for (;;) {
        uint32_t r = 0, p = 0;
        glint_status st = glint_server_poll(s, &r, &p);
        if (st) return st;
        if (!r && !p) return GLINT_OK;
        for (Slot_ &sl : s->slots)
            if (sl.query) SV_TRY(s, hipStreamSynchronize(sl.stream));
        SV_TRY(s, hipStreamSynchronize(s->view.stream));
    }
}

glint_status glint_server_query_info(glint_server *s, uint64_t id, glint_query_info *out)
{
    if (!s || !out) return GLINT_ERR_INVALID_ARGUMENT;
    auto it = s->queries.find(id);
    if (it == s->queries.end()) return GLINT_ERR_INVALID_ARGUMENT;
    const Query &q = it->second;
    *out = glint_query_info{};
    out->id = q.id;
    out->state = q.state;
    out->hops = q.hops;
    out->hits1 = q.hits1.size();
    out->hits2 = q.hits2.size();
    out->footprint_bytes = q.bytes ? q.bytes : footprint(s, q);
    out->submit_tick = q.T;
    out->admitted_tick = q.admitted_tick;
    out->retries = q.retries;
    out->submit_ms = q.t_submit;
    out->admit_ms = q.t_admit;
    out->done_ms = q.t_done;
    return GLINT_OK;
}

glint_status glint_server_result(glint_server *s, uint64_t id, glint_hit *out, uint64_t cap, uint64_t *n_out)
{
    if (!s || !n_out) return GLINT_ERR_INVALID_ARGUMENT;
    auto it = s->queries.find(id);
    if (it == s->queries.end()) return GLINT_ERR_INVALID_ARGUMENT;
    Query &q = it->second;
    if (q.state == kFailed) return GLINT_ERR_OUTPUT_CAPACITY;
    if (q.state != kDone) return GLINT_ERR_STATE;
    *n_out = q.hits1.size() + q.hits2.size();
    if (!out) return GLINT_OK;
    if (cap < *n_out) return GLINT_ERR_OUTPUT_CAPACITY;
    std::sort(q.hits1.begin(), q.hits1.end(), hit_less);
    std::sort(q.hits2.begin(), q.hits2.end(), hit_less);
    std::copy(q.hits1.begin(), q.hits1.end(), out);
    std::copy(q.hits2.begin(), q.hits2.end(), out + q.hits1.size());
    s->queries.erase(it);
    return GLINT_OK;
}

}  // extern "C"
