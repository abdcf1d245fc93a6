// Contact-tracing path for MI355X (gfx950): partition -> refine -> evaluate
// -> schema update, behind the C ABI of include/glint.h.
//
// Reference behaviour restated here (paths relative to the reference repo):
//   descent and last_valid        src/tracing/filter.cpp:229-243, geometry.cu:234-249
//   lookup rule                   src/tracing/filter.cpp:130-152, geometry.cu:350-386
//   distance test                 src/geometry/geometry.h:40-48, refinement.cpp:30-33
//   meeting rule                  src/tracing/meeting.cpp:134-155
//   schema collect/split/merge    src/cuda/geometry.cu:1114-1234
//   initial build                 src/index/QTree.h:157-214
//
// Compiled with -ffp-contract=off: every FP64 expression rounds as written,
// so the descent, the box distances and the squared distances are bit-equal
// to the reference's.  sqrt(s) <= r is evaluated as s <= t64, where t64 is
// the largest double whose correctly rounded square root is <= r (computed
// on the host), which is the same predicate without a device sqrt.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "glint.h"
#include "glint_device.h"
#include "kv_engine.h"
#include "ctx_view.h"

using namespace glint;

namespace {

constexpr int kBlock = 256;
constexpr int kScanItems = 4;                       // per thread
constexpr int kScanTile = kBlock * kScanItems;      // elements per scan block
constexpr int kTile = 256;                          // points per LDS tile of the join (default)
constexpr int kStage = 432;                         // staged contacts per wave
constexpr int kFlushRound = 4;                      // join columns between stage checks
constexpr int kMeetStage = 128;                     // staged meetings per wave (table update)
constexpr uint32_t kChunk = 4096;                   // most contact-list entries reserved per atomic
constexpr int kRegion = 2048;                       // meeting-table slots per region (LDS-resident)
constexpr uint32_t kRegionFill = 1024;              // stored contacts per region the table shape aims at
constexpr uint32_t kMaxRegionBits = 21;             // at most 2^21 regions (coarse 2^10 x fine 2^11)
constexpr int kBucketThreads = 1024;                // workgroup of the two bucketing passes
constexpr int kBucketTile = 16384;                  // contacts per LDS tile of the coarse pass (128 KB)
constexpr int kFineTile = 12288;                    // contacts per LDS tile of the fine pass (96 KB)
constexpr uint32_t kMaxCoarse = 1024;               // coarse bins of pass 1
constexpr uint32_t kMaxFine = 2048;                 // fine bins of pass 2
constexpr int kBorderStage = 128;                   // staged border objects per wave (assign)
constexpr uint32_t kHeavyLeaf = 256;                // leaves above this size are joined first

// ------------------------------------------------------------ predicates

// Degrees per metre (util.h:38,57-66), each operation in the reference's
// order: PI is the reference's 3.14159265, 90 degrees is taken as 89.
constexpr double kRefPi = 3.14159265;
constexpr double kDegPerMeterLat = 360.0 / (40076.0 * 1000.0);

// Double-double arithmetic (Dekker, Knuth): error-free sums and products.
// Contraction is off in each so that no product is fused into a later sum,
// which would break the error-free property.
struct DD {
    double hi, lo;
};

__device__ inline DD dd_quick(double a, double b)   // |a| >= |b|
{
#pragma clang fp contract(off)
    const double s = a + b;
    return DD{s, b - (s - a)};
}

__device__ inline DD dd_two_sum(double a, double b)
{
#pragma clang fp contract(off)
    const double s = a + b, bb = s - a;
    return DD{s, (a - (s - bb)) + (b - bb)};
}

__device__ inline DD dd_add(DD a, DD b)
{
#pragma clang fp contract(off)
    DD s = dd_two_sum(a.hi, b.hi);
    const DD t = dd_two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = dd_quick(s.hi, s.lo);
    s.lo += t.lo;
    return dd_quick(s.hi, s.lo);
}

__device__ inline DD dd_mul(DD a, DD b)
{
#pragma clang fp contract(off)
    const double p = a.hi * b.hi;
    double e = fma(a.hi, b.hi, -p);
    e = fma(a.hi, b.lo, e);
    e = fma(a.lo, b.hi, e);
    return dd_quick(p, e);
}

// (-1)^k / (2k+1)! as double-double (head correctly rounded, tail the
// rounded remainder), k = 0..16.
constexpr DD kSinTaylor[17] = {
    {0x1.0000000000000p+0, 0x0.0p+0},
    {-0x1.5555555555555p-3, -0x1.5555555555555p-57},
    {0x1.1111111111111p-7, 0x1.1111111111111p-63},
    {-0x1.a01a01a01a01ap-13, -0x1.a01a01a01a01ap-73},
    {0x1.71de3a556c734p-19, -0x1.c154f8ddc6c00p-73},
    {-0x1.ae64567f544e4p-26, 0x1.c062e06d1f209p-80},
    {0x1.6124613a86d09p-33, 0x1.f28e0cc748ebep-87},
    {-0x1.ae7f3e733b81fp-41, -0x1.1d8656b0ee8cbp-97},
    {0x1.952c77030ad4ap-49, 0x1.ac981465ddc6cp-103},
    {-0x1.2f49b46814157p-57, -0x1.2650f61dbdcb4p-112},
    {0x1.71b8ef6dcf572p-66, -0x1.d043ae40c4647p-120},
    {-0x1.761b41316381ap-75, 0x1.3423c7d91404fp-130},
    {0x1.3f3ccdd165fa9p-84, -0x1.58ddadf344487p-139},
    {-0x1.d1ab1c2dccea3p-94, -0x1.054d0c78aea14p-149},
    {0x1.259f98b4358adp-103, 0x1.eaf8c39dd9bc5p-157},
    {-0x1.434d2e783f5bcp-113, -0x1.0b87b91be9affp-167},
    {0x1.3981254dd0d52p-123, -0x1.2b1f4c8015a2fp-177},
};

// sin(x) correctly rounded for |x| <= 1.6 (every angle a latitude inside
// (-90, 90) gives): the Taylor series to x^33 in double-double, whose
// truncation (x^35/35! < 2^-110 at pi/2) and rounding errors stay below
// 2^-100 relative, then one rounding to double.  Outside that range (not
// reached by positions inside a valid box) the device libm's sin.
__device__ inline double sin_correctly_rounded(double x)
{
    if (!(fabs(x) <= 1.6)) return sin(x);
    const DD y = dd_mul(DD{x, 0.0}, DD{x, 0.0});
    DD p = kSinTaylor[16];
#pragma unroll
    for (int k = 15; k >= 0; --k) p = dd_add(dd_mul(p, y), kSinTaylor[k]);
    return dd_mul(DD{x, 0.0}, p).hi;
}

// The device evaluates the sine correctly rounded.  The host libm the
// reference calls (glibc 2.35 here) returns the other neighbour of the
// correctly rounded value for about 0.1% of latitudes, those whose exact
// sine lies within about 0.005 ulp of a rounding midpoint; the factors then
// differ by one ulp.
__device__ inline double deg_per_meter_lon(double latitude)
{
    double absla = fabs(latitude);
    if (absla == 90) absla = 89;
    return 360.0 / (sin_correctly_rounded((90 - absla) * kRefPi / 180) * 40076.0 * 1000.0);
}

// The same expression on the host, with the host libm's sin as the
// reference computes it (x_buffer, filter.cpp:30).
inline double deg_per_meter_lon_host(double latitude)
{
    double absla = std::fabs(latitude);
    if (absla == 90) absla = 89;
    return 360.0 / (std::sin((90 - absla) * kRefPi / 180) * 40076.0 * 1000.0);
}

// box::distance(p, true) <= r (geometry.h:174-185): the clamped offsets in
// degrees divided by the degrees per metre, longitude at the point's
// latitude (f = deg_per_meter_lon(py)).
__device__ __forceinline__ bool box_within_geo(const Node &b, double px, double py, double t64, double f)
{
    if (px >= b.lx && px <= b.hx && py >= b.ly && py <= b.hy) return true;
    double dx = fabs(px - (b.lx + b.hx) / 2) - (b.hx - b.lx) / 2;
    double dy = fabs(py - (b.ly + b.hy) / 2) - (b.hy - b.ly) / 2;
    dx = dx < 0.0 ? 0.0 : dx;
    dy = dy < 0.0 ? 0.0 : dy;
    dy = dy / kDegPerMeterLat;
    dx = dx / f;
    return dx * dx + dy * dy <= t64;
}

__device__ __forceinline__ bool box_within(const Node &b, double px, double py, double t64)
{
    // box::distance (geometry.h:174-185) <= r, with contain() first.
    if (px >= b.lx && px <= b.hx && py >= b.ly && py <= b.hy) return true;
    double dx = fabs(px - (b.lx + b.hx) / 2) - (b.hx - b.lx) / 2;
    double dy = fabs(py - (b.ly + b.hy) / 2) - (b.hy - b.ly) / 2;
    dx = dx < 0.0 ? 0.0 : dx;
    dy = dy < 0.0 ? 0.0 : dy;
    double s = dx * dx + dy * dy;
    return s <= t64;
}

__device__ __forceinline__ bool point_within64(double x1, double y1, double x2, double y2, double t64)
{
    // Point::distance (geometry.h:40-48) <= r
    double dx = x1 - x2;
    double dy = y1 - y2;
    double s = dx * dx + dy * dy;
    return s <= t64;
}

// Workgroup barrier for LDS data only: global loads and stores stay in
// flight across it (__syncthreads also waits for every outstanding global
// access, which drains loads issued ahead for the next step).
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Orders LDS accesses of the lanes of one wave (staging buffers).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ------------------------------------------------------------ fused launches
//
// Every kernel of the tick sequence is a body run by one workgroup for one
// context: `blk` is the workgroup's index among that context's workgroups
// and their count.  k_<name> runs one context, its state in the kernel
// arguments; kf_<name> runs many contexts in one launch (independent
// scenarios of a batch): workgroups [start[i], start[i+1]) run member i,
// whose state is read from device memory.  The members' shares are the
// grids their own launches would have, so a fused launch does the same work
// as the members' launches one after another, without the gaps between them.

// The workgroup's place in a fused launch.
struct Blk {
    uint32_t id_, count_;
    __device__ __forceinline__ uint32_t id() const { return id_; }
    __device__ __forceinline__ uint32_t count() const { return count_; }
};

// The workgroup's place in a launch of its own: the hardware's indices, read
// where they are used.
struct ThisGrid {
    __device__ __forceinline__ uint32_t id() const { return blockIdx.x; }
    __device__ __forceinline__ uint32_t count() const { return gridDim.x; }
};

__device__ __forceinline__ ThisGrid this_blk() { return ThisGrid{}; }

constexpr uint32_t kMaxFused = 64;   // contexts per fused launch

struct Many {
    const DevState *__restrict__ ss;   // member states (device memory)
    uint32_t n;                        // members
    uint32_t start[kMaxFused + 1];     // first workgroup of each member; start[n] = the grid
};

struct Member {
    uint32_t i;
    Blk blk;
};

// The member running this workgroup: the last i with start[i] <= blockIdx.x
// (members with no workgroups are skipped over).
__device__ __forceinline__ Member member_of(const Many &m)
{
    const uint32_t b = blockIdx.x;
    uint32_t lo = 0, hi = m.n;   // start[lo] <= b < start[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (m.start[mid] <= b)
            lo = mid;
        else
            hi = mid;
    }
    return Member{lo, Blk{b - m.start[lo], m.start[lo + 1] - m.start[lo]}};
}

// ------------------------------------------------------------ small kernels

template <class P>
__device__ __forceinline__ void fill_body(uint32_t *p, uint32_t v, uint64_t n, const P blk)
{
    for (uint64_t i = blk.id() * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)blk.count() * blockDim.x)
        p[i] = v;
}

__global__ void k_fill_u32(uint32_t *p, uint32_t v, uint64_t n) { fill_body(p, v, n, this_blk()); }

__global__ void k_set_tick(TickParams *tp, const double *xy, TableGen tab_old, TableGen tab_cur, float4 *mbr_old,
                           float4 *mbr_cur, double *xy_prev, double *xy_save, uint4 *cand_prev, uint4 *cand,
                           double *sx_old, double *sy_old, double *sx, double *sy, uint32_t *sleaf_old,
                           uint32_t *sleaf, uint32_t tick, uint32_t prev_tick, uint32_t has_prev)
{
    tp->sleaf = sleaf;
    tp->sleaf_old = sleaf_old;
    tp->sx = sx;
    tp->sy = sy;
    tp->sx_old = sx_old;
    tp->sy_old = sy_old;
    tp->cand = cand;
    tp->cand_prev = cand_prev;
    tp->xy = xy;
    tp->tab_old = tab_old;
    tp->tab_cur = tab_cur;
    tp->mbr_old = mbr_old;
    tp->mbr_cur = mbr_cur;
    tp->xy_prev = xy_prev;
    tp->xy_save = xy_save;
    tp->hist_prev = xy_prev;
    tp->hist_save = xy_save;
    tp->tick = tick;
    tp->prev_tick = prev_tick;
    tp->has_prev = has_prev;
}

// Next tick of a device-resident trace: positions move one snapshot on and
// the generations swap (this tick's table and positions become the previous
// ones).  Every tick of a run starts with it, so a run of ticks is a fixed
// sequence of launches that can be captured in a graph and replayed.  Inside
// a run the positions at prev_tick are the trace's previous snapshot; only
// the run's last tick copies its positions into the history buffer, for the
// tick after the run.
__device__ __forceinline__ void advance_tick(TickParams *tp, uint64_t stride)
{
    uint32_t k = tp->trace_index + 1;
    tp->trace_index = k;
    if (tp->stamps && 2 * k < tp->nstamps) tp->stamps[2 * k] = wall_clock64();
    tp->xy = tp->trace_base + stride * k;
    tp->prev_tick = tp->tick;
    tp->tick = tp->next_tick;
    tp->next_tick = tp->next_tick + 1;
    tp->has_prev = tp->has_prev_next;
    tp->has_prev_next = 1;
    const TableGen t = tp->tab_old;
    tp->tab_old = tp->tab_cur;
    tp->tab_cur = t;
    float4 *m = tp->mbr_old;
    tp->mbr_old = tp->mbr_cur;
    tp->mbr_cur = m;
    double *h = tp->hist_prev;
    tp->hist_prev = tp->hist_save;
    tp->hist_save = h;
    tp->xy_prev = k == 0 ? tp->hist_prev : tp->trace_base + stride * (k - 1);
    tp->xy_save = k + 1 == tp->nticks ? tp->hist_save : nullptr;
    uint4 *cb = tp->cand_prev;
    tp->cand_prev = tp->cand;
    tp->cand = cb;
    uint32_t *sl = tp->sleaf_old;
    tp->sleaf_old = tp->sleaf;
    tp->sleaf = sl;
    double *q = tp->sx_old;
    tp->sx_old = tp->sx;
    tp->sx = q;
    q = tp->sy_old;
    tp->sy_old = tp->sy;
    tp->sy = q;
}

__global__ void k_advance_tick(TickParams *tp, uint64_t stride) { advance_tick(tp, stride); }

__global__ void kf_advance_tick(const Many m)
{
    const Member w = member_of(m);
    const DevState &s = m.ss[w.i];
    if (threadIdx.x == 0) advance_tick(s.tp, 2ull * s.n);
}

// State before the first tick of a run, one k_advance_tick short of it; the
// positions already point at the first snapshot for the initial build.
__global__ void k_begin_run(TickParams *tp, const double *base, TableGen tab_old, TableGen tab_cur, float4 *mbr_old,
                            float4 *mbr_cur, double *xy_prev, double *xy_save, uint4 *cand_prev, uint4 *cand,
                            double *sx_old, double *sy_old, double *sx, double *sy, uint32_t *sleaf_old,
                            uint32_t *sleaf, uint32_t tick, uint32_t next_tick, uint32_t has_prev_next,
                            unsigned long long *stamps, uint32_t nstamps, uint32_t nticks)
{
    tp->sleaf = sleaf;
    tp->sleaf_old = sleaf_old;
    tp->sx = sx;
    tp->sy = sy;
    tp->sx_old = sx_old;
    tp->sy_old = sy_old;
    tp->nticks = nticks;
    tp->cand = cand;
    tp->cand_prev = cand_prev;
    tp->trace_base = base;
    tp->trace_index = 0xFFFFFFFFu;
    tp->xy = base;
    tp->tab_old = tab_old;
    tp->tab_cur = tab_cur;
    tp->mbr_old = mbr_old;
    tp->mbr_cur = mbr_cur;
    tp->xy_prev = xy_prev;
    tp->xy_save = xy_save;
    tp->hist_prev = xy_prev;
    tp->hist_save = xy_save;
    tp->tick = tick;
    tp->next_tick = next_tick;
    tp->has_prev_next = has_prev_next;
    tp->stamps = stamps;
    tp->nstamps = nstamps;
}

// Tick k of a run is stamped at its start (k_advance_tick, entry 2k) and at
// its end (entry 2k+1), so a tick's latency excludes time spent waiting for
// other work between ticks.
__device__ __forceinline__ void stamp_end(TickParams *tp)
{
    uint32_t k = tp->trace_index;
    if (tp->stamps && 2 * k + 1 < tp->nstamps) tp->stamps[2 * k + 1] = wall_clock64();
}

__global__ void k_stamp(TickParams *tp) { stamp_end(tp); }

__global__ void kf_stamp(const Many m)
{
    const Member w = member_of(m);
    if (threadIdx.x == 0) stamp_end(m.ss[w.i].tp);
}

template <class P>
__device__ __forceinline__ void reset_tick_body(const DevState s, const P blk)
{
    Counters *c = s.ctr;
    c->n_units = 0;
    c->n_contacts = 0;
    c->n_contacts_border = 0;
    c->n_dup = 0;
    c->n_stored = 0;
    c->n_meet = 0;
    c->n_rechecks = 0;
    c->n_clist = 0;
    c->work_border = 0;
    c->n_split = 0;
    c->n_merge = 0;
    c->n_border = 0;
    c->n_leaf_list = 0;
    c->n_heavy = 0;
    c->n_leaves = 0;
    c->max_leaf = 0;
    c->n_drop_leaves = 0;
    c->n_drop_objs = 0;
    c->n_tests = 0;
    c->n_evals = 0;
    c->work_inside = 0;
    c->g_heavy = 0;
}

__global__ void k_reset_tick(DevState s) { reset_tick_body(s, this_blk()); }
__global__ void kf_reset_tick(const Many m)
{
    const Member w = member_of(m);
    reset_tick_body(m.ss[w.i], w.blk);
}

// ------------------------------------------------------------ tree build

__global__ void k_init_tree(DevState s)
{
    uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t k = gid; k < s.node_cap; k += stride) {
        // free stack holds nodes 1..node_cap-1; pops return 1, 2, 3, ...
        if (k + 1 < s.node_cap) s.free_stack[k] = (int32_t)(s.node_cap - 1 - k);
        NodeMeta m;
        m.type = kInvalid;
        m.level = 0;
        m.overflow = 0;
        m.underflow = 0;
        s.meta[k] = m;
    }
    if (gid == 0) {
        Node root;
        root.lx = s.lx;
        root.ly = s.ly;
        root.hx = s.hx;
        root.hy = s.hy;
        root.mx = (s.hx + s.lx) / 2;   // QTree.h:103-104
        root.my = (s.hy + s.ly) / 2;
        root.child[0] = kLeafTag;
        root.child[1] = root.child[2] = root.child[3] = 0;
        s.nodes[0] = root;
        s.parent[0] = kNone;
        NodeMeta m;
        m.type = kLeaf;
        m.level = 0;
        m.overflow = 0;
        m.underflow = 0;
        s.meta[0] = m;
        s.ctr->free_top = (int32_t)s.node_cap - 1;
        s.ctr->error = 0;
    }
}

// Node fields read as 16-byte vectors.  A descent picks child c from the
// four ids with selects: indexing a private array by a runtime value would
// place it in scratch memory.
__device__ __forceinline__ double2 node_mid(const Node *nodes, uint32_t n)
{
    return reinterpret_cast<const double2 *>(nodes + n)[0];
}
__device__ __forceinline__ double2 node_high(const Node *nodes, uint32_t n)
{
    return reinterpret_cast<const double2 *>(nodes + n)[2];
}
__device__ __forceinline__ uint4 node_children(const Node *nodes, uint32_t n)
{
    return reinterpret_cast<const uint4 *>(nodes + n)[3];
}
__device__ __forceinline__ uint32_t pick_child(uint4 ch, int c)
{
    return c < 2 ? (c == 0 ? ch.x : ch.y) : (c == 2 ? ch.z : ch.w);
}

__device__ __forceinline__ uint32_t descend_to_leaf(const Node *nodes, double x, double y)
{
    uint32_t cur = 0;
    double2 mid = node_mid(nodes, 0);
    uint4 ch = node_children(nodes, 0);
    while (ch.x != kLeafTag) {
        cur = pick_child(ch, (y > mid.y) * 2 + (x > mid.x));
        mid = node_mid(nodes, cur);
        ch = node_children(nodes, cur);
    }
    return cur;
}

// One count per lane added to p[key] (kNone: nothing), one atomic per run of
// equal keys in the wave.  Every lane of the wave must call it.
__device__ __forceinline__ void wave_count(uint32_t *p, uint32_t key)
{
    bool need = key != kNone;
    for (;;) {
        const unsigned long long pend = __ballot(need);
        if (!pend) break;
        const uint32_t leader = __ffsll((long long)pend) - 1;
        const uint32_t K = __shfl(key, leader);
        const bool mine = need && key == K;
        const unsigned long long mm = __ballot(mine);
        if (lane_id() == leader) atomicAdd(&p[K], (uint32_t)__popcll(mm));
        if (mine) need = false;
    }
}

// Leaf populations of the tree built so far.  The first levels put most
// objects in a few leaves, so the counts are aggregated per wave.
__global__ void k_build_count(DevState s)
{
    const double *xy = s.tp->xy;
    for (uint32_t base = blockIdx.x * blockDim.x; base < s.n; base += gridDim.x * blockDim.x) {
        const uint32_t i = base + threadIdx.x;
        uint32_t leaf = kNone;
        if (i < s.n) {
            const double x = xy[2 * (size_t)i], y = xy[2 * (size_t)i + 1];
            if (x >= s.lx && x <= s.hx && y >= s.ly && y <= s.hy)
                leaf = descend_to_leaf(s.nodes, x, y);
            else
                atomicOr(&s.ctr->error, kErrBadInput);
        }
        wave_count(s.count, leaf);
    }
}

// QTNode::split (QTree.h:157-160): count >= cap && width/2 > r.
__global__ void k_build_collect(DevState s)
{
    for (uint32_t n = blockIdx.x * blockDim.x + threadIdx.x; n < s.node_cap; n += gridDim.x * blockDim.x) {
        bool want = false;
        if (s.meta[n].type == kLeaf) {
            Node nd = s.nodes[n];
            // QTNode::split (QTree.h:158): mbr.width(geography) / 2 > r
            const double wd = s.geography ? (nd.hx - nd.lx) / deg_per_meter_lon(nd.ly) : nd.hx - nd.lx;
            want = s.count[n] >= s.cap && wd / 2 > s.r;
        }
        uint32_t idx = wave_append_u32(&s.ctr->n_split, want);
        if (want) s.split_list[idx] = n;
    }
}

// Split the listed leaves one level.  build_boxes selects the child boxes of
// the initial build, (lo, mid) and (mid, hi) (QTree.h:163-166); otherwise the
// elastic update's lo + half and mid + half (geometry.cu:1151-1156).
template <class P>
__device__ __forceinline__ void split_body(const DevState s, const P blk, int build_boxes)
{
    uint32_t ns = s.ctr->n_split;
    for (uint32_t k = blk.id() * blockDim.x + threadIdx.x; k < ns; k += blk.count() * blockDim.x) {
        uint32_t node = s.split_list[k];
        int32_t top = atomicSub(&s.ctr->free_top, 4);
        if (top < 4) {
            atomicAdd(&s.ctr->free_top, 4);   // give the entries back: the pool never goes negative
            atomicOr(&s.ctr->error, kErrNodeCap);
            continue;
        }
        Node p = s.nodes[node];
        NodeMeta pm = s.meta[node];
        double xhalf = p.mx - p.lx;
        double yhalf = p.my - p.ly;
        Node np = p;
        for (int i = 0; i < 4; i++) {
            uint32_t c = (uint32_t)s.free_stack[top - 1 - i];
            Node ch;
            if (build_boxes) {
                ch.lx = (i & 1) ? p.mx : p.lx;
                ch.hx = (i & 1) ? p.hx : p.mx;
                ch.ly = (i & 2) ? p.my : p.ly;
                ch.hy = (i & 2) ? p.hy : p.my;
            } else {
                ch.lx = p.lx + (double)(i % 2 == 1) * xhalf;
                ch.ly = p.ly + (double)(i / 2 == 1) * yhalf;
                ch.hx = p.mx + (double)(i % 2 == 1) * xhalf;
                ch.hy = p.my + (double)(i / 2 == 1) * yhalf;
            }
            ch.mx = (ch.lx + ch.hx) / 2;
            ch.my = (ch.ly + ch.hy) / 2;
            ch.child[0] = kLeafTag;
            ch.child[1] = ch.child[2] = ch.child[3] = 0;
            s.nodes[c] = ch;
            s.parent[c] = node;
            NodeMeta m;
            m.type = kLeaf;
            m.level = pm.level + 1;
            m.overflow = 0;
            m.underflow = 0;
            s.meta[c] = m;
            np.child[i] = c;
        }
        s.nodes[node] = np;
        pm.type = kBranch;
        s.meta[node] = pm;
    }
}

__global__ void k_split(DevState s, int build_boxes) { split_body(s, this_blk(), build_boxes); }
__global__ void kf_split(const Many m, int build_boxes)
{
    const Member w = member_of(m);
    split_body(m.ss[w.i], w.blk, build_boxes);
}

// geometry.cu:1166-1190: the four leaf children return to the pool.
template <class P>
__device__ __forceinline__ void merge_body(const DevState s, const P blk)
{
    uint32_t nm = s.ctr->n_merge;
    for (uint32_t k = blk.id() * blockDim.x + threadIdx.x; k < nm; k += blk.count() * blockDim.x) {
        uint32_t node = s.merge_list[k];
        Node p = s.nodes[node];
        int32_t top = atomicAdd(&s.ctr->free_top, 4);
        if (top < 0 || top + 4 > (int32_t)s.node_cap) {
            atomicOr(&s.ctr->error, kErrNodeCap);
            continue;
        }
        for (int i = 0; i < 4; i++) {
            uint32_t c = p.child[i];
            s.meta[c].type = kInvalid;
            s.free_stack[top + i] = (int32_t)c;
        }
        p.child[0] = kLeafTag;
        p.child[1] = p.child[2] = p.child[3] = 0;
        s.nodes[node] = p;
        NodeMeta pm = s.meta[node];
        pm.type = kLeaf;
        pm.overflow = 0;
        s.meta[node] = pm;
    }
}

__global__ void k_merge(DevState s) { merge_body(s, this_blk()); }
__global__ void kf_merge(const Many m)
{
    const Member w = member_of(m);
    merge_body(m.ss[w.i], w.blk);
}

// geometry.cu:1193-1234: decisions from this tick's counts, before any change.
template <class P>
__device__ __forceinline__ void collect_body(const DevState s, const P blk)
{
    for (uint32_t n = blk.id() * blockDim.x + threadIdx.x; n < s.node_cap; n += blk.count() * blockDim.x) {
        NodeMeta m = s.meta[n];
        bool split = false, merge = false;
        if (m.type == kLeaf) {
            Node nd = s.nodes[n];
            // height(mbr) > 2r (geometry.cu:1200), height in metres (geometry.cu:30) with geography
            const double ht = s.geography ? (nd.hy - nd.ly) / kDegPerMeterLat : nd.hy - nd.ly;
            if (ht > 2 * s.r && s.count[n] > s.cap) {
                if (++m.overflow >= s.delay) {
                    split = true;
                    m.overflow = 0;
                }
            } else {
                m.overflow = 0;
            }
            s.meta[n] = m;
        } else if (m.type == kBranch) {
            Node nd = s.nodes[n];
            int leafchild = 0;
            int ncounter = 0;
            for (int i = 0; i < 4; i++) {
                uint32_t c = nd.child[i];
                if (s.meta[c].type == kLeaf) {
                    leafchild++;
                    ncounter += (int)s.count[c];
                }
            }
            if (leafchild == 4 && ncounter < (int)s.cap) {
                if (++m.underflow >= s.delay) {
                    merge = true;
                    m.underflow = 0;
                }
            } else {
                m.underflow = 0;
            }
            s.meta[n] = m;
        }
        uint32_t is = wave_append_u32(&s.ctr->n_split, split);
        if (split) s.split_list[is] = n;
        uint32_t im = wave_append_u32(&s.ctr->n_merge, merge);
        if (merge) s.merge_list[im] = n;
    }
}

__global__ void k_collect(DevState s) { collect_body(s, this_blk()); }
__global__ void kf_collect(const Many m)
{
    const Member w = member_of(m);
    collect_body(m.ss[w.i], w.blk);
}

// End of a tick: schema change counts and the running totals.
template <class P>
__device__ __forceinline__ void schema_done_body(const DevState s, const P blk)
{
    Counters *c = s.ctr;
    c->last_splits = c->n_split;
    c->last_merges = c->n_merge;
    c->acc_ticks += 1;
    c->acc_contacts += c->n_contacts;
    c->acc_contacts_border += c->n_contacts_border;
    const unsigned long long distinct = c->n_contacts - c->n_dup;
    c->acc_distinct += distinct;
    c->acc_meet += c->n_meet;
    c->acc_units += c->n_units;
    c->acc_border += c->n_border;
    c->acc_rechecks += c->n_rechecks;
    if (c->n_contacts > c->max_contacts) c->max_contacts = c->n_contacts;
    if (distinct > c->max_distinct) c->max_distinct = distinct;
    c->acc_stored += c->n_stored;
    if (c->n_stored > c->max_stored) c->max_stored = c->n_stored;
    if (c->n_meet > c->max_meet) c->max_meet = c->n_meet;
    c->acc_drop_leaves += c->n_drop_leaves;
    c->acc_drop_objs += c->n_drop_objs;
    c->acc_tests += c->n_tests;
    c->acc_evals += c->n_evals;
    if (c->n_drop_objs > c->max_drop_objs) c->max_drop_objs = c->n_drop_objs;
    if (c->max_leaf > c->max_leaf_run) c->max_leaf_run = c->max_leaf;
    if (s.mlog) c->n_mlog += c->n_meet < s.meet_cap ? c->n_meet : s.meet_cap;
}

__global__ void k_schema_done(DevState s) { schema_done_body(s, this_blk()); }
__global__ void kf_schema_done(const Many m)
{
    const Member w = member_of(m);
    schema_done_body(m.ss[w.i], w.blk);
}

// The tick's meetings appended to the meeting log ring (k_schema_done
// advances its total afterwards).
template <class P>
__device__ __forceinline__ void log_append_body(const DevState s, const P blk)
{
    const unsigned long long n = min(s.ctr->n_meet, s.meet_cap), base = s.ctr->n_mlog;
    for (unsigned long long i = blk.id() * (unsigned long long)kBlock + threadIdx.x; i < n;
         i += (unsigned long long)blk.count() * kBlock)
        s.mlog[(base + i) % s.mlog_cap] = s.meet[i];
}

__global__ __launch_bounds__(kBlock) void k_log_append(DevState s) { log_append_body(s, this_blk()); }
__global__ __launch_bounds__(kBlock) void kf_log_append(const Many m)
{
    const Member w = member_of(m);
    log_append_body(m.ss[w.i], w.blk);
}

// ------------------------------------------------------------ grid rebuild
//
// GLINT_SCHEMA_GRID_REBUILD: the reference's GPU path without a dynamic
// schema builds the tree from scratch every tick (geometry.cu:1629-1645):
// points are counted in a 16384 x 16384 cell grid (cuda_build_qtree,
// geometry.cu:1270-1279) and merged bottom-up (cuda_merge_qtree,
// geometry.cu:1295-1393), a block of cells becoming a branch iff it holds
// more than grid_capacity points, a branch cell getting four leaf children.
// The branch set is closed upwards, so the same tree is built here
// top-down from block counts, without the reference's two 1 GiB cell arrays:
//   - k_grid_cells: each object's cell; counts of the 1024^2 coarse blocks
//     (depth 10, 16x16 cells each, 4 MiB);
//   - k_grid_pyramid, k_grid_top: block counts of depths 9..0 from them, and
//     a 16x16 fine histogram for every coarse block over capacity (only those
//     can hold branches below depth 10);
//   - k_grid_fine: the fine histograms;
//   - k_grid_level(d), d = 0..14: every branch block of depth d writes its
//     node (grid-formula box) and four children: branch blocks, listed for
//     depth d+1, or leaves (the xhalf box formula, geometry.cu:1311-1326).
// A fixed launch sequence, so a tick with the rebuild can be captured in a
// graph.
constexpr uint32_t kGridBits = 14;
constexpr uint32_t kGridDim = 1u << kGridBits;             // one_dim (geometry.cu:1266)
constexpr uint32_t kCoarseDepth = 10;
constexpr uint32_t kCoarseDim = 1u << kCoarseDepth;
constexpr uint32_t kFineDim = kGridDim / kCoarseDim;       // cells per coarse block side
constexpr uint32_t kFineCells = kFineDim * kFineDim;

// Per-context arrays cleared by the tick sequence.
enum FillArray : int { kFillCount, kFillUcount, kFillMeta, kFillCoarse };

template <int F>
__global__ void kf_fill(const Many m)
{
    const Member w = member_of(m);
    const DevState &s = m.ss[w.i];
    if (F == kFillCount) fill_body(s.count, 0u, s.node_cap, w.blk);
    if (F == kFillUcount) fill_body(s.ucount, 0u, s.node_cap, w.blk);
    if (F == kFillMeta) fill_body(reinterpret_cast<uint32_t *>(s.meta), 0u, 4ull * s.node_cap, w.blk);
    if (F == kFillCoarse) fill_body(s.g10, 0u, kCoarseDim * kCoarseDim, w.blk);
}

__host__ __device__ constexpr uint32_t pyr_off(uint32_t d) { return ((1u << (2 * d)) - 1u) / 3u; }

// Object k of the pass: the leaf order of the tick before when there is one.
__device__ __forceinline__ uint32_t grid_object(const DevState &s, uint32_t k, bool ordered)
{
    if (!ordered) return k;
    const uint32_t i = s.spid[k];
    return i < s.n ? i : k;
}

// The cell as the reference computes it (geometry.cu:1276-1278): x + y*16384
// in 32-bit arithmetic, so x = 16384 (a point on the high x edge) is the
// first cell of the next row.  y = 16384 has no cell (the reference writes
// past its histogram there): flagged.
template <class P>
__device__ __forceinline__ void grid_cells_body(const DevState s, const P blk)
{
    const double *xy = s.tp->xy;
    const bool ordered = s.ctr->sorted_valid != 0;
    for (uint32_t base = blk.id() * kBlock; base < s.n; base += blk.count() * kBlock) {
        const uint32_t k = base + threadIdx.x;
        uint32_t key = kNone;
        if (k < s.n) {
            const uint32_t i = grid_object(s, k, ordered);
            const double x = xy[2 * (size_t)i], y = xy[2 * (size_t)i + 1];
            uint32_t lin = kNone;
            if (x >= s.lx && x <= s.hx && y >= s.ly && y <= s.hy) {
                const uint32_t cx = (uint32_t)((x - s.lx) / (s.hx - s.lx) * (double)kGridDim);
                const uint32_t cy = (uint32_t)((y - s.ly) / (s.hy - s.ly) * (double)kGridDim);
                lin = cx + cy * kGridDim;
                if (lin >= kGridDim * kGridDim) {
                    atomicOr(&s.ctr->error, kErrGridEdge);
                    lin = kNone;
                }
            } else {
                atomicOr(&s.ctr->error, kErrBadInput);
            }
            s.gcell[i] = lin;
            if (lin != kNone) {
                const uint32_t gx = lin & (kGridDim - 1), gy = lin >> kGridBits;
                key = (gy / kFineDim) * kCoarseDim + gx / kFineDim;
            }
        }
        wave_count(s.g10, key);
    }
}

__global__ __launch_bounds__(kBlock) void k_grid_cells(DevState s) { grid_cells_body(s, this_blk()); }
__global__ __launch_bounds__(kBlock) void kf_grid_cells(const Many m)
{
    const Member w = member_of(m);
    grid_cells_body(m.ss[w.i], w.blk);
}

// One workgroup per 32x32 patch of coarse blocks (1024 workgroups of 256):
// depths 9..5 of the pyramid, and a fine histogram for each coarse block
// over capacity.
template <class P>
__device__ __forceinline__ void grid_pyramid_body(const DevState s, const P blk)
{
    __shared__ uint32_t lv[256];
    const uint32_t t = threadIdx.x;
    const uint32_t px = blk.id() & 31, py = blk.id() >> 5;
    const uint32_t gx = px * 16 + (t & 15), gy = py * 16 + (t >> 4);   // depth-9 block
    uint32_t sum = 0;
    for (uint32_t i = 0; i < 4; i++) {
        const uint32_t b = (2 * gy + (i >> 1)) * kCoarseDim + 2 * gx + (i & 1);
        const uint32_t v = s.g10[b];
        sum += v;
        const bool heavy = v > s.cap;
        const uint32_t h = wave_append_u32(&s.ctr->g_heavy, heavy);
        uint32_t id = kNone;
        if (heavy) {
            if (h < s.heavy_cap)
                id = h;
            else
                atomicOr(&s.ctr->error, kErrNodeCap);   // not reached: blocks over capacity are disjoint
        }
        s.gheavy[b] = id;
    }
    lv[t] = sum;
    s.gpyr[pyr_off(9) + gy * (1u << 9) + gx] = sum;
    uint32_t side = 16;
    for (uint32_t d = 8; d >= 5; d--) {
        const uint32_t ns = side / 2;
        const uint32_t x = t % ns, y = t / ns;
        uint32_t v = 0;
        __syncthreads();
        if (t < ns * ns)
            v = lv[2 * y * side + 2 * x] + lv[2 * y * side + 2 * x + 1] + lv[(2 * y + 1) * side + 2 * x] +
                lv[(2 * y + 1) * side + 2 * x + 1];
        __syncthreads();
        if (t < ns * ns) {
            lv[t] = v;
            s.gpyr[pyr_off(d) + (py * ns + y) * (1u << d) + px * ns + x] = v;
        }
        side = ns;
    }
}

__global__ __launch_bounds__(kBlock) void k_grid_pyramid(DevState s) { grid_pyramid_body(s, this_blk()); }
__global__ __launch_bounds__(kBlock) void kf_grid_pyramid(const Many m)
{
    const Member w = member_of(m);
    grid_pyramid_body(m.ss[w.i], w.blk);
}

// Depths 4..0 from the 32x32 blocks of depth 5, and the root: a branch when
// the points exceed capacity (the list of depth 0), else one leaf.
template <class P>
__device__ __forceinline__ void grid_top_body(const DevState s, const P blk)
{
    __shared__ uint32_t lv[1024];
    const uint32_t t = threadIdx.x;
    lv[t] = s.gpyr[pyr_off(5) + t];
    uint32_t side = 32;
    for (int d = 4; d >= 0; d--) {
        const uint32_t ns = side / 2;
        const uint32_t x = t % ns, y = t / ns;
        uint32_t v = 0;
        __syncthreads();
        if (t < ns * ns)
            v = lv[2 * y * side + 2 * x] + lv[2 * y * side + 2 * x + 1] + lv[(2 * y + 1) * side + 2 * x] +
                lv[(2 * y + 1) * side + 2 * x + 1];
        __syncthreads();
        if (t < ns * ns) {
            lv[t] = v;
            s.gpyr[pyr_off((uint32_t)d) + y * ns + x] = v;
        }
        side = ns;
    }
    if (t == 0) {
        Counters *c = s.ctr;
        c->g_nodes = 1;
        for (int d = 0; d < 16; d++) c->g_list[d] = 0;
        if (lv[0] > s.cap) {
            s.glist[0][0] = make_uint2(0u, 0u);
            s.parent[0] = kNone;
            c->g_list[0] = 1;
        } else {
            // no branch anywhere (the reference writes no node then): the root block as one leaf
            Node r;
            r.lx = s.lx + 0.0 * (s.hx - s.lx) / 1.0;
            r.ly = s.ly + 0.0 * (s.hy - s.ly) / 1.0;
            r.hx = s.lx + 1.0 * (s.hx - s.lx) / 1.0;
            r.hy = s.ly + 1.0 * (s.hy - s.ly) / 1.0;
            r.mx = (r.lx + r.hx) / 2;
            r.my = (r.ly + r.hy) / 2;
            r.child[0] = kLeafTag;
            r.child[1] = r.child[2] = r.child[3] = 0;
            s.nodes[0] = r;
            s.parent[0] = kNone;
            NodeMeta m;
            m.type = kLeaf;
            m.level = 0;
            m.overflow = 0;
            m.underflow = 0;
            s.meta[0] = m;
        }
    }
}

__global__ __launch_bounds__(1024) void k_grid_top(DevState s) { grid_top_body(s, this_blk()); }
__global__ __launch_bounds__(1024) void kf_grid_top(const Many m)
{
    const Member w = member_of(m);
    grid_top_body(m.ss[w.i], w.blk);
}

template <class P>
__device__ __forceinline__ void grid_fine_zero_body(const DevState s, const P blk)
{
    const uint64_t n = (uint64_t)min(s.ctr->g_heavy, s.heavy_cap) * kFineCells;
    for (uint64_t i = blk.id() * (uint64_t)kBlock + threadIdx.x; i < n; i += (uint64_t)blk.count() * kBlock)
        s.gfine[i] = 0;
}

__global__ __launch_bounds__(kBlock) void k_grid_fine_zero(DevState s) { grid_fine_zero_body(s, this_blk()); }
__global__ __launch_bounds__(kBlock) void kf_grid_fine_zero(const Many m)
{
    const Member w = member_of(m);
    grid_fine_zero_body(m.ss[w.i], w.blk);
}

template <class P>
__device__ __forceinline__ void grid_fine_body(const DevState s, const P blk)
{
    const bool ordered = s.ctr->sorted_valid != 0;
    for (uint32_t base = blk.id() * kBlock; base < s.n; base += blk.count() * kBlock) {
        const uint32_t k = base + threadIdx.x;
        uint32_t key = kNone;
        if (k < s.n) {
            const uint32_t lin = s.gcell[grid_object(s, k, ordered)];
            if (lin != kNone) {
                const uint32_t gx = lin & (kGridDim - 1), gy = lin >> kGridBits;
                const uint32_t h = s.gheavy[(gy / kFineDim) * kCoarseDim + gx / kFineDim];
                if (h != kNone) key = h * kFineCells + (gy % kFineDim) * kFineDim + gx % kFineDim;
            }
        }
        wave_count(s.gfine, key);
    }
}

__global__ __launch_bounds__(kBlock) void k_grid_fine(DevState s) { grid_fine_body(s, this_blk()); }
__global__ __launch_bounds__(kBlock) void kf_grid_fine(const Many m)
{
    const Member w = member_of(m);
    grid_fine_body(m.ss[w.i], w.blk);
}

// Points in block (bx, by) of depth e (2^e blocks per side).  Below the
// coarse depth the block lies in a coarse block over capacity (its parent
// is a branch), whose fine histogram holds its cells.
__device__ uint32_t grid_count(const DevState &s, uint32_t e, uint32_t bx, uint32_t by)
{
    if (e < kCoarseDepth) return s.gpyr[pyr_off(e) + by * (1u << e) + bx];
    if (e == kCoarseDepth) return s.g10[by * kCoarseDim + bx];
    const uint32_t sh = e - kCoarseDepth;
    const uint32_t h = s.gheavy[(by >> sh) * kCoarseDim + (bx >> sh)];
    if (h == kNone) return 0;
    const uint32_t sz = 1u << (kGridBits - e), m = (1u << sh) - 1u;
    const uint32_t ox = (bx & m) * sz, oy = (by & m) * sz;
    const uint32_t *f = s.gfine + (size_t)h * kFineCells;
    uint32_t sum = 0;
    for (uint32_t yy = 0; yy < sz; yy++)
        for (uint32_t xx = 0; xx < sz; xx++) sum += f[(oy + yy) * kFineDim + ox + xx];
    return sum;
}

// The branch blocks of depth d: node boxes by the grid formula
// (geometry.cu:1304-1307,1335-1338), four children each, from one
// wave-aggregated reservation; a child over capacity is listed for depth
// d+1, any other is a leaf.  A failed reservation leaves the block a leaf
// (and flags the error), so the tree stays well formed.
template <class P>
__device__ __forceinline__ void grid_level_body(const DevState s, const P blk, uint32_t d)
{
    const uint32_t nb = min(s.ctr->g_list[d], s.glist_cap);
    // selects, not an index into the state (which would place it in scratch memory)
    const uint2 *in = (d & 1) ? s.glist[1] : s.glist[0];
    uint2 *out = (d & 1) ? s.glist[0] : s.glist[1];
    const double xdim = (double)(1u << d);
    const double wx = s.hx - s.lx, wy = s.hy - s.ly;
    for (uint32_t base = blk.id() * kBlock; base < nb; base += blk.count() * kBlock) {
        const uint32_t j = base + threadIdx.x;
        const bool act = j < nb;
        const unsigned long long mask = __ballot(act);
        const uint32_t leader = __ffsll((long long)mask) - 1;
        uint32_t c0 = 0;
        if (lane_id() == leader) c0 = atomicAdd(&s.ctr->g_nodes, 4u * (uint32_t)__popcll(mask));
        c0 = __shfl(c0, leader) + 4u * (uint32_t)__popcll(mask & ((1ull << lane_id()) - 1ull));
        if (!act) continue;
        const uint2 e = in[j];
        const uint32_t node = e.x, bx = e.y & 0xFFFFu, by = e.y >> 16;
        Node b;
        b.lx = s.lx + (double)bx * wx / xdim;
        b.ly = s.ly + (double)by * wy / xdim;
        b.hx = s.lx + (double)(bx + 1) * wx / xdim;
        b.hy = s.ly + (double)(by + 1) * wy / xdim;
        b.mx = (b.lx + b.hx) / 2;
        b.my = (b.ly + b.hy) / 2;
        NodeMeta bm;
        bm.type = kBranch;
        bm.level = (int32_t)d;
        bm.overflow = 0;
        bm.underflow = 0;
        if ((uint64_t)c0 + 4 > s.node_cap) {
            atomicOr(&s.ctr->error, kErrNodeCap);
            b.child[0] = kLeafTag;
            b.child[1] = b.child[2] = b.child[3] = 0;
            bm.type = kLeaf;
        } else {
            const double xhalf = b.mx - b.lx, yhalf = b.my - b.ly;
            for (uint32_t i = 0; i < 4; i++) {
                const uint32_t c = c0 + i;
                b.child[i] = c;
                s.parent[c] = node;
                const uint32_t cbx = 2 * bx + (i & 1), cby = 2 * by + (i >> 1);
                if (d < kGridBits && grid_count(s, d + 1, cbx, cby) > s.cap) {
                    const uint32_t slot = atomicAdd(&s.ctr->g_list[d + 1], 1u);
                    if (slot < s.glist_cap) {
                        out[slot] = make_uint2(c, cbx | (cby << 16));
                        continue;
                    }
                    atomicOr(&s.ctr->error, kErrNodeCap);   // not reached: branches of a depth are disjoint
                }
                Node l;
                l.lx = b.lx + (double)(i % 2 == 1) * xhalf;
                l.ly = b.ly + (double)(i / 2 == 1) * yhalf;
                l.hx = b.mx + (double)(i % 2 == 1) * xhalf;
                l.hy = b.my + (double)(i / 2 == 1) * yhalf;
                l.mx = (l.lx + l.hx) / 2;
                l.my = (l.ly + l.hy) / 2;
                l.child[0] = kLeafTag;
                l.child[1] = l.child[2] = l.child[3] = 0;
                s.nodes[c] = l;
                NodeMeta lm;
                lm.type = kLeaf;
                lm.level = (int32_t)d + 1;
                lm.overflow = 0;
                lm.underflow = 0;
                s.meta[c] = lm;
            }
        }
        s.nodes[node] = b;
        s.meta[node] = bm;
    }
}

__global__ __launch_bounds__(kBlock) void k_grid_level(DevState s, uint32_t d) { grid_level_body(s, this_blk(), d); }
__global__ __launch_bounds__(kBlock) void kf_grid_level(const Many m, uint32_t d)
{
    const Member w = member_of(m);
    grid_level_body(m.ss[w.i], w.blk, d);
}

// ------------------------------------------------------------ partition

// The region of every node of the tree: the points the descent (child =
// (y > my)*2 + (x > mx) at every branch) sends into it, the intersection of
// its ancestors' half-planes, as (x.lo, x.hi] x (y.lo, y.hi] (infinite where
// no ancestor bounds it).
template <class P>
__device__ __forceinline__ void regions_body(const DevState s, const P blk)
{
    for (uint32_t n = blk.id() * blockDim.x + threadIdx.x; n < s.node_cap; n += blk.count() * blockDim.x) {
        const int32_t t = s.meta[n].type;
        if (t != kLeaf && t != kBranch) continue;
        double xlo = -INFINITY, xhi = INFINITY, ylo = -INFINITY, yhi = INFINITY;
        for (uint32_t c = n; c != 0;) {
            const uint32_t a = s.parent[c];
            const double2 mid = node_mid(s.nodes, a);
            const uint4 ch = node_children(s.nodes, a);
            const uint32_t i = ch.x == c ? 0u : ch.y == c ? 1u : ch.z == c ? 2u : 3u;
            if (i & 1u)
                xlo = fmax(xlo, mid.x);
            else
                xhi = fmin(xhi, mid.x);
            if (i & 2u)
                ylo = fmax(ylo, mid.y);
            else
                yhi = fmin(yhi, mid.y);
            c = a;
        }
        s.region[n] = make_double4(xlo, xhi, ylo, yhi);
    }
}

__global__ __launch_bounds__(kBlock) void k_regions(DevState s) { regions_body(s, this_blk()); }
__global__ __launch_bounds__(kBlock) void kf_regions(const Many m)
{
    const Member w = member_of(m);
    regions_body(m.ss[w.i], w.blk);
}

__device__ __forceinline__ bool in_region(const double4 &r, double x, double y)
{
    return x > r.x && x <= r.y && y > r.z && y <= r.w;
}

// Descent with last_valid (filter.cpp:229-243): last_valid is the deepest
// node on the path, below the root, with x + x_buffer < high.x &&
// y + y_buffer < high.y (both r in planar metres).
// Objects are taken in the leaf order of the tick before (spid), so the
// lanes of a wave mostly share a leaf, whose count one atomic per leaf and
// wave advances.  The order used is kept for the scatter.
// An object starts from its leaf of the tick before (incremental mode): the
// schema update since then left that node a leaf, split it (a branch) or
// merged it (freed; its parent is now a leaf).  The object climbs from there
// to the first node whose region holds it and descends from that node; the
// leaf is the one the descent from the root would reach, since every node's
// region is exactly the set of points the descent sends into it.
// last_valid is then the first node below the root, climbing from the leaf,
// that meets the condition (the deepest one on the path).  Without a leaf of
// the tick before (the first tick, the grid rebuild) the descent starts at
// the root.
template <class P>
__device__ __forceinline__ void assign_body(const DevState s, const P blk)
{
    __shared__ BorderObj stage[kBlock / 64][kBorderStage];
    BorderObj *st = stage[threadIdx.x / 64];
    uint32_t nst = 0;
    const uint32_t lane = lane_id();
    const double *xy = s.tp->xy;
    const double xbuf = s.xbuf, ybuf = s.ybuf;   // x_buffer, y_buffer (filter.cpp:30-31)
    const bool root_leaf = s.nodes[0].child[0] == kLeafTag;
    const bool ordered = s.ctr->sorted_valid != 0;
    // Each thread takes two objects per step (positions base and base +
    // stride, both wave-uniform blocks) and descends for both at once, so
    // that the node reads of one overlap those of the other.
    constexpr uint32_t kU = 2;
    const uint32_t stride = blk.count() * blockDim.x;
    for (uint32_t base0 = blk.id() * blockDim.x; base0 < s.n; base0 += kU * stride) {
        uint32_t kk[kU], id[kU], lv[kU], cur[kU], pleaf[kU];
        bool active[kU], border[kU];
        double x[kU], y[kU];
        uint4 ch[kU];
        double2 mid[kU];
#pragma unroll
        for (uint32_t q = 0; q < kU; q++) {
            const uint32_t k = base0 + q * stride + threadIdx.x;
            kk[q] = k;
            active[q] = k < s.n;
            id[q] = k;
            if (active[q] && ordered) {
                id[q] = s.spid[k];
                if (id[q] >= s.n) id[q] = k;   // not a permutation: only after an earlier error
            }
            border[q] = false;
            lv[q] = 0;
            cur[q] = kNone;
            pleaf[q] = kNone;
            x[q] = 0.0;
            y[q] = 0.0;
            ch[q] = make_uint4(kLeafTag, 0u, 0u, 0u);
            mid[q] = make_double2(0.0, 0.0);
            if (active[q]) {
                x[q] = xy[2 * (size_t)id[q]];
                y[q] = xy[2 * (size_t)id[q] + 1];
                pleaf[q] = ordered ? s.tp->sleaf_old[k] : kNone;   // position k held the object then
                if (!(x[q] >= s.lx && x[q] <= s.hx && y[q] >= s.ly && y[q] <= s.hy)) {
                    atomicOr(&s.ctr->error, kErrBadInput);
                    active[q] = false;
                } else {
                    cur[q] = 0;
                    mid[q] = node_mid(s.nodes, 0);
                    ch[q] = node_children(s.nodes, 0);
                }
            }
        }
        // incremental mode: start from the leaf of the tick before (see above)
        const bool incr = ordered && s.incremental;
        if (incr) {
#pragma unroll
            for (uint32_t q = 0; q < kU; q++) {
                if (!active[q] || pleaf[q] >= s.node_cap) continue;
                uint32_t p = pleaf[q];
                if (s.meta[p].type == kInvalid) p = s.parent[p];   // merged: the parent is the leaf now
                while (p != 0 && !in_region(s.region[p], x[q], y[q])) p = s.parent[p];
                cur[q] = p;
                mid[q] = node_mid(s.nodes, p);
                ch[q] = node_children(s.nodes, p);
            }
        }
        // both descents, one level of each per step
        for (;;) {
            bool more = false;
#pragma unroll
            for (uint32_t q = 0; q < kU; q++) {
                if (ch[q].x == kLeafTag) continue;
                cur[q] = pick_child(ch[q], (y[q] > mid[q].y) * 2 + (x[q] > mid[q].x));
                mid[q] = node_mid(s.nodes, cur[q]);
                const double2 hi = node_high(s.nodes, cur[q]);
                ch[q] = node_children(s.nodes, cur[q]);
                if (!incr && x[q] + xbuf < hi.x && y[q] + ybuf < hi.y) lv[q] = cur[q];
                more = true;
            }
            if (!more) break;
        }
        if (incr) {
#pragma unroll
            for (uint32_t q = 0; q < kU; q++) {
                if (!active[q]) continue;
                for (uint32_t n = cur[q]; n != 0; n = s.parent[n]) {
                    const double2 hi = node_high(s.nodes, n);
                    if (x[q] + xbuf < hi.x && y[q] + ybuf < hi.y) {
                        lv[q] = n;
                        break;
                    }
                }
            }
        }
#pragma unroll
        for (uint32_t q = 0; q < kU; q++) {
            if (active[q]) {
                border[q] = s.complete ? !root_leaf : (lv[q] != cur[q]);
                if (s.complete) lv[q] = 0;
            }
            // ranks: the lanes that share a leaf take consecutive ranks with one atomic
            uint32_t rank = 0;
            bool need = active[q];
            for (;;) {
                const unsigned long long pend = __ballot(need);
                if (!pend) break;
                const uint32_t leader = __ffsll((long long)pend) - 1;
                const uint32_t L = __shfl(cur[q], leader);
                const bool mine = need && cur[q] == L;
                const unsigned long long mm = __ballot(mine);
                uint32_t b = 0;
                if (lane == leader) b = atomicAdd(&s.count[L], (uint32_t)__popcll(mm));
                b = __shfl(b, leader);
                if (mine) {
                    rank = b + (uint32_t)__popcll(mm & ((1ull << lane) - 1ull));
                    need = false;
                }
            }
            const uint32_t k = kk[q];
            if (k < s.n) {
                s.order[k] = id[q];
                s.ord_xy[k] = make_double2(x[q], y[q]);
                s.ord_leaf[k] = active[q] ? cur[q] : kNone;
                s.ord_rank[k] = rank;
                s.ord_pleaf[k] = pleaf[q];
            }
            // border objects (with what the lookup needs) are staged per wave and
            // appended one flush at a time
            const unsigned long long m = __ballot(border[q]);
            if (border[q])
                st[nst + __popcll(m & ((1ull << lane) - 1ull))] =
                    BorderObj{id[q], lv[q], cur[q], pleaf[q], ordered ? k : kNone, rank, x[q], y[q]};
            nst += (uint32_t)__popcll(m);
            if (nst > kBorderStage - 64) {
                wave_sync();
                uint32_t b0 = 0;
                if (lane == 0) b0 = atomicAdd(&s.ctr->n_border, nst);
                b0 = __shfl(b0, 0);
                for (uint32_t t = lane; t < nst; t += 64) s.border[b0 + t] = st[t];
                wave_sync();
                nst = 0;
            }
        }
    }
    if (nst) {
        wave_sync();
        uint32_t b0 = 0;
        if (lane == 0) b0 = atomicAdd(&s.ctr->n_border, nst);
        b0 = __shfl(b0, 0);
        for (uint32_t q = lane; q < nst; q += 64) s.border[b0 + q] = st[q];
    }
}

__global__ __launch_bounds__(kBlock) void k_assign(DevState s) { assign_body(s, this_blk()); }
__global__ __launch_bounds__(kBlock) void kf_assign(const Many m)
{
    const Member w = member_of(m);
    assign_body(m.ss[w.i], w.blk);
}

__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t x)
{
    uint32_t lane = lane_id();
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(x, d);
        if (lane >= (uint32_t)d) x += y;
    }
    return x;
}

// Exclusive scan of a 256-thread block; returns the prefix of this thread.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *total)
{
    __shared__ uint32_t wsum[kBlock / 64];
    uint32_t w = threadIdx.x / 64;
    uint32_t inc = wave_inclusive_scan(v);
    if (lane_id() == 63) wsum[w] = inc;
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (uint32_t k = 0; k < kBlock / 64; k++) {
        if (k < w) before += wsum[k];
        all += wsum[k];
    }
    __syncthreads();
    *total = all;
    return before + inc - v;
}

// Three-kernel exclusive scan over per-node counters (node_cap entries).
template <class P>
__device__ __forceinline__ void scan_reduce_body(const uint32_t *in, uint32_t n, uint32_t *block_sums, const P blk)
{
    uint32_t base = blk.id() * kScanTile + threadIdx.x * kScanItems;
    uint32_t v = 0;
    for (int k = 0; k < kScanItems; k++)
        if (base + k < n) v += in[base + k];
    uint32_t total;
    block_exclusive_scan(v, &total);
    if (threadIdx.x == 0) block_sums[blk.id()] = total;
}

__global__ __launch_bounds__(kBlock) void k_scan_reduce(const uint32_t *in, uint32_t n, uint32_t *block_sums)
{
    scan_reduce_body(in, n, block_sums, this_blk());
}

// The counter arrays the tick sequence scans: leaf object counts, leaf unit
// counts, and the region bins of the contact grouping.
enum ScanArray : int { kScanCount, kScanUcount, kScanBins };

__host__ __device__ inline uint32_t scan_length(const DevState &s, int a)
{
    return a == kScanBins ? (s.coarse_max + 1) * s.bucket_blocks : s.node_cap;
}

__host__ __device__ inline uint32_t scan_blocks_for(uint32_t n) { return (n + kScanTile - 1) / kScanTile; }

template <int A>
__global__ __launch_bounds__(kBlock) void kf_scan_reduce(const Many m)
{
    const Member w = member_of(m);
    const DevState &s = m.ss[w.i];
    const uint32_t *in = A == kScanCount ? s.count : (A == kScanUcount ? s.ucount : s.bcount);
    scan_reduce_body(in, scan_length(s, A), s.block_sums, w.blk);
}

__device__ __forceinline__ void scan_top_body(uint32_t *block_sums, uint32_t nblocks)
{
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < nblocks; b0 += kBlock) {
        uint32_t b = b0 + threadIdx.x;
        uint32_t v = b < nblocks ? block_sums[b] : 0;
        uint32_t total;
        uint32_t ex = block_exclusive_scan(v, &total);
        if (b < nblocks) block_sums[b] = carry + ex;
        carry += total;
    }
}

__global__ __launch_bounds__(kBlock) void k_scan_top(uint32_t *block_sums, uint32_t nblocks)
{
    scan_top_body(block_sums, nblocks);
}

template <int A>
__global__ __launch_bounds__(kBlock) void kf_scan_top(const Many m)
{
    const Member w = member_of(m);
    const DevState &s = m.ss[w.i];
    scan_top_body(s.block_sums, scan_blocks_for(scan_length(s, A)));
}

// Object offsets of every leaf (node-index order) and the leaf statistics.
template <class P>
__device__ __forceinline__ void scan_objects_body(const DevState s, const P blk)
{
    uint32_t base = blk.id() * kScanTile + threadIdx.x * kScanItems;
    uint32_t c[kScanItems];
    uint32_t v = 0;
    for (int k = 0; k < kScanItems; k++) {
        c[k] = base + k < s.node_cap ? s.count[base + k] : 0;
        v += c[k];
    }
    uint32_t total;
    uint32_t ex = block_exclusive_scan(v, &total) + s.block_sums[blk.id()];
    // leaves beyond the reference's leaf storage of 2*cap objects
    // (grid_amplify, workbench.cpp:36), where it would drop objects
    // (filter.cpp:100-104): only leaves hold objects
    const uint32_t store = 2 * s.cap;
    uint32_t mx = 0, nleaf = 0, over = 0, drop = 0;
    for (int k = 0; k < kScanItems; k++) {
        uint32_t n = base + k;
        if (n < s.node_cap) {
            s.offset[n] = ex;
            nleaf += s.meta[n].type == kLeaf ? 1u : 0u;
        }
        ex += c[k];
        mx = max(mx, c[k]);
        over += c[k] > store ? 1u : 0u;
        drop += c[k] > store ? c[k] - store : 0u;
    }
    for (int d = 32; d >= 1; d >>= 1) {
        mx = max(mx, (uint32_t)__shfl_xor(mx, d));
        nleaf += (uint32_t)__shfl_xor(nleaf, d);
        over += (uint32_t)__shfl_xor(over, d);
        drop += (uint32_t)__shfl_xor(drop, d);
    }
    if (lane_id() == 0) {
        atomicMax(&s.ctr->max_leaf, mx);
        atomicAdd(&s.ctr->n_leaves, nleaf);
        if (over) {
            atomicAdd(&s.ctr->n_drop_leaves, over);
            atomicAdd(&s.ctr->n_drop_objs, (unsigned long long)drop);
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_scan_objects(DevState s) { scan_objects_body(s, this_blk()); }
__global__ __launch_bounds__(kBlock) void kf_scan_objects(const Many m)
{
    const Member w = member_of(m);
    scan_objects_body(m.ss[w.i], w.blk);
}

// Unit offsets of every leaf, and the join work list: leaves with >= 2
// objects (inside pairs) or with border units aimed at them.  Leaves of more
// than kHeavyLeaf objects are listed from the back, so that the join takes
// them first and ends on small leaves.
template <class P>
__device__ __forceinline__ void scan_units_body(const DevState s, const P blk)
{
    uint32_t base = blk.id() * kScanTile + threadIdx.x * kScanItems;
    uint32_t c[kScanItems];
    uint32_t v = 0;
    for (int k = 0; k < kScanItems; k++) {
        c[k] = base + k < s.node_cap ? s.ucount[base + k] : 0;
        v += c[k];
    }
    uint32_t total;
    uint32_t ex = block_exclusive_scan(v, &total) + s.block_sums[blk.id()];
    for (int k = 0; k < kScanItems; k++) {
        uint32_t n = base + k;
        bool in = n < s.node_cap;
        if (in) s.ustart[n] = ex;
        ex += c[k];
        const uint32_t cnt = in ? s.count[n] : 0u;
        const bool work = in && (c[k] > 0 || cnt >= 2);
        const bool heavy = work && cnt > kHeavyLeaf;
        const uint32_t idx = wave_append_u32(&s.ctr->n_leaf_list, work && !heavy);
        const uint32_t hdx = wave_append_u32(&s.ctr->n_heavy, heavy);
        if (work) s.leaf_list[heavy ? s.node_cap - 1 - hdx : idx] = n;
    }
}

__global__ __launch_bounds__(kBlock) void k_scan_units(DevState s) { scan_units_body(s, this_blk()); }
__global__ __launch_bounds__(kBlock) void kf_scan_units(const Many m)
{
    const Member w = member_of(m);
    scan_units_body(m.ss[w.i], w.blk);
}

// Counting sort by leaf, in the order of the assignment: objects that share
// a leaf were given consecutive ranks there, so the sorted arrays are
// written in runs.
template <bool GEO, class P>
__device__ __forceinline__ void scatter_body(const DevState s, const P blk)
{
    const double *xy_prev = s.tp->xy_prev;
    double *xy_save = s.tp->xy_save;
    double *sx = s.tp->sx, *sy = s.tp->sy;
    const double *sx_old = s.tp->sx_old, *sy_old = s.tp->sy_old;
    const bool has_prev = s.tp->has_prev != 0;
    // in the previous tick's leaf order, position k held the object now at k
    const bool ordered = s.ctr->sorted_valid != 0;
    for (uint32_t k = blk.id() * blockDim.x + threadIdx.x; k < s.n; k += blk.count() * blockDim.x) {
        const uint32_t leaf = s.ord_leaf[k];
        if (leaf == kNone) continue;
        const uint32_t i = s.order[k];
        const uint32_t pos = s.offset[leaf] + s.ord_rank[k];
        const double2 p = s.ord_xy[k];
        const double x = p.x, y = p.y;
        sx[pos] = x;
        sy[pos] = y;
        s.spid[pos] = i;
        s.sprev[pos] = ordered ? k : kNone;
        s.tp->sleaf[pos] = leaf;
        s.splf[pos] = s.ord_pleaf[k];
        if (GEO) s.sgeo[pos] = deg_per_meter_lon(y);   // geography: the object's longitude scale
        double px = 0.0, py = 0.0;
        if (has_prev) {
            px = ordered ? sx_old[k] : xy_prev[2 * (size_t)i];
            py = ordered ? sy_old[k] : xy_prev[2 * (size_t)i + 1];
        }
        if (xy_save) {
            xy_save[2 * (size_t)i] = x;
            xy_save[2 * (size_t)i + 1] = y;
        }
        if (s.precision == GLINT_PRECISION_FP32_GUARD) {
            Node nd = s.nodes[leaf];
            s.fx[pos] = (float)(x - nd.lx);
            s.fy[pos] = (float)(y - nd.ly);
            if (has_prev) {
                const double ox = px - nd.lx, oy = py - nd.ly;
                s.pfx[pos] = fabs(ox) <= s.pf_lim ? (float)ox : __builtin_nanf("");
                s.pfy[pos] = fabs(oy) <= s.pf_lim ? (float)oy : __builtin_nanf("");
            }
        }
    }
}

template <bool GEO>
__global__ __launch_bounds__(kBlock) void k_scatter(DevState s) { scatter_body<GEO>(s, this_blk()); }
template <bool GEO>
__global__ __launch_bounds__(kBlock) void kf_scatter(const Many m)
{
    const Member w = member_of(m);
    scatter_body<GEO>(m.ss[w.i], w.blk);
}

__device__ __forceinline__ uint32_t wave_exclusive_scan(uint32_t v, uint32_t *total)
{
    uint32_t inc = wave_inclusive_scan(v);
    *total = __shfl(inc, 63);
    return inc - v;
}

// Rank of this lane's count in p[key] (kNone: none): the lanes of the wave
// that share a key take consecutive ranks with one atomic.  Every lane of the
// wave must call it.
__device__ __forceinline__ uint32_t wave_rank(uint32_t *p, uint32_t key)
{
    uint32_t rank = 0;
    bool need = key != kNone;
    for (;;) {
        const unsigned long long pend = __ballot(need);
        if (!pend) break;
        const uint32_t leader = __ffsll((long long)pend) - 1;
        const uint32_t K = __shfl(key, leader);
        const bool mine = need && key == K;
        const unsigned long long mm = __ballot(mine);
        uint32_t b = 0;
        if (lane_id() == leader) b = atomicAdd(&p[K], (uint32_t)__popcll(mm));
        b = __shfl(b, leader);
        if (mine) {
            rank = b + (uint32_t)__popcll(mm & ((1ull << lane_id()) - 1ull));
            need = false;
        }
    }
    return rank;
}

// The children of a node as the lookup sees them: the four child nodes are
// read together; accepted leaves are passed to take(leaf) when `takes`, and
// the mask of branch children within r is returned.
template <bool GEO, class Take>
__device__ __forceinline__ uint32_t lookup_children(const DevState &s, uint4 ids, double x, double y, double f,
                                                    uint32_t own, bool takes, Take &&take)
{
    const Node n0 = s.nodes[ids.x], n1 = s.nodes[ids.y], n2 = s.nodes[ids.z], n3 = s.nodes[ids.w];
    uint32_t desc = 0;
    auto one = [&](const Node &cn, uint32_t c, uint32_t bit) {
        if (!(GEO ? box_within_geo(cn, x, y, s.t64, f) : box_within(cn, x, y, s.t64))) return;
        if (cn.child[0] != kLeafTag) {
            desc |= bit;
        } else if (takes && (s.complete ? (c != own) : (y < cn.ly || (y < cn.hy && x < cn.lx)))) {
            take(c);
        }
    };
    one(n0, ids.x, 1u);
    one(n1, ids.y, 2u);
    one(n2, ids.z, 4u);
    one(n3, ids.w, 8u);
    return desc;
}

// lookup_rec (filter.cpp:130-152) below node `from`: every leaf within r of
// (x, y) that the top/right rule accepts (in complete mode every leaf within
// r other than `own`) is passed to take(leaf).  A depth-first walk without a
// stack in memory: the branch children still to visit at each of the first
// 16 levels below `from` are kept as 4-bit masks in one 64-bit register, and
// the walk climbs back through the parent links (deeper levels recompute
// their mask from the children), so nothing sits in scratch memory.
// GEO: geography, f the degrees of longitude per metre at y.
template <bool GEO, class Take>
__device__ __forceinline__ void lookup_walk(const DevState &s, double x, double y, double f, uint32_t from,
                                            uint32_t own, Take &&take)
{
    uint32_t node = from;
    uint4 ids = node_children(s.nodes, from);
    uint32_t desc = lookup_children<GEO>(s, ids, x, y, f, own, true, take);
    unsigned long long masks = 0;   // level d (< 16): masks >> 4d
    uint32_t depth = 0;             // of `node` below `from`
    for (;;) {
        if (desc) {
            const uint32_t ci = __builtin_ctz(desc);
            if (depth < 16) {
                const unsigned long long rest = desc & ~((2u << ci) - 1u);
                masks = (masks & ~(0xFull << (4 * depth))) | (rest << (4 * depth));
            }
            node = pick_child(ids, (int)ci);
            depth++;
            ids = node_children(s.nodes, node);
            desc = lookup_children<GEO>(s, ids, x, y, f, own, true, take);
            continue;
        }
        if (node == from) break;
        const uint32_t p = s.parent[node];
        if (p >= s.node_cap) {   // not reached: every node below `from` has its parent
            atomicOr(&s.ctr->error, kErrStack);
            break;
        }
        depth--;
        ids = node_children(s.nodes, p);
        if (depth < 16) {
            desc = (uint32_t)(masks >> (4 * depth)) & 0xFu;
        } else {
            const uint32_t ci = ids.x == node ? 0u : (ids.y == node ? 1u : (ids.z == node ? 2u : 3u));
            desc = lookup_children<GEO>(s, ids, x, y, f, own, false, take) & ~((2u << ci) - 1u);
        }
        node = p;
    }
}

// The lookup from last_valid for every border object.  Each unit (object,
// candidate leaf) takes a rank among the units of its leaf.  The object's
// first three candidates (kCandOverflow in the third when there are more)
// are kept, stamped with the tick, for the joins of this tick and the next.
// Objects with up to three candidates (nearly all) write their units from
// registers after one atomic per wave; the others walk a second time.
template <bool GEO, class P>
__device__ __forceinline__ void lookup_body(const DevState s, const P blk)
{
    const uint32_t nb = s.ctr->n_border;
    // the scatter before has put the objects in leaf order (read by the next tick)
    if (blk.id() == 0 && threadIdx.x == 0) s.ctr->sorted_valid = 1;
    for (uint32_t base = blk.id() * blockDim.x; base < nb; base += blk.count() * blockDim.x) {
        const uint32_t k = base + threadIdx.x;
        uint32_t total = 0;
        uint32_t f0 = kNone, f1 = kNone, f2 = kNone;
        uint32_t pid = 0, lv = 0, own = 0, pleaf = kNone, prevk = kNone, pos = 0;
        double x = 0.0, y = 0.0, f = 0.0;
        if (k < nb) {
            const BorderObj b = s.border[k];
            pid = b.pid;
            lv = b.lv;
            x = b.x;
            y = b.y;
            if (GEO) f = deg_per_meter_lon(y);
            own = b.leaf;
            pleaf = b.pleaf;
            prevk = b.prevk;
            pos = s.offset[own] + b.rank;
            lookup_walk<GEO>(s, x, y, f, lv, own, [&](uint32_t c) {
                f0 = total == 0 ? c : f0;
                f1 = total == 1 ? c : f1;
                f2 = total == 2 ? c : f2;
                total++;
            });
            if (!s.store_all) s.tp->cand[pos] = make_uint4(f0, f1, total > 3 ? kCandOverflow : f2, s.tp->tick);
        }
        // one atomic per wave reserves the units of the lists of up to three;
        // lanes that share a candidate leaf take consecutive ranks there with
        // one atomic (objects come in leaf order, so neighbours share leaves)
        const uint32_t nc = total <= 3 ? total : 0u;
        uint32_t wtotal;
        const uint32_t wex = wave_exclusive_scan(nc, &wtotal);
        unsigned long long wbase = 0;
        if (lane_id() == 0 && wtotal) wbase = atomicAdd(&s.ctr->n_units, (unsigned long long)wtotal);
        wbase = __shfl(wbase, 0);
        const uint32_t r0 = wave_rank(s.ucount, nc > 0 ? f0 : kNone);
        const uint32_t r1 = wave_rank(s.ucount, nc > 1 ? f1 : kNone);
        const uint32_t r2 = wave_rank(s.ucount, nc > 2 ? f2 : kNone);
        unsigned long long u = wbase + wex;
        // what the join's border rows need of the object, beside each unit
        const uint4 aux = make_uint4(own | (total > 3 ? 0x80000000u : 0u), pleaf, prevk, pos);
        auto write = [&](uint32_t c, uint32_t rank) {
            if (u < s.unit_cap) {
                s.units[u] = make_uint2(pid, c);
                s.uaux[u] = aux;
                s.urank[u] = rank;
            } else {
                atomicOr(&s.ctr->error, kErrUnitCap);
            }
            u++;
        };
        if (nc > 0) write(f0, r0);
        if (nc > 1) write(f1, r1);
        if (nc > 2) write(f2, r2);
        if (total > 3) {
            u = atomicAdd(&s.ctr->n_units, (unsigned long long)total);
            lookup_walk<GEO>(s, x, y, f, lv, own, [&](uint32_t c) { write(c, atomicAdd(&s.ucount[c], 1u)); });
        }
    }
}

template <bool GEO>
__global__ __launch_bounds__(kBlock) void k_lookup(DevState s) { lookup_body<GEO>(s, this_blk()); }
template <bool GEO>
__global__ __launch_bounds__(kBlock) void kf_lookup(const Many m)
{
    const Member w = member_of(m);
    lookup_body<GEO>(m.ss[w.i], w.blk);
}

// Units grouped by target leaf: unit_pid[ustart[leaf] + rank] = object, and
// its leaf and previous leaf beside it (unit_lf).
template <class P>
__device__ __forceinline__ void unit_scatter_body(const DevState s, const P blk)
{
    if (s.ctr->error & kErrUnitCap) return;   // units were dropped: the tick is void
    const unsigned long long nu = s.ctr->n_units;
    for (unsigned long long u = blk.id() * (unsigned long long)blockDim.x + threadIdx.x; u < nu;
         u += (unsigned long long)blk.count() * blockDim.x) {
        const uint2 un = s.units[u];
        const uint32_t q = s.ustart[un.y] + s.urank[u];
        s.unit_pid[q] = un.x;
        s.unit_lf[q] = s.uaux[u];   // 16 bytes: leaf, previous leaf, previous and current sorted position
    }
}

__global__ __launch_bounds__(kBlock) void k_unit_scatter(DevState s) { unit_scatter_body(s, this_blk()); }
__global__ __launch_bounds__(kBlock) void kf_unit_scatter(const Many m)
{
    const Member w = member_of(m);
    unit_scatter_body(m.ss[w.i], w.blk);
}

// ------------------------------------------------------------ meeting table
//
// The table is an open-addressed hash of the pair key (pid_lo << 32) | pid_hi
// in two generations of B 16-byte slots.  h = pair_hash(key): the high bits pick
// a region of kRegion slots, the 11 bits below them the home slot inside it,
// and linear probing wraps inside the region.  A tick's contacts are bucketed by region
// and every region is then updated by one workgroup in LDS: the region of the
// previous generation is read once, the new region is written once.

// The join writes the tick's contacts where the first bucketing pass reads
// them: into the current generation's list when there are two passes (the
// second pass writes the grouped list back there), else into the bucketing
// buffer (the one pass writes the grouped list into the generation).
__device__ __forceinline__ unsigned long long *contact_out(const DevState &s)
{
    return s.tp->tab_cur.keys;
}

// Region of a pair in a table of 2^bits regions: the high bits of its hash
// (the low bits pick its slot in the region's LDS hash).
__device__ __forceinline__ unsigned long long pair_hash(unsigned long long key)
{
    return key * 0x9E3779B97F4A7C15ull;   // multiplicative: every key bit reaches the high bits
}
__device__ __forceinline__ uint32_t region_of(unsigned long long key, uint32_t bits)
{
    // a one-region table has no region bits (a shift by 64 is undefined)
    return bits ? (uint32_t)(pair_hash(key) >> (64 - bits)) : 0u;
}
// Home slot of a pair in its region's LDS hash: the kRegion-bit field of
// the hash just below the region bits.
static_assert(kRegion == 2048, "slot_of takes 11 bits");
__device__ __forceinline__ uint32_t slot_of(unsigned long long key, uint32_t bits)
{
    return (uint32_t)(pair_hash(key) >> (64 - 11 - bits)) & (kRegion - 1);
}

// The table's shape at this tick, from the contacts the join stored: the
// fewest regions (at least 2^rbits_min) that keep kRegionFill contacts per
// region on average, half a region's LDS slots, so that a tick denser than
// the configuration expected still fits.  Every kernel after the join
// computes the same shape from the same count.  Regions split into
// 2^(bits - lo_bits) <= kMaxCoarse coarse bins of 2^lo_bits each.
struct TableShape {
    uint32_t bits, lo_bits, coarse;
};

__device__ __forceinline__ TableShape table_shape(const DevState &s)
{
    const unsigned long long n = s.ctr->n_stored;
    uint32_t b = s.rbits_min;
    while (b < s.rbits_max && ((unsigned long long)kRegionFill << b) < n) b++;
    TableShape t;
    t.bits = b;
    t.lo_bits = b > 10 ? b - 10 : 0;
    t.coarse = 1u << (b - t.lo_bits);
    return t;
}

// Contacts found by one wave are staged in LDS as (row lane, column) index
// pairs and appended to the tick's contact list in 4096-entry chunks
// reserved with one atomic each; the pair key (and its flags) is formed when
// the stage is written out, one lane per contact.
struct Stager {
    unsigned long long *out;    // the tick's contact list
    uint32_t *buf;
    uint32_t n;                 // staged entries (uniform across the wave)
    unsigned long long chunk;   // current chunk in the contact list
    uint32_t used;              // entries of the chunk written
    uint32_t kept;              // entries written in all (uniform)
};

// The lanes of mask m (bit i: lane i) append e in lane order (mbcnt counts the
// lanes of m below each, from st.n); an empty mask writes nothing.
__device__ __forceinline__ void stage_push_mask(Stager &st, unsigned long long m, uint32_t e)
{
    if (m == 0) return;
    if (__builtin_amdgcn_inverse_ballot_w64(m))
        st.buf[st.n + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u))] = e;
    st.n += (uint32_t)__popcll(m);
}

// Writes the staged entries out: key(e, valid) gives the contact-list key of
// entry e, or kEmptyKey for an entry that is not kept (and for !valid);
// kept keys are packed.
template <class Key>
__device__ __forceinline__ void stage_flush(Stager &st, const DevState &s, Key key)
{
    if (st.n == 0) return;
    wave_sync();
    const uint32_t lane = lane_id();
    for (uint32_t done = 0; done < st.n; done += 64) {
        const uint32_t i = done + lane;
        // every lane calls key (it may exchange registers across the wave)
        const unsigned long long k = key(i < st.n ? st.buf[i] : 0u, i < st.n);
        const bool keep = k != kEmptyKey;
        const unsigned long long m = __ballot(keep);
        const uint32_t cnt = (uint32_t)__popcll(m);
        if (cnt == 0) continue;
        st.kept += cnt;
        const uint32_t C = s.chunk;
        if (st.used + cnt > C) {
            // close the open chunk (its tail holds kEmptyKey) and reserve the next
            if (st.chunk + C <= s.clist_cap)
                for (uint32_t t = st.used + lane; t < C; t += 64) st.out[st.chunk + t] = kEmptyKey;
            unsigned long long c = 0;
            if (lane == 0) c = atomicAdd(&s.ctr->n_clist, (unsigned long long)C);
            st.chunk = __shfl(c, 0);
            st.used = 0;
        }
        if (st.chunk + C <= s.clist_cap) {
            if (keep) st.out[st.chunk + st.used + __popcll(m & ((1ull << lane) - 1ull))] = k;
        } else if (lane == 0) {
            atomicOr(&s.ctr->error, kErrContactCap);
        }
        st.used += cnt;
    }
    wave_sync();
    st.n = 0;
}

// The unused tail of a wave's last chunk holds kEmptyKey (skipped later).
__device__ __forceinline__ void stage_close(Stager &st, const DevState &s)
{
    if (st.used < s.chunk && st.chunk + s.chunk <= s.clist_cap)
        for (uint32_t k = st.used + lane_id(); k < s.chunk; k += 64) st.out[st.chunk + k] = kEmptyKey;
}

// ------------------------------------------------------------ refine: join

__device__ __forceinline__ unsigned long long pair_key(uint32_t a, uint32_t b)
{
    uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
    return ((unsigned long long)lo << 32) | hi;
}

// Candidate list of the object sorted at `at` in the tick that wrote buf, as
// the join tests it: up to three leaves and w = 0, or w = 4 when the lookup
// had more (an entry not written at `stamp` is empty).
__device__ __forceinline__ uint4 cand_read(const uint4 *buf, uint32_t at, uint32_t stamp)
{
    const uint4 c = buf[at];
    if (c.w != stamp) return make_uint4(kNone, kNone, kNone, 0u);
    return c.z == kCandOverflow ? make_uint4(kNone, kNone, kNone, 4u) : make_uint4(c.x, c.y, c.z, 0u);
}

// The same at a previous sorted position that may be unknown (at >= n): an
// empty list.  Positions are unknown only before the first lookup, when no
// tick before is evaluated.
__device__ __forceinline__ uint4 cand_at(const uint4 *buf, uint32_t at, uint32_t n, uint32_t stamp)
{
    return at < n ? cand_read(buf, at, stamp) : make_uint4(kNone, kNone, kNone, 0u);
}

// One bit per leaf id: a 32-bit summary of a candidate list whose clear bits
// prove a leaf is not on it (all bits set for a list that overflowed).
__device__ __forceinline__ uint32_t leaf_bit(uint32_t leaf) { return 1u << ((leaf * 0x9E3779B1u) >> 27); }
__device__ __forceinline__ uint32_t cand_mask(uint4 c)
{
    if (c.w > 3) return 0xFFFFFFFFu;
    return (c.x != kNone ? leaf_bit(c.x) : 0u) | (c.y != kNone ? leaf_bit(c.y) : 0u) |
           (c.z != kNone ? leaf_bit(c.z) : 0u);
}

// Squared FP32 distance of a guard-band test: dy*dy rounded, then dx*dx added
// in one fused operation (one rounding fewer than a product and a sum; the
// band's error bound covers both).  The column loop and the write-out use it
// alike, so they take the same decisions.
__device__ __forceinline__ float dist2_f32(float dx, float dy) { return __builtin_fmaf(dx, dx, dy * dy); }

// Positions of a tile object: FP64 for a leaf joined in FP64, else the FP32
// offsets from the leaf's low corner at tick and at prev_tick (NaN: not
// bounded).  The FP64 positions of a guard-band leaf stay in HBM (sx, sy),
// read only for the rare band re-test.
struct TileF {
    float2 f, pf;
};

// Bounding box of a group of tile objects in the tile's frame, in FP32:
// exact for FP32 offsets, rounded outward for FP64 positions (so it holds
// them); an empty group's box is (+inf, +inf, -inf, -inf).
struct GBox {
    float lx, ly, hx, hy;
};

// The float next to finite f towards -inf (+inf): one step of its bit
// pattern (zero steps to the smallest subnormal of the other sign).
__device__ __forceinline__ float f32_prev(float f)
{
    const uint32_t b = __builtin_bit_cast(uint32_t, f);
    return f > 0.f ? __builtin_bit_cast(float, b - 1u)
                   : (f < 0.f ? __builtin_bit_cast(float, b + 1u) : __builtin_bit_cast(float, 0x80000001u));
}

__device__ __forceinline__ float f32_next(float f) { return -f32_prev(-f); }

// The largest float <= v and the smallest float >= v.
__device__ __forceinline__ float f32_down(double v)
{
    const float f = (float)v;
    return (double)f > v ? f32_prev(f) : f;
}

__device__ __forceinline__ float f32_up(double v)
{
    const float f = (float)v;
    return (double)f < v ? f32_next(f) : f;
}

constexpr uint32_t kGroup = 16;      // tile columns per box
constexpr uint32_t kSubCells = 64;   // 8 x 8 sub-cells of a leaf, in Morton order

template <int TILE>
struct JoinTile {
    union {
        double2 xy[TILE];
        TileF fp[TILE];
    };
    uint32_t g[TILE];              // index in the sorted arrays
    uint32_t pl[TILE];             // leaf at prev_tick
    uint32_t p[TILE];              // object ids
    uint4 pc[TILE];                // candidate leaves at prev_tick (DevState::prev_cand)
    GBox box[TILE / kGroup];       // box of every kGroup consecutive slots
};

// The row objects of a wave's current task (read when its contacts are
// written out).
struct JoinRows {
    float2 pf[64];      // FP32 offsets at prev_tick (guard mode; NaN: unbounded)
    uint32_t p[64];
    uint32_t pl[64];    // border rows: the object's leaf, bit 31 set when its candidates overflowed
    uint32_t ppl[64];   // the object's leaf at prev_tick
};

// One object of a tile as read from the sorted arrays: FP32 offsets of this
// and the previous tick (use32) or FP64 coordinates, leaf at prev_tick, id,
// candidate leaves at prev_tick (or the geography scale), and this tick's
// candidate list.
struct TileObj {
    uint4 pos;   // the TileF (use32) or the double2, as bits
    uint32_t pl, p;
    uint4 pc, cc;
    __device__ float fx() const { return __uint_as_float(pos.x); }
    __device__ float fy() const { return __uint_as_float(pos.y); }
    __device__ double x() const { return __hiloint2double((int)pos.y, (int)pos.x); }
    __device__ double y() const { return __hiloint2double((int)pos.w, (int)pos.z); }
};

// Sorted object `from` of a leaf joined in FP32 (use32) or FP64; every load
// is issued here, before any is used.
__device__ __forceinline__ TileObj tile_fetch(const DevState &s, uint32_t from, bool has_prev, bool use32,
                                              bool want_pc, bool cand)
{
    TileObj o;
    if (use32) {
        o.pos = make_uint4(__float_as_uint(s.fx[from]), __float_as_uint(s.fy[from]), 0u, 0u);
        if (has_prev) {
            o.pos.z = __float_as_uint(s.pfx[from]);
            o.pos.w = __float_as_uint(s.pfy[from]);
        }
    } else {
        const double x = s.tp->sx[from], y = s.tp->sy[from];
        o.pos = make_uint4((uint32_t)__double2loint(x), (uint32_t)__double2hiint(x), (uint32_t)__double2loint(y),
                           (uint32_t)__double2hiint(y));
    }
    o.pl = s.splf[from];
    o.p = s.spid[from];
    o.pc = make_uint4(kNone, kNone, kNone, 0u);
    if (want_pc) o.pc = cand_at(s.tp->cand_prev, s.sprev[from], s.n, s.tp->prev_tick);
    if (s.sgeo) {   // geography keeps every contact (no candidate lists): the slot holds the longitude scale
        const unsigned long long f = __double_as_longlong(s.sgeo[from]);
        o.pc = make_uint4((uint32_t)f, (uint32_t)(f >> 32), 0u, 0u);
    }
    o.cc = cand ? cand_read(s.tp->cand, from, s.tp->tick) : make_uint4(kNone, kNone, kNone, 0u);
    return o;
}

__device__ __forceinline__ double tile_geo(const uint4 &pc)
{
    return __longlong_as_double((long long)(((unsigned long long)pc.y << 32) | pc.x));
}

__device__ __forceinline__ uint32_t spread3(uint32_t v) { return (v & 1u) | ((v & 2u) << 1) | ((v & 4u) << 2); }

// Sorted objects from .. from+n-1 of one leaf into tile slots 0..n-1, in the
// Morton order of the leaf's 8 x 8 sub-cells (a counting sort in LDS), and
// the box of every kGroup consecutive slots: a wave task then skips the
// column groups whose box is out of reach of its rows' box.  With cand, the
// candidate lists written at this tick go to ccand/cmask in the same order.
// Each thread reads all of its objects' fields at once (one round of loads,
// plus the candidate list at prev_tick behind its sorted index) and keeps
// them in registers through the sort.  Every thread of the block calls it;
// it ends with the tile complete.
template <int TILE>
__device__ __forceinline__ void load_tile_sorted(JoinTile<TILE> &t, const DevState &s, uint32_t from, uint32_t n,
                                                 bool has_prev, bool use32, double lx, double ly, double w,
                                                 bool cand, uint4 *ccand, uint32_t *cmask, uint32_t *hist)
{
    constexpr int kPer = (TILE + kBlock - 1) / kBlock;
    const bool want_pc = has_prev && !s.store_all;
    const double scale = w > 0.0 ? 8.0 / w : 0.0;
    if (threadIdx.x < kSubCells) hist[threadIdx.x] = 0;
    TileObj obj[kPer];
#pragma unroll
    for (int q = 0; q < kPer; q++) {
        const uint32_t k = threadIdx.x + q * kBlock;
        if (k < n) obj[q] = tile_fetch(s, from + k, has_prev, use32, want_pc, cand);
    }
    __syncthreads();
    uint32_t cell[kPer], rank[kPer];
#pragma unroll
    for (int q = 0; q < kPer; q++) {
        const uint32_t k = threadIdx.x + q * kBlock;
        cell[q] = 0;
        if (k < n) {
            // the cell only orders the tile; any value is correct
            const double ox = use32 ? (double)obj[q].fx() : obj[q].x() - lx;
            const double oy = use32 ? (double)obj[q].fy() : obj[q].y() - ly;
            const uint32_t cx = (uint32_t)fmin(fmax(ox * scale, 0.0), 7.0);
            const uint32_t cy = (uint32_t)fmin(fmax(oy * scale, 0.0), 7.0);
            cell[q] = spread3(cx) | (spread3(cy) << 1);
            rank[q] = atomicAdd(&hist[cell[q]], 1u);
        }
    }
    __syncthreads();
    if (threadIdx.x < 64) {   // exclusive offsets of the 64 cells, in place
        const uint32_t v = hist[threadIdx.x];
        hist[threadIdx.x] = wave_inclusive_scan(v) - v;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kPer; q++) {
        const uint32_t k = threadIdx.x + q * kBlock;
        if (k < n) {
            const uint32_t d = hist[cell[q]] + rank[q];
            // (the same 16 bytes either way)
            reinterpret_cast<uint4 *>(t.fp)[d] = obj[q].pos;
            t.g[d] = from + k;
            t.pl[d] = obj[q].pl;
            t.p[d] = obj[q].p;
            if (want_pc || s.sgeo) t.pc[d] = obj[q].pc;
            if (cand) cmask[d] = cand_mask(ccand[d] = obj[q].cc);
        } else if (k < (n + kGroup - 1) / kGroup * kGroup) {
            // the rest of the last group: positions no distance test accepts
            reinterpret_cast<uint4 *>(t.fp)[k] =
                use32 ? make_uint4(0x7F800000u, 0x7F800000u, 0x7F800000u, 0x7F800000u)
                      : make_uint4(0u, 0x7FF00000u, 0u, 0x7FF00000u);
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kPer; q++) {
        const uint32_t k = threadIdx.x + q * kBlock;
        float x0 = __builtin_inff(), y0 = __builtin_inff(), x1 = -__builtin_inff(), y1 = -__builtin_inff();
        if (k < n) {
            if (use32) {
                x0 = x1 = t.fp[k].f.x;
                y0 = y1 = t.fp[k].f.y;
            } else {
                x0 = f32_down(t.xy[k].x);
                x1 = f32_up(t.xy[k].x);
                y0 = f32_down(t.xy[k].y);
                y1 = f32_up(t.xy[k].y);
            }
        }
        for (int d = kGroup / 2; d >= 1; d >>= 1) {
            x0 = fminf(x0, __shfl_xor(x0, d));
            y0 = fminf(y0, __shfl_xor(y0, d));
            x1 = fmaxf(x1, __shfl_xor(x1, d));
            y1 = fmaxf(y1, __shfl_xor(y1, d));
        }
        if (k < (uint32_t)TILE && k % kGroup == 0) t.box[k / kGroup] = GBox{x0, y0, x1, y1};
    }
    __syncthreads();
}

// True when no pair of a row in box a and a column in box b can pass the
// distance test of the column loop: the squared gap between the boxes,
// computed with the loop's own operations (FP32 fused, or FP64), is past
// the threshold.  Rounding is monotone, so every pair's computed squared
// distance is at least the gap's.
// For FP64 positions the boxes hold them (rounded outward), and the gap
// computed in FP64 from their corners is at most the pair's own difference.
template <bool USE32>
__device__ __forceinline__ bool box_out_of_reach(const GBox &a, const GBox &b, float t32_out, double t64)
{
    if constexpr (USE32) {
        const float gx = fmaxf(fmaxf(b.lx - a.hx, a.lx - b.hx), 0.f);
        const float gy = fmaxf(fmaxf(b.ly - a.hy, a.ly - b.hy), 0.f);
        return dist2_f32(gx, gy) > t32_out;
    } else {
        const double gx = fmax(fmax((double)b.lx - (double)a.hx, (double)a.lx - (double)b.hx), 0.0);
        const double gy = fmax(fmax((double)b.ly - (double)a.hy, (double)a.ly - (double)b.hy), 0.0);
        return gx * gx + gy * gy > t64;
    }
}

// Stage entries: row lane << 16 | column.  In guard mode the write-out
// repeats the FP32 distance of the entry (the row lane's offsets exchanged
// across the wave, the same operations as the column loop, so the same
// result): a pair inside the guard band is re-tested in FP64 there.

// One wave task's column loop: the row lane's object against columns
// [jbeg, jend) of tile C.  An idle lane's position is infinite, which no
// distance test accepts.
// In guard mode (USE32) pairs are decided on the FP32 offsets: inside
// r - eps a hit, past r + eps a miss, in between a band entry.  Sure hits
// are counted here and staged only if they may be stored: when every
// contact is kept, or when the FP32 offsets at prev_tick do not place the
// pair surely out of reach then (the exact test follows at write-out).  A
// sure border hit that is surely untagged is staged only when a candidate
// list overflowed (it is stored then) or when the column object's candidate
// summary does not rule out the row's leaf (the pair may be found twice,
// which key_of counts); any other is neither stored nor a second copy.
struct JoinCounts {
    uint32_t sure;   // sure contacts (per lane)
    uint32_t pass;   // band entries within reach in FP64 (per lane)
    uint32_t band;   // band entries re-tested (per lane)
    uint32_t dup;    // pairs found twice and not stored, counted once (per lane)
    uint32_t cols;   // columns the wave evaluated (uniform; the rest were skipped by box)
};

// Register x of lane `src` (every lane of the wave must take part).
__device__ __forceinline__ uint4 lane_read4(uint4 x, uint32_t src)
{
    return make_uint4((uint32_t)__shfl((int)x.x, (int)src), (uint32_t)__shfl((int)x.y, (int)src),
                      (uint32_t)__shfl((int)x.z, (int)src), (uint32_t)__shfl((int)x.w, (int)src));
}

// Positions of the tick read by the write-out, loaded once per kernel (the
// tick parameters live in device memory and would be re-read after stores).
struct TickView {
    const double *xy, *xy_prev, *sx, *sy, *sx_old, *sy_old;
};

// ROT: a diagonal inside task (rows and columns the same block of n objects
// from jbeg): lane i tests the objects i+1 .. i+lim(i) (mod n), lim(i) =
// (n-1)/2, plus one for i < n/2 when n is even, so each unordered pair is
// tested once and no lane idles on a lower triangle.
// GEO (geography, FP64 only): the distance of the reference's test
// p1->distance(p2, true) (refinement.cpp:33, geometry.h:40-48): dy divided
// by the degrees of latitude per metre and dx by the degrees of longitude
// per metre at p1's latitude, p1 the unit's object (border) or the smaller
// id (inside, where pid < q is tested); fi and pi are the row's scale and id.
template <bool USE32, bool BORDER, int TILE, bool ROT = false, bool GEO = false>
__device__ __forceinline__ JoinCounts join_columns(Stager &st, const DevState &s, const TickView &tv,
                                                   const JoinTile<TILE> &C, double xi,
                                                   double yi, float fxi, float fyi, float pfxi, float pfyi,
                                                   uint32_t jbeg, uint32_t jend, float t32_in,
                                                   float t32_out, double t64, const JoinRows &rows,
                                                   bool has_prev, const uint4 *ccand, const uint32_t *cmask,
                                                   uint32_t leaf, uint4 rpc, uint32_t rpl, const GBox &rbox,
                                                   double fi = 0.0, uint32_t pi = 0)
{
    uint32_t n_sure = 0, n_pass = 0, band = 0, n_dup = 0, cols = 0;
    // the task's columns are the same for every lane: scalar loop bounds
    jbeg = __builtin_amdgcn_readfirstlane(jbeg);
    jend = __builtin_amdgcn_readfirstlane(jend);
    has_prev = __builtin_amdgcn_readfirstlane((uint32_t)has_prev) != 0u;   // (uniform: a scalar branch)
    const uint32_t tag = lane_id() << 16;
    const unsigned long long flag = BORDER ? kTagBorder : 0ull;
    const bool tested_check = has_prev && !s.store_all;   // wave-uniform
    // border rows in guard mode (pairs not all kept): the row leaf's summary
    // bit, and whether the row's candidates overflowed
    const bool settle = BORDER && USE32 && !s.store_all;
    const uint32_t rbit = leaf_bit(rpl & 0x7FFFFFFFu);
    const bool rovf = (rpl >> 31) != 0;
    // the stage's contacts as keys (kEmptyKey: a band entry that is not
    // within reach in FP64, or a contact that is not stored)
    auto key_of = [&](uint32_t e, bool valid) {
        const uint32_t rl = e >> 16, j = e & 0xFFFFu;
        uint4 ca = make_uint4(kNone, kNone, kNone, 0u);
        if (tested_check) ca = lane_read4(rpc, rl);   // the row object's candidates at prev_tick
        float rfx = 0.f, rfy = 0.f;
        if (USE32) {
            rfx = __shfl(fxi, (int)rl);
            rfy = __shfl(fyi, (int)rl);
        }
        if (!valid) return kEmptyKey;
        // every LDS read of the entry at once (one round trip)
        const uint32_t pa = rows.p[rl], pb = C.p[j];
        TileF cf = TileF{make_float2(0.f, 0.f), make_float2(0.f, 0.f)};
        float2 rpf = make_float2(0.f, 0.f);
        if (USE32) {
            cf = C.fp[j];
            if (has_prev) rpf = rows.pf[rl];
        }
        uint32_t la = kNone, lb = kNone;
        uint4 cb = make_uint4(kNone, kNone, kNone, 0u);
        if (tested_check) {
            la = rows.ppl[rl];
            lb = C.pl[j];
            cb = C.pc[j];
        }
        uint4 qc = make_uint4(kNone, kNone, kNone, 0u);
        uint32_t pl = 0;
        if (BORDER && !s.store_all) {
            qc = ccand[j];
            pl = rows.pl[rl];
        }
        bool in_band = false;
        if (USE32) {
            const float dx = rfx - cf.f.x;
            const float dy = rfy - cf.f.y;
            const float s32 = dist2_f32(dx, dy);
            in_band = !(s32 <= t32_in);
        }
        if (in_band) {
            band++;
            const uint32_t gj = C.g[j];
            const double ex = tv.xy[2 * (size_t)pa] - tv.sx[gj];
            const double ey = tv.xy[2 * (size_t)pa + 1] - tv.sy[gj];
            if (!(ex * ex + ey * ey <= t64)) return kEmptyKey;
            n_pass++;
        }
        unsigned long long k = pair_key(pa, pb) | flag;
        if (has_prev) {
            // within reach at prev_tick: decided on the FP32 offsets outside
            // the guard band, else by the FP64 test
            int dec = -1;
            if (USE32) {
                const float px = rpf.x - cf.pf.x;
                const float py = rpf.y - cf.pf.y;
                const float p32 = dist2_f32(px, py);
                dec = p32 <= t32_in ? 1 : (p32 > t32_out ? 0 : -1);   // NaN: -1
            }
            if (GEO) {
                dec = 1;   // every contact is kept and looked up (the scaled test is not symmetric)
            } else if (dec < 0) {
                const double *xp = tv.xy_prev;   // (by id: this path is rare)
                const double dx = xp[2 * (size_t)pa] - xp[2 * (size_t)pb];
                const double dy = xp[2 * (size_t)pa + 1] - xp[2 * (size_t)pb + 1];
                dec = dx * dx + dy * dy <= t64;
            }
            if (dec) k |= kTagPrev;
        }
        if ((k & kTagPrev) && tested_check) {
            // tested by the join of the tick before: both objects in one leaf
            // then, or one's lookup selected the other's leaf (candidate
            // lists that overflowed stored the pair, so they do not count)
            const bool tested = (la == lb && la != kNone) ||   // kNone: no leaf known at the tick before
                                (ca.w <= 3 && cb.w <= 3 &&
                                 (ca.x == lb || ca.y == lb || ca.z == lb || cb.x == la || cb.y == la || cb.z == la));
            if (tested) k |= kTagTested;
        }
        if (s.store_all || (k & kTagPrev)) return k;
        if (!BORDER) return kEmptyKey;   // untagged inside contact: not stored
        // Untagged border contact (row object p's lookup selected this leaf):
        // not stored unless a candidate list overflowed.  It is found twice
        // iff q's lookup selected p's leaf as well; that pair is counted as
        // a second copy once, by the side whose row leaf is the smaller.
        if ((pl >> 31) || qc.w > 3) return k;
        const uint32_t lp = pl & 0x7FFFFFFFu;
        const bool twice = qc.x == lp || qc.y == lp || qc.z == lp;
        n_dup += twice && lp < leaf;
        return kEmptyKey;
    };
    const uint32_t nblk = jend - jbeg, lane = lane_id();
    const uint32_t dmax = ROT ? (nblk - 1) / 2 + (nblk % 2 == 0 ? 1u : 0u) : 0u;
    const uint32_t lim = ROT ? (lane < nblk ? (nblk - 1) / 2 + ((nblk % 2 == 0 && lane < nblk / 2) ? 1u : 0u) : 0u)
                             : 0u;
    // ROT: every lim(i) is dmax or dmax - 1, so the lanes with a column at
    // rotation step d are those of the block for d < dmax, those with
    // lim(i) = dmax at d = dmax, and none past it
    const unsigned long long mrot_all = ROT ? __builtin_amdgcn_ballot_w64(lane < nblk) : 0ull;
    const unsigned long long mrot_top = ROT ? __builtin_amdgcn_ballot_w64(lane < nblk && lim == dmax) : 0ull;
    // guard mode: every hit is staged when every contact is kept and for
    // border rows that are not settled here; settled border rows stage their
    // sure hits also when their candidates overflowed
    // (uniform flags as all-or-nothing masks: the loop has no branch on them)
    const bool stage_all_hits = !settle && (BORDER || s.store_all);
    const unsigned long long mall = stage_all_hits ? ~0ull : 0ull;
    const unsigned long long msettle = settle ? ~0ull : 0ull;
    // sure hits staged whatever their previous offsets (the rows' previous
    // offsets are infinite without a tick before: never within reach)
    const unsigned long long mfix = mall | (__builtin_amdgcn_ballot_w64(rovf) & msettle);
    // one round: columns jb .. jb+kFlushRound-1 (below jlim; ROT: rotation steps)
    auto round = [&](uint32_t jb, uint32_t jlim) {
        TileF cfp[kFlushRound];
        double2 cx[kFlushRound];
        uint32_t cm[kFlushRound];
        uint32_t cj[kFlushRound];   // ROT: each lane's column
#pragma unroll
        for (uint32_t jj = 0; jj < kFlushRound; jj++) {
            uint32_t jc;
            if constexpr (ROT) {
                const uint32_t dd = min(jb + jj, dmax);
                uint32_t jr = (lane < nblk ? lane : 0u) + dd;
                jr = jr >= nblk ? jr - nblk : jr;
                jc = jbeg + jr;
                cj[jj] = jc;
            } else {
                // jb is in a group of kGroup slots that starts at a multiple
                // of kGroup below TILE, so jb + jj is a slot of the tile; a
                // slot past the tile's objects holds an infinite position
                // (geography, which reads a scale from the slot, clamps the
                // column and masks its lanes out)
                jc = GEO ? min(jb + jj, jlim - 1) : jb + jj;
            }
            if constexpr (USE32)
                cfp[jj] = C.fp[jc];
            else
                cx[jj] = C.xy[jc];
            if constexpr (BORDER && USE32) cm[jj] = cmask[jc];   // (used when settled)
        }
        // Every decision is a mask of lanes (a compare's own result), combined
        // with scalar operations; the vector unit only forms the distances.
        unsigned long long mst[kFlushRound];
#pragma unroll
        for (uint32_t jj = 0; jj < kFlushRound; jj++) {
            // the lanes with a candidate pair at this column
            unsigned long long mcand;
            if constexpr (ROT) {
                const uint32_t d = jb + jj;
                mcand = d < dmax ? mrot_all : (d == dmax ? mrot_top : 0ull);
            } else {
                // slots past the tile's objects hold infinite positions (no
                // distance test passes) up to the end of their group
                mcand = GEO ? (jb + jj < jlim ? ~0ull : 0ull) : ~0ull;
            }
            unsigned long long msure, mstage;
            if constexpr (USE32) {
                const float dx = fxi - cfp[jj].f.x;
                const float dy = fyi - cfp[jj].f.y;
                const float s32 = dist2_f32(dx, dy);
                msure = __builtin_amdgcn_ballot_w64(s32 <= t32_in) & mcand;
                const unsigned long long mhit = __builtin_amdgcn_ballot_w64(s32 <= t32_out) & mcand;
                // a sure hit is staged when it may be stored
                const float px = pfxi - cfp[jj].pf.x;
                const float py = pfyi - cfp[jj].pf.y;
                const float p32 = dist2_f32(px, py);
                // NaN (unbounded offsets) compares false: may be tagged
                unsigned long long mkeep = mfix | __builtin_amdgcn_ballot_w64(!(p32 > t32_out));
                if constexpr (BORDER) mkeep |= __builtin_amdgcn_ballot_w64((cm[jj] & rbit) != 0u) & msettle;
                mstage = mhit & (~msure | mkeep);
            } else {
                double ex = xi - cx[jj].x;
                double ey = yi - cx[jj].y;
                if constexpr (GEO) {
                    const uint32_t jc = ROT ? cj[jj] : min(jb + jj, jlim - 1);
                    const double f = BORDER ? fi : (pi < C.p[jc] ? fi : tile_geo(C.pc[jc]));
                    ey = ey / kDegPerMeterLat;
                    ex = ex / f;
                }
                msure = __builtin_amdgcn_ballot_w64(ex * ex + ey * ey <= t64) & mcand;
                mstage = (BORDER || s.store_all || has_prev) ? msure : 0ull;
            }
            n_sure += (uint32_t)__popcll(msure);
            mst[jj] = mstage;
        }
        // the round's stage writes together, after its distance tests
#pragma unroll
        for (uint32_t jj = 0; jj < kFlushRound; jj++) stage_push_mask(st, mst[jj], tag | (ROT ? cj[jj] : jb + jj));
        if (st.n > kStage - 64 * kFlushRound) stage_flush(st, s, key_of);
    };
    if constexpr (ROT) {
        for (uint32_t jb = 1; jb <= dmax; jb += kFlushRound) round(jb, 0u);
        cols = dmax;
    } else {
        // column groups of kGroup (jbeg is a multiple of kGroup); a group whose
        // box is out of reach of the rows' box holds no contact
        for (uint32_t g0 = jbeg; g0 < jend; g0 += kGroup) {
            const uint32_t g1 = min(g0 + kGroup, jend);
            // every lane holds the same boxes: the decision is taken as a scalar
            // (the boxes are in degrees without the scale: no skip with geography)
            if (!GEO && __builtin_amdgcn_readfirstlane(
                            (uint32_t)box_out_of_reach<USE32>(rbox, C.box[g0 / kGroup], t32_out, t64)))
                continue;
            cols += g1 - g0;
            for (uint32_t jb = g0; jb < g1; jb += kFlushRound) round(jb, g1);
        }
    }
    stage_flush(st, s, key_of);
    // n_sure is the wave's count (the same on every lane): lane 0 reports it
    return JoinCounts{lane == 0 ? n_sure : 0u, n_pass, band, n_dup, cols};
}

// A leaf of the join's current chunk of work.
struct JoinLeaf {
    uint32_t off, cnt;        // its objects in the sorted arrays
    uint32_t uoff, ucnt;      // its border units
    uint32_t use32;
    uint32_t node;            // the leaf
    double lx, ly;            // low corner (FP32 offsets of border units)
    double w;                 // side (sub-cells of the tile sort)
};

template <int TILE>
struct JoinLds {
    JoinTile<TILE> a, b;
    JoinRows rows[kBlock / 64];
    uint32_t stage[kBlock / 64][kStage];
    uint4 ccand[TILE];                   // candidate leaves of the objects of tile a (see DevState::cand)
    uint32_t cmask[TILE];                // and a 32-bit summary of them (cand_mask)
    JoinLeaf leaf;                       // the leaf being joined
    uint32_t hist[kSubCells];            // sub-cell counts of a tile being loaded
    uint32_t item;
    uint32_t next;                       // the next task of the tile pair being joined
};


// Upper-triangle task t of n row blocks -> (r <= q), row-major.
__device__ __forceinline__ void triangle_task(uint32_t t, uint32_t n, uint32_t &r, uint32_t &q)
{
    r = 0;
    while (t >= n - r) {
        t -= n - r;
        r++;
    }
    q = r + t;
}

// The refine stage (work queue over chunks of the leaf list):
//  - inside units (filter.cpp:107-109 with refinement.cpp:30, pid < q): every
//    unordered pair of a leaf's objects once;
//  - border units (pid, L, inside=false) (refinement.cpp:30 with !inside):
//    each object whose lookup selected L against every object of L.
// Work is split into 64x64 wave tasks (64 row lanes x 64 broadcast columns).
// Consecutive leaves of a chunk whose objects fit one tile are loaded
// together and their tasks are shared by all waves at once; a leaf larger
// than a tile is joined tile pair by tile pair.  The column loop only tests
// distances and stages (row lane, column) of each hit; when a task's stage
// is written out, one lane per contact forms the pair key, adds kTagPrev
// when the pair was also within reach at prev_tick (the only case that
// needs the previous table generation) and kTagBorder when it came from a
// border unit (the only case that can find a pair twice).
template <int TILE, bool GEO = false, class P>
__device__ __forceinline__ void join_body(const DevState s, const P blk)
{
    __shared__ JoinLds<TILE> lds;
    const double *xy = s.tp->xy;
    const double *xy_prev = s.tp->xy_prev;
    const bool has_prev = s.tp->has_prev != 0;
    const TickView tv{xy, xy_prev, s.tp->sx, s.tp->sy, s.tp->sx_old, s.tp->sy_old};
    const double t64 = s.t64;
    const uint32_t nh = s.ctr->n_heavy;
    const uint32_t nl = (s.ctr->error & kErrUnitCap) ? 0 : s.ctr->n_leaf_list + nh;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64), lane = lane_id();
    Stager st{contact_out(s), lds.stage[wave], 0, 0, s.chunk, 0};
    JoinRows &rows = lds.rows[wave];
    if (threadIdx.x == 0) lds.next = 0;   // (the first leaf's barriers come before any use)
    uint32_t n_in = 0, n_bd = 0;                   // sure contacts (per lane)
    uint32_t pass_in = 0, pass_bd = 0, band = 0;   // band entries within reach / re-tested (per lane)
    uint32_t dups = 0;                             // unstored pairs found twice (per lane)
    unsigned long long tests = 0;                  // candidate pairs of the lane's tasks
    unsigned long long evals = 0;                  // columns the wave evaluated (uniform)

    // One task: an inside task (row slot `row` of tile A) or a border task
    // (unit object `pid`, columns in tile a), against columns [jbeg, jend)
    // of tile C (rot: a diagonal task, each pair once); `leaf` is the leaf
    // being joined; vi false for an idle lane.
    auto run = [&](const JoinTile<TILE> &A, const JoinTile<TILE> &C, bool inside, uint32_t row, uint32_t pid,
                   uint4 ulf, bool vi, uint32_t jbeg, uint32_t jend, bool rot, bool use32, double lx,
                   double ly, uint32_t leaf) {
        double xi = 0.0, yi = 0.0;
        float fxi = 0.f, fyi = 0.f, pfxi = 0.f, pfyi = 0.f;
        uint32_t pi = 0;
        uint4 rpc = make_uint4(kNone, kNone, kNone, 0u);   // the row's candidates at prev_tick
        const bool want_pc = has_prev && !s.store_all;
        if (inside) {
            const uint32_t rr = vi ? row : jbeg;
            if (use32) {
                const TileF af = A.fp[rr];
                fxi = af.f.x;
                fyi = af.f.y;
                pfxi = af.pf.x;
                pfyi = af.pf.y;
            } else {
                const double2 a = A.xy[rr];
                xi = a.x;
                yi = a.y;
            }
            pi = A.p[rr];
            if (want_pc) rpc = A.pc[rr];
        } else {
            pi = pid;
            double pxi = 0.0, pyi = 0.0;
            if (vi) {
                // read at the object's sorted positions (units of a leaf come
                // from a few neighbouring leaves), not by id
                xi = tv.sx[ulf.w];
                yi = tv.sy[ulf.w];
                if (has_prev) {
                    // the sorted positions of the tick before, at the object's
                    // sorted position then (by id when that is not known)
                    pxi = ulf.z < s.n ? tv.sx_old[ulf.z] : tv.xy_prev[2 * (size_t)pid];
                    pyi = ulf.z < s.n ? tv.sy_old[ulf.z] : tv.xy_prev[2 * (size_t)pid + 1];
                }
                if (want_pc) rpc = cand_at(s.tp->cand_prev, ulf.z, s.n, s.tp->prev_tick);
            }
            if (use32) {
                fxi = (float)(xi - lx);
                fyi = (float)(yi - ly);
                if (has_prev) {
                    const double ox = pxi - lx, oy = pyi - ly;
                    pfxi = fabs(ox) <= s.pf_lim ? (float)ox : __builtin_nanf("");
                    pfyi = fabs(oy) <= s.pf_lim ? (float)oy : __builtin_nanf("");
                }
            }
        }
        if (!has_prev) pfxi = pfyi = __builtin_inff();   // (the tile's previous offsets are 0)
        if (!vi) {   // an idle lane: no column is within reach
            xi = yi = __builtin_inf();
            fxi = fyi = __builtin_inff();
        }
        // the rows' box: the boxes of the row block's column groups for an
        // inside task, else the units' positions (idle lanes left out)
        GBox rbox;
        if (inside) {
            const uint32_t g0 = (row - lane) / kGroup;   // the row block's first group, for every lane
            rbox = A.box[g0];
            for (uint32_t g = g0 + 1; g < g0 + 64 / kGroup && g < (uint32_t)TILE / kGroup; g++) {
                const GBox &b = A.box[g];
                rbox = GBox{fminf(rbox.lx, b.lx), fminf(rbox.ly, b.ly), fmaxf(rbox.hx, b.hx), fmaxf(rbox.hy, b.hy)};
            }
        } else {
            float x0 = __builtin_inff(), y0 = __builtin_inff(), x1 = -__builtin_inff(), y1 = -__builtin_inff();
            if (vi) {
                x0 = use32 ? fxi : f32_down(xi);
                x1 = use32 ? fxi : f32_up(xi);
                y0 = use32 ? fyi : f32_down(yi);
                y1 = use32 ? fyi : f32_up(yi);
            }
            for (int d = 32; d >= 1; d >>= 1) {
                x0 = fminf(x0, __shfl_xor(x0, d));
                y0 = fminf(y0, __shfl_xor(y0, d));
                x1 = fmaxf(x1, __shfl_xor(x1, d));
                y1 = fmaxf(y1, __shfl_xor(y1, d));
            }
            rbox = GBox{x0, y0, x1, y1};
        }
        // the same on every lane: held in scalar registers
        rbox.lx = __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(uint32_t, rbox.lx)));
        rbox.ly = __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(uint32_t, rbox.ly)));
        rbox.hx = __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(uint32_t, rbox.hx)));
        rbox.hy = __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(uint32_t, rbox.hy)));
        rows.p[lane] = pi;
        rows.pf[lane] = make_float2(pfxi, pfyi);
        if (want_pc) rows.ppl[lane] = inside ? A.pl[vi ? row : jbeg] : (vi ? ulf.y : kNone);
        uint32_t rpl = 0;
        if (!inside && !s.store_all && vi) {
            rpl = ulf.x;   // the unit object's leaf, bit 31 when its candidates overflowed
            rows.pl[lane] = rpl;
        }
        double fi = 0.0;   // geography: the row object's longitude scale
        if constexpr (GEO) {
            if (inside)
                fi = tile_geo(A.pc[vi ? row : jbeg]);
            else if (vi)
                fi = s.sgeo[ulf.w];
        }
        const JoinCounts jc = GEO ?
            ((inside && rot) ? join_columns<false, false, TILE, true, GEO>(
                                   st, s, tv, C, xi, yi, fxi, fyi, pfxi, pfyi, jbeg, jend, s.t32_in, s.t32_out, t64,
                                   rows, has_prev, lds.ccand, lds.cmask, leaf, rpc, rpl, rbox, fi, pi)
             : inside ? join_columns<false, false, TILE, false, GEO>(
                            st, s, tv, C, xi, yi, fxi, fyi, pfxi, pfyi, jbeg, jend, s.t32_in, s.t32_out, t64, rows,
                            has_prev, lds.ccand, lds.cmask, leaf, rpc, rpl, rbox, fi, pi)
                      : join_columns<false, true, TILE, false, GEO>(
                            st, s, tv, C, xi, yi, fxi, fyi, pfxi, pfyi, jbeg, jend, s.t32_in, s.t32_out, t64, rows,
                            has_prev, lds.ccand, lds.cmask, leaf, rpc, rpl, rbox, fi, pi)) :
            (inside && rot)
                ? (use32 ? join_columns<true, false, TILE, true>(st, s, tv, C, xi, yi, fxi, fyi, pfxi, pfyi,
                                                                jbeg, jend, s.t32_in, s.t32_out, t64, rows, has_prev,
                                                                lds.ccand, lds.cmask, leaf, rpc, rpl, rbox)
                         : join_columns<false, false, TILE, true>(st, s, tv, C, xi, yi, fxi, fyi, pfxi, pfyi,
                                                                 jbeg, jend, s.t32_in, s.t32_out, t64, rows,
                                                                 has_prev, lds.ccand, lds.cmask, leaf, rpc, rpl, rbox)) :
            use32 ? (inside ? join_columns<true, false>(st, s, tv, C, xi, yi, fxi, fyi, pfxi, pfyi, jbeg, jend,
                                                        s.t32_in, s.t32_out, t64, rows, has_prev, lds.ccand, lds.cmask, leaf,
                                                        rpc, rpl, rbox)
                            : join_columns<true, true>(st, s, tv, C, xi, yi, fxi, fyi, pfxi, pfyi, jbeg, jend,
                                                       s.t32_in, s.t32_out, t64, rows, has_prev, lds.ccand, lds.cmask, leaf,
                                                       rpc, rpl, rbox))
                  : (inside ? join_columns<false, false>(st, s, tv, C, xi, yi, fxi, fyi, pfxi, pfyi, jbeg, jend,
                                                         s.t32_in, s.t32_out, t64, rows, has_prev, lds.ccand, lds.cmask, leaf,
                                                         rpc, rpl, rbox)
                            : join_columns<false, true>(st, s, tv, C, xi, yi, fxi, fyi, pfxi, pfyi, jbeg, jend,
                                                        s.t32_in, s.t32_out, t64, rows, has_prev, lds.ccand, lds.cmask, leaf,
                                                        rpc, rpl, rbox));
        // the lane's candidate pairs: its columns, or on a diagonal task the
        // lim(i) columns of join_columns' rotation (each pair once)
        const uint32_t ncol = jend - jbeg;
        if (vi)
            tests += (inside && rot) ? (ncol - 1) / 2 + ((ncol % 2 == 0 && lane < ncol / 2) ? 1u : 0u) : ncol;
        evals += jc.cols;
        n_in += inside ? jc.sure : 0u;
        n_bd += inside ? 0u : jc.sure;
        pass_in += inside ? jc.pass : 0u;
        pass_bd += inside ? 0u : jc.pass;
        band += jc.band;
        dups += jc.dup;
    };

    // thread 0: takes the next leaf from the queue and leaves it in LDS
    auto fetch_leaf = [&]() {
        const uint32_t w = atomicAdd(&s.ctr->work_inside, 1u);
        lds.item = w;
        if (w < nl) {
            const uint32_t leaf = s.leaf_list[w < nh ? s.node_cap - 1 - w : w - nh];
            JoinLeaf L;
            L.off = s.offset[leaf];
            L.cnt = s.count[leaf];
            L.uoff = s.ustart[leaf];
            L.ucnt = s.ucount[leaf];
            const double2 lo = reinterpret_cast<const double2 *>(s.nodes + leaf)[1];
            const double2 hi = node_high(s.nodes, leaf);
            L.lx = lo.x;
            L.ly = lo.y;
            L.w = fmax(hi.x - lo.x, hi.y - lo.y);
            L.use32 = s.precision == GLINT_PRECISION_FP32_GUARD && (hi.x - lo.x) <= s.w32_max &&
                      (hi.y - lo.y) <= s.w32_max;
            L.node = leaf;
            lds.leaf = L;
        }
    };
    if (threadIdx.x == 0) fetch_leaf();
    __syncthreads();
    for (;;) {
        const uint32_t item = lds.item;
        const JoinLeaf F = lds.leaf;
        __syncthreads();   // read by every thread before thread 0 takes the next leaf
        if (item >= nl) break;
        // The next leaf is taken once this one's first tile is loaded: thread
        // 0 waits on the queue and the leaf's record while the other waves
        // start on the tasks (read after the last barrier of this leaf).
        if (F.cnt == 0) {
            if (threadIdx.x == 0) fetch_leaf();
            __syncthreads();
            continue;
        }
        const uint32_t ntile = (F.cnt + TILE - 1) / TILE;
        for (uint32_t I = 0; I < ntile; I++) {
            const uint32_t ni = min((uint32_t)TILE, F.cnt - I * TILE);
            load_tile_sorted(lds.a, s, F.off + I * TILE, ni, has_prev, F.use32, F.lx, F.ly, F.w, !s.store_all,
                             lds.ccand, lds.cmask, lds.hist);
            if (I == 0 && threadIdx.x == 0) fetch_leaf();
            for (uint32_t J = I; J < ntile; J++) {
                const uint32_t nj = min((uint32_t)TILE, F.cnt - J * TILE);
                const bool diag = (J == I);
                const JoinTile<TILE> &B = diag ? lds.a : lds.b;
                if (!diag)
                    load_tile_sorted(lds.b, s, F.off + J * TILE, nj, has_prev, F.use32, F.lx, F.ly, F.w, false,
                                     nullptr, nullptr, lds.hist);
                const uint32_t nri = (ni + 63) / 64, nrj = (nj + 63) / 64;
                const uint32_t ntask_in = diag ? nri * (nri + 1) / 2 : nri * nrj;
                // border tasks: 64 units against 64 columns of the tile, taken
                // first (every one a full 64 x 64 block), then the inside tasks;
                // each wave takes the next task when it is done with one, so
                // the tile pair ends when the work does, not when the wave with
                // the most column steps does
                const uint32_t nut = diag ? (F.ucnt + 63) / 64 * nri : 0;
                for (;;) {
                    uint32_t t = 0;
                    if (lane == 0) t = atomicAdd(&lds.next, 1u);
                    t = __builtin_amdgcn_readfirstlane(t);
                    if (t >= nut + ntask_in) break;
                    if (t >= nut) {
                        t -= nut;
                        uint32_t r, q;
                        if (diag) {
                            // the blocks against each other first, then each
                            // block against itself (half the pairs) at the end
                            const uint32_t noff = nri * (nri - 1) / 2;
                            if (t < noff) {
                                triangle_task(t, nri - 1, r, q);
                                q++;
                            } else {
                                r = q = t - noff;
                            }
                        } else {
                            r = t / nrj;
                            q = t % nrj;
                        }
                        const uint32_t row = r * 64 + lane;
                        const uint32_t jbeg = q * 64;
                        run(lds.a, B, true, row, 0, make_uint4(0u, 0u, 0u, 0u), row < ni, jbeg, min(jbeg + 64, nj),
                            diag && r == q, F.use32, F.lx, F.ly, F.node);
                    } else {
                        const uint32_t u = t / nri * 64 + lane, jbeg = t % nri * 64;
                        const bool vi = u < F.ucnt;
                        run(lds.a, lds.a, false, 0, vi ? s.unit_pid[F.uoff + u] : 0u,
                            vi ? s.unit_lf[F.uoff + u] : make_uint4(0u, 0u, 0u, 0u), vi, jbeg, min(jbeg + 64, ni),
                            false, F.use32, F.lx, F.ly, F.node);
                    }
                }
                __syncthreads();
                if (threadIdx.x == 0) lds.next = 0;   // read again only after the next tile's barriers
            }
        }
    }
    stage_close(st, s);
    // one atomic per wave for each counter
    uint32_t pi_ = pass_in + n_in, pb = pass_bd + n_bd, bd = band, du = dups;
    for (int d = 32; d >= 1; d >>= 1) {
        pi_ += __shfl_xor(pi_, d);
        pb += __shfl_xor(pb, d);
        bd += __shfl_xor(bd, d);
        du += __shfl_xor(du, d);
    }
    tests = wave_sum_u64(tests);
    if (lane == 0) {
        const uint32_t ci = pi_, cb = pb;
        if (tests) atomicAdd(&s.ctr->n_tests, tests);
        if (evals) atomicAdd(&s.ctr->n_evals, 64ull * evals);
        if (ci + cb) atomicAdd(&s.ctr->n_contacts, (unsigned long long)(ci + cb));
        if (cb) atomicAdd(&s.ctr->n_contacts_border, (unsigned long long)cb);
        if (bd) atomicAdd(&s.ctr->n_rechecks, (unsigned long long)bd);
        if (st.kept) atomicAdd(&s.ctr->n_stored, (unsigned long long)st.kept);
        if (du) atomicAdd(&s.ctr->n_dup, (unsigned long long)du);
    }
}

template <int TILE, bool GEO = false>
__global__ __launch_bounds__(kBlock) void k_join(DevState s) { join_body<TILE, GEO>(s, this_blk()); }
template <int TILE, bool GEO = false>
__global__ __launch_bounds__(kBlock, 4) void kf_join(const Many m)   // (4 waves per SIMD: at most 128 VGPRs)
{
    const Member w = member_of(m);
    join_body<TILE, GEO>(m.ss[w.i], w.blk);
}

// ------------------------------------------------------------ bucketing
//
// Two passes group the contact list by region.  Pass 1 (k_bucket_count, a
// scan, k_bucket_scatter) splits it by the high bits of the region index into
// at most 1024 coarse bins, sentinels going to one extra bin that is never
// written out; each block sorts 8192-entry tiles in LDS so its writes form
// runs.  Pass 2 (k_bucket_fine) splits every coarse bin by the low bits with
// one workgroup per bin and records the first contact of every region.

__device__ __forceinline__ uint32_t coarse_bin(const TableShape &sh, unsigned long long tagged)
{
    if (tagged == kEmptyKey) return sh.coarse;
    return region_of(tagged & kKeyBits, sh.bits) >> sh.lo_bits;
}

template <class P>
__device__ __forceinline__ void block_range(const P blk, unsigned long long n, unsigned long long &b0,
                                            unsigned long long &b1)
{
    const unsigned long long per = ((n + blk.count() - 1) / blk.count() + kBucketTile - 1) / kBucketTile * kBucketTile;
    b0 = min(n, per * blk.id());
    b1 = min(n, b0 + per);
}

template <class P>
__device__ __forceinline__ void bucket_count_body(const DevState s, const P blk)
{
    __shared__ uint32_t hist[kMaxCoarse + 1];
    const unsigned long long n = min(s.ctr->n_clist, s.clist_cap);
    const TableShape sh = table_shape(s);
    for (uint32_t k = threadIdx.x; k <= sh.coarse; k += kBlock) hist[k] = 0;
    __syncthreads();
    unsigned long long b0, b1;
    block_range(blk, n, b0, b1);
    const unsigned long long *in = contact_out(s);
    // four independent loads in flight per thread
    constexpr int kU = 4;
    for (unsigned long long i0 = b0 + threadIdx.x; i0 < b1; i0 += kU * kBlock) {
        unsigned long long key[kU];
#pragma unroll
        for (int q = 0; q < kU; q++) key[q] = i0 + q * kBlock < b1 ? in[i0 + q * kBlock] : 0ull;
#pragma unroll
        for (int q = 0; q < kU; q++)
            if (i0 + q * kBlock < b1) atomicAdd(&hist[coarse_bin(sh, key[q])], 1u);
    }
    __syncthreads();
    // bins past this tick's (the empty-slot bin is sh.coarse) count nothing
    for (uint32_t k = threadIdx.x; k <= s.coarse_max; k += kBlock)
        s.bcount[(size_t)k * blk.count() + blk.id()] = k <= sh.coarse ? hist[k] : 0u;
}

__global__ __launch_bounds__(kBlock) void k_bucket_count(DevState s) { bucket_count_body(s, this_blk()); }
__global__ __launch_bounds__(kBlock) void kf_bucket_count(const Many m)
{
    const Member w = member_of(m);
    bucket_count_body(m.ss[w.i], w.blk);
}

// Exclusive offsets of a scanned counter array (third kernel of the scan).
template <class P>
__device__ __forceinline__ void scan_write_body(const uint32_t *in, uint32_t n, const uint32_t *block_sums,
                                                uint32_t *out, const P blk)
{
    uint32_t base = blk.id() * kScanTile + threadIdx.x * kScanItems;
    uint32_t c[kScanItems];
    uint32_t v = 0;
    for (int k = 0; k < kScanItems; k++) {
        c[k] = base + k < n ? in[base + k] : 0;
        v += c[k];
    }
    uint32_t total;
    uint32_t ex = block_exclusive_scan(v, &total) + block_sums[blk.id()];
    for (int k = 0; k < kScanItems; k++) {
        if (base + k < n) out[base + k] = ex;
        ex += c[k];
    }
}

__global__ __launch_bounds__(kBlock) void k_scan_write(const uint32_t *in, uint32_t n, const uint32_t *block_sums,
                                                       uint32_t *out)
{
    scan_write_body(in, n, block_sums, out, this_blk());
}

__global__ __launch_bounds__(kBlock) void kf_scan_write_bins(const Many m)
{
    const Member w = member_of(m);
    const DevState &s = m.ss[w.i];
    scan_write_body(s.bcount, scan_length(s, kScanBins), s.block_sums, s.boff, w.blk);
}

template <class P>
__device__ __forceinline__ void bucket_scatter_body(const DevState s, const P blk)
{
    __shared__ unsigned long long tile[kBucketTile];
    __shared__ uint32_t hist[kMaxCoarse + 1], start[kMaxCoarse + 1], running[kMaxCoarse + 1];
    __shared__ uint32_t wsum[kBucketThreads / 64];
    const unsigned long long n = min(s.ctr->n_clist, s.clist_cap);
    const TableShape sh = table_shape(s);
    const uint32_t nb = sh.coarse + 1;
    const unsigned long long *in = contact_out(s);
    unsigned long long *out = s.clist2;
    unsigned long long b0, b1;
    block_range(blk, n, b0, b1);
    for (uint32_t k = threadIdx.x; k < nb; k += kBucketThreads) running[k] = s.boff[(size_t)k * blk.count() + blk.id()];
    for (unsigned long long t0 = b0; t0 < b1; t0 += kBucketTile) {
        const uint32_t tn = (uint32_t)min((unsigned long long)kBucketTile, b1 - t0);
        for (uint32_t k = threadIdx.x; k < nb; k += kBucketThreads) hist[k] = 0;
        __syncthreads();
        // each key's rank in its bin is the count it found (no second pass of atomics)
        unsigned long long keys[kBucketTile / kBucketThreads];
        uint32_t rk[kBucketTile / kBucketThreads];
#pragma unroll
        for (int q = 0; q < kBucketTile / kBucketThreads; q++) {   // every load before the first atomic
            const uint32_t i = threadIdx.x + q * kBucketThreads;
            keys[q] = i < tn ? in[t0 + i] : kEmptyKey;
        }
#pragma unroll
        for (int q = 0; q < kBucketTile / kBucketThreads; q++) {
            const uint32_t i = threadIdx.x + q * kBucketThreads;
            rk[q] = i < tn ? atomicAdd(&hist[coarse_bin(sh, keys[q])], 1u) : 0u;
        }
        __syncthreads();
        // exclusive scan of the tile histogram (nb <= 1025 bins, kPer per thread)
        {
            constexpr int kPer = (kMaxCoarse + 1 + kBucketThreads - 1) / kBucketThreads;
            uint32_t c[kPer], v = 0;
            for (int k = 0; k < kPer; k++) {
                uint32_t b = threadIdx.x * kPer + k;
                c[k] = b < nb ? hist[b] : 0;
                v += c[k];
            }
            uint32_t inc = v;
            for (int d = 1; d < 64; d <<= 1) {
                uint32_t y = __shfl_up(inc, d);
                if (lane_id() >= (uint32_t)d) inc += y;
            }
            if (lane_id() == 63) wsum[threadIdx.x / 64] = inc;
            __syncthreads();
            uint32_t ex = inc - v;
            for (uint32_t w = 0; w < threadIdx.x / 64; w++) ex += wsum[w];
            for (int k = 0; k < kPer; k++) {
                uint32_t b = threadIdx.x * kPer + k;
                if (b < nb) start[b] = ex;
                ex += c[k];
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kBucketTile / kBucketThreads; q++) {
            uint32_t i = threadIdx.x + q * kBucketThreads;
            if (i < tn) tile[start[coarse_bin(sh, keys[q])] + rk[q]] = keys[q];
        }
        __syncthreads();
        for (uint32_t p = threadIdx.x; p < tn; p += kBucketThreads) {
            unsigned long long key = tile[p];
            uint32_t b = coarse_bin(sh, key);
            if (b < sh.coarse) out[running[b] + (p - start[b])] = key;
        }
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < nb; k += kBucketThreads) running[k] += hist[k];
        __syncthreads();
    }
}

__global__ __launch_bounds__(kBucketThreads) void k_bucket_scatter(DevState s) { bucket_scatter_body(s, this_blk()); }
__global__ __launch_bounds__(kBucketThreads) void kf_bucket_scatter(const Many m)
{
    const Member w = member_of(m);
    bucket_scatter_body(m.ss[w.i], w.blk);
}

// Pass 2: one workgroup per coarse bin c splits its segment of clist2 by the
// low region bits into the generation's list and writes the first contact of
// each region.
// Tiles of the segment are sorted by fine bin in LDS so that the writes form
// runs (a per-item scatter into 512-2048 streams would write partial lines).
__device__ __forceinline__ uint32_t block_scan_bins(uint32_t *bins, uint32_t nbins, uint32_t *out_excl,
                                                    uint32_t add, uint32_t *wsum)
{
    // exclusive scan of bins[0..nbins) (nbins <= kMaxFine) into out_excl (+add)
    constexpr int kPer = kMaxFine / kBucketThreads;
    uint32_t cc[kPer], v = 0;
    for (int k = 0; k < kPer; k++) {
        uint32_t b = threadIdx.x * kPer + k;
        cc[k] = b < nbins ? bins[b] : 0;
        v += cc[k];
    }
    uint32_t inc = wave_inclusive_scan(v);
    if (lane_id() == 63) wsum[threadIdx.x / 64] = inc;
    __syncthreads();
    uint32_t ex = inc - v, total = 0;
    for (uint32_t w = 0; w < kBucketThreads / 64; w++) {
        if (w < threadIdx.x / 64) ex += wsum[w];
        total += wsum[w];
    }
    for (int k = 0; k < kPer; k++) {
        uint32_t b = threadIdx.x * kPer + k;
        if (b < nbins) out_excl[b] = add + ex;
        ex += cc[k];
    }
    __syncthreads();
    return total;
}

template <class P>
__device__ __forceinline__ void bucket_fine_body(const DevState s, const P blk)
{
    __shared__ unsigned long long tile[kFineTile];
    __shared__ uint32_t thist[kMaxFine], tstart[kMaxFine], gcur[kMaxFine];
    __shared__ uint32_t wsum[kBucketThreads / 64];
    const uint32_t G1 = s.bucket_blocks;
    const TableShape sh = table_shape(s);
    const uint32_t nf = 1u << sh.lo_bits;
    unsigned long long *out = s.tp->tab_cur.keys;
    uint32_t *rstart = s.tp->tab_cur.rstart;
    for (uint32_t c = blk.id(); c < sh.coarse; c += blk.count()) {
        const uint32_t seg0 = s.boff[(size_t)c * G1], seg1 = s.boff[(size_t)(c + 1) * G1];
        for (uint32_t k = threadIdx.x; k < nf; k += kBucketThreads) thist[k] = 0;
        __syncthreads();
        for (uint32_t i0 = seg0 + threadIdx.x; i0 < seg1; i0 += 4 * kBucketThreads) {
            unsigned long long key[4];
#pragma unroll
            for (int q = 0; q < 4; q++) key[q] = i0 + q * kBucketThreads < seg1 ? s.clist2[i0 + q * kBucketThreads] : 0ull;
#pragma unroll
            for (int q = 0; q < 4; q++)
                if (i0 + q * kBucketThreads < seg1) atomicAdd(&thist[region_of(key[q] & kKeyBits, sh.bits) & (nf - 1)], 1u);
        }
        __syncthreads();
        block_scan_bins(thist, nf, gcur, seg0, wsum);
        for (uint32_t k = threadIdx.x; k < nf; k += kBucketThreads) rstart[(size_t)c * nf + k] = gcur[k];
        __syncthreads();
        for (uint32_t t0 = seg0; t0 < seg1; t0 += kFineTile) {
            const uint32_t tn = min((uint32_t)kFineTile, seg1 - t0);
            for (uint32_t k = threadIdx.x; k < nf; k += kBucketThreads) thist[k] = 0;
            __syncthreads();
            unsigned long long keys[kFineTile / kBucketThreads];
            uint32_t fb[kFineTile / kBucketThreads], rk[kFineTile / kBucketThreads];
#pragma unroll
            for (int q = 0; q < kFineTile / kBucketThreads; q++) {   // every load before the first atomic
                const uint32_t i = threadIdx.x + q * kBucketThreads;
                keys[q] = i < tn ? s.clist2[t0 + i] : 0ull;
            }
#pragma unroll
            for (int q = 0; q < kFineTile / kBucketThreads; q++) {
                const uint32_t i = threadIdx.x + q * kBucketThreads;
                fb[q] = region_of(keys[q] & kKeyBits, sh.bits) & (nf - 1);
                rk[q] = i < tn ? atomicAdd(&thist[fb[q]], 1u) : 0u;   // its rank in the region
            }
            __syncthreads();
            block_scan_bins(thist, nf, tstart, 0, wsum);
            __syncthreads();
#pragma unroll
            for (int q = 0; q < kFineTile / kBucketThreads; q++) {
                const uint32_t i = threadIdx.x + q * kBucketThreads;
                if (i < tn) tile[tstart[fb[q]] + rk[q]] = keys[q];
            }
            __syncthreads();
            for (uint32_t p = threadIdx.x; p < tn; p += kBucketThreads) {
                const unsigned long long key = tile[p];
                const uint32_t f = region_of(key & kKeyBits, sh.bits) & (nf - 1);
                out[gcur[f] + (p - tstart[f])] = key;
            }
            __syncthreads();
            for (uint32_t k = threadIdx.x; k < nf; k += kBucketThreads) gcur[k] += thist[k];
            __syncthreads();
        }
    }
    if (blk.id() == 0 && threadIdx.x == 0) {
        // the generation's entry count and region bits, where the next tick
        // (and the readers of a generation) find them whatever its shape
        const uint32_t total = s.boff[(size_t)sh.coarse * G1];
        rstart[(size_t)sh.coarse * nf] = total;
        rstart[s.rs_total] = total;
        rstart[s.rs_total + 1] = sh.bits;
        s.ctr->tab_bits = sh.bits;
    }
}

__global__ __launch_bounds__(kBucketThreads) void k_bucket_fine(DevState s) { bucket_fine_body(s, this_blk()); }
__global__ __launch_bounds__(kBucketThreads) void kf_bucket_fine(const Many m)
{
    const Member w = member_of(m);
    bucket_fine_body(m.ss[w.i], w.blk);
}

// ------------------------------------------------------------ table update
//
// One workgroup per region: the upsert of every contact of the tick
// (refinement.cpp:36-55 with set semantics) and the evaluation of the
// previous generation (meeting.cpp:134-155) in LDS.  A pair of the previous
// generation keeps its start tick if it is seen again; a pair not seen again
// has ended at prev_tick and is emitted iff tick - start >= min + 1.
//
// A generation is not a slot table in HBM but the tick's bucketed contact
// list itself, with one start tick per entry.  Only a contact tagged as
// within reach at prev_tick can continue a run, and only one found through
// a border unit can be a second copy of a pair (an inside pair is found once,
// by one leaf).  The workgroup therefore hashes just those contacts of its
// region into LDS (removing second copies), streams the previous
// generation's entries of the region past the hash (a hit continues a run,
// a miss has ended), and writes every contact's start beside it.  HBM
// traffic is 12 bytes per entry of each generation plus the key read; no
// slot is ever empty.
constexpr uint32_t kNoSlot = 0xFFFFFFFFu;

template <bool KV, int NT>
struct TableLds {
    unsigned long long hk[kRegion];   // hashed contacts of the tick (kEmptyKey: free)
    uint32_t hc[kRegion];             // their list index
    uint32_t hs[kRegion];             // their start tick
    uint32_t ho[KV ? kRegion : 1];    // list index of the old entry they continue (kNoSlot: none)
    glint_meeting stage[NT / 64][kMeetStage];
    float4 sbox[KV ? NT / 64 : 1][KV ? kMeetStage : 1];   // the staged meetings' boxes
};

// Writes a wave's staged meetings out with one reservation; with the
// key-value output, those whose box lies in the map box (geometry.cu:665-670,
// box_contain geometry.cu:53-58) are also logged with their boxes, with one
// more reservation.
template <bool KV, int NT>
__device__ __forceinline__ void table_flush_stage(const DevState &s, glint_meeting *st, const float4 *sb,
                                                  uint32_t &nst)
{
    wave_sync();
    const uint32_t lane = lane_id();
    unsigned long long b = 0;
    if (lane == 0) b = atomicAdd(&s.ctr->n_meet, (unsigned long long)nst);
    b = __shfl(b, 0);
    for (uint32_t j = lane; j < nst; j += 64) {
        if (b + j < s.meet_cap)
            s.meet[b + j] = st[j];
        else
            atomicOr(&s.ctr->error, kErrMeetCap);
    }
    if constexpr (KV) {
        auto in_map = [&](uint32_t j) {
            if (j >= nst) return false;
            const float4 bx = sb[j];
            return bx.x >= s.lx && bx.z <= s.hx && bx.y >= s.ly && bx.w <= s.hy;
        };
        uint32_t total = 0;
        for (uint32_t j0 = 0; j0 < nst; j0 += 64) total += (uint32_t)__popcll(__ballot(in_map(j0 + lane)));
        if (total) {
            unsigned long long k = 0;
            if (lane == 0) k = atomicAdd(&s.ctr->n_kvlog, (unsigned long long)total);
            k = __shfl(k, 0);
            for (uint32_t j0 = 0; j0 < nst; j0 += 64) {
                const uint32_t j = j0 + lane;
                const bool put = in_map(j);
                const unsigned long long m = __ballot(put);
                if (put) {
                    const unsigned long long at = k + __popcll(m & ((1ull << lane) - 1ull));
                    if (at < s.kvlog_cap) {
                        KvLogEntry e;
                        e.pid_a = st[j].pid_a;
                        e.pid_b = st[j].pid_b;
                        e.start = st[j].start;
                        e.end = st[j].end;
                        e.box = sb[j];
                        s.kvlog[at] = e;
                    } else {
                        atomicOr(&s.ctr->error, kErrKvCap);
                    }
                }
                k += (unsigned long long)__popcll(m);
            }
        }
    }
    wave_sync();
    nst = 0;
}

// Hashes contact c (tagged or border) of the tick; returns its slot, or
// kNoSlot for a second copy of a pair (its start is marked) or a full region.
template <bool KV, int NT>
__device__ __forceinline__ uint32_t table_put(TableLds<KV, NT> &L, const DevState &s, const TableGen &tcur,
                                              unsigned long long key, uint32_t c, uint32_t tick, uint32_t bits)
{
    const uint32_t h = slot_of(key, bits);
    for (uint32_t p = 0; p < kRegion; p++) {
        const uint32_t q = (h + p) & (kRegion - 1);
        const unsigned long long prev = atomicCAS(&L.hk[q], kEmptyKey, key);
        if (prev == kEmptyKey) {
            L.hc[q] = c;
            L.hs[q] = tick;
            if constexpr (KV) L.ho[q] = kNoSlot;
            return q;
        }
        if (prev == key) {
            tcur.starts[c] = kDupStart;
            return kNoSlot;
        }
    }
    atomicOr(&s.ctr->error, kErrTableFull);
    return kNoSlot;
}

template <bool KV, int NT>
__device__ __forceinline__ uint32_t table_find(const TableLds<KV, NT> &L, unsigned long long key, uint32_t bits)
{
    const uint32_t h = slot_of(key, bits);
    for (uint32_t p = 0; p < kRegion; p++) {
        const uint32_t q = (h + p) & (kRegion - 1);
        const unsigned long long k = L.hk[q];
        if (k == key) return q;
        if (k == kEmptyKey) return kNoSlot;
    }
    return kNoSlot;
}

// One entry of the previous generation: continues the run of the tick's
// contact of the same pair, or has ended (emitted if long enough).
template <bool KV, int NT>
__device__ __forceinline__ void table_old(TableLds<KV, NT> &L, const DevState &s, bool valid,
                                          unsigned long long tagged, uint32_t start, uint32_t i, uint32_t tick,
                                          uint32_t last, glint_meeting *st, float4 *sb, uint32_t &nst, uint32_t bits)
{
    valid = valid && start != kDupStart;
    const unsigned long long key = tagged & kKeyBits;
    bool emit = false;
    if (valid) {
        const uint32_t q = table_find(L, key, bits);
        if (q != kNoSlot) {
            L.hs[q] = start;
            if constexpr (KV) L.ho[q] = i;
        } else {
            emit = tick - start >= s.min_meet + 1;
        }
    }
    const uint32_t lane = lane_id();
    const unsigned long long m = __ballot(emit);
    if (emit) {
        glint_meeting mt;
        mt.pid_a = (uint32_t)(key >> 32);
        mt.pid_b = (uint32_t)key;
        mt.start = start;
        mt.end = last;
        const uint32_t at = nst + __popcll(m & ((1ull << lane) - 1ull));
        st[at] = mt;
        if constexpr (KV) sb[at] = s.tp->mbr_old[i];   // the run's box, logged when the stage is written
    }
    nst += (uint32_t)__popcll(m);
    if (nst > kMeetStage - 64) table_flush_stage<KV, NT>(s, st, sb, nst);
}

// Start (and, with the key-value output, box) of contact c, whose hash slot
// is q (kNoSlot: a contact that was not hashed, new at this tick).  A tagged
// contact (within reach at the tick before) with no entry in the previous
// generation either is new or was a contact of the tick before that was not
// stored; the join set kTagTested when the join of the tick before tested
// the pair, which tells which.
// The pair's two positions at this tick and at the tick before (key-value
// output only; both in one 16-byte gather per object).
struct PairPos {
    float4 a, b;
};

template <bool KV>
__device__ __forceinline__ PairPos pair_pos(const DevState &s, unsigned long long tagged)
{
    PairPos p{};
    if constexpr (KV) {
        p.a = s.pos32[(uint32_t)((tagged & kKeyBits) >> 32)];
        p.b = s.pos32[(uint32_t)(tagged & kKeyBits)];
    }
    return p;
}

template <bool KV, int NT>
__device__ __forceinline__ void table_write(const TableLds<KV, NT> &L, const DevState &s, const TableGen &tcur,
                                            unsigned long long tagged, uint32_t c, uint32_t q, uint32_t tick,
                                            uint32_t last, const PairPos pp)
{
    uint32_t start = q == kNoSlot ? tick : L.hs[q];
    bool was_unstored = false;
    if (start == tick && (tagged & kTagPrev) && (tagged & kTagTested) && !s.store_all) {
        start = last;
        was_unstored = true;
    }
    tcur.starts[c] = start;
    if constexpr (KV) {
        const float4 pa = pp.a, pb = pp.b;
        float4 m = make_float4(fminf(pa.x, pb.x), fminf(pa.y, pb.y), fmaxf(pa.x, pb.x), fmaxf(pa.y, pb.y));
        if (q != kNoSlot && L.ho[q] != kNoSlot) {
            const float4 o = s.tp->mbr_old[L.ho[q]];
            m = make_float4(fminf(m.x, o.x), fminf(m.y, o.y), fmaxf(m.z, o.z), fmaxf(m.w, o.w));
        } else if (was_unstored) {   // the run's box so far: both points at the tick before
            m = make_float4(fminf(m.x, fminf(pa.z, pb.z)), fminf(m.y, fminf(pa.w, pb.w)),
                            fmaxf(m.z, fmaxf(pa.z, pb.z)), fmaxf(m.w, fmaxf(pa.w, pb.w)));
        }
        s.tp->mbr_cur[c] = m;
    }
}

// With the key-value output on, every pair also carries the float box of
// both objects over its run (geometry.cu:469-494): the box of a new pair is
// its two points at this tick, a continuing pair adds them to its old box.
// min/max of the rounded coordinates equals the rounded min/max, so the box
// does not depend on the order of the updates.  Boxes stay in HBM beside the
// list entries; the points come from a float copy of the positions that
// fits the last-level cache.
// RESHAPED: the variant for a tick whose table shape differs from the
// previous generation's; the host launches both variants and the one that
// does not apply returns at once (the common variant then needs no registers
// for the mapping between shapes).
template <bool KV, int NT, bool RESHAPED, class P>
__device__ __forceinline__ void table_body(const DevState s, const P blk)
{
    __shared__ TableLds<KV, NT> L;
    const TableGen told = s.tp->tab_old, tcur = s.tp->tab_cur;
    const uint32_t tick = s.tp->tick, last = s.tp->prev_tick;
    // regions of this tick, and of the previous generation (its shape came
    // from its own tick's count): region r holds the pairs whose hash has r
    // in its high `bits` bits, so an old region of fewer bits holds the
    // pairs of 2^(bits - pbits) regions (filtered here by the hash), and a
    // region of this tick holds those of 2^(pbits - bits) consecutive old ones
    const uint32_t bits = table_shape(s).bits, pbits_gen = told.rstart[s.rs_total + 1];
    if ((pbits_gen != bits) != RESHAPED) return;
    const uint32_t pbits = RESHAPED ? pbits_gen : bits;
    const uint32_t nreg = 1u << bits;
    const bool filter_old = pbits < bits;
    // old region range of region rr: rstart[(rr >> dn) << up] .. rstart[((rr >> dn) + 1) << up]
    const uint32_t up = pbits > bits ? pbits - bits : 0u, dn = bits > pbits ? bits - pbits : 0u;
    glint_meeting *st = L.stage[threadIdx.x / 64];
    float4 *sb = KV ? L.sbox[threadIdx.x / 64] : nullptr;
    uint32_t nst = 0, dups = 0;
    const bool void_tick = (s.ctr->error & kErrContactCap) != 0;
    // the previous generation's entries and the tick's contacts of the next
    // region are loaded into registers while the current region is processed;
    // the bounds of the region after that are loaded one region earlier still,
    // so that the entry loads do not wait for them
    constexpr int kPer = kRegion / NT;
    unsigned long long pk[kPer], nk[kPer], nkn[kPer];
    uint32_t ps[kPer], ns[kPer];
    uint32_t o0 = 0, o1 = 0, c0n = 0, c1n = 0;
    uint32_t bo0 = 0, bo1 = 0, bc0 = 0, bc1 = 0;
    auto bounds = [&](uint32_t rr) {
        if (rr >= nreg) return;
        bo0 = told.rstart[(rr >> dn) << up];
        bo1 = told.rstart[((rr >> dn) + 1) << up];
        bc0 = void_tick ? 0 : tcur.rstart[rr];
        bc1 = void_tick ? 0 : tcur.rstart[rr + 1];
    };
    auto prefetch = [&](uint32_t rr) {
        o0 = bo0;
        o1 = bo1;
        c0n = bc0;
        c1n = bc1;
        bounds(rr + blk.count());
#pragma unroll
        for (int q = 0; q < kPer; q++) {
            const uint32_t i = o0 + threadIdx.x + q * NT;
            pk[q] = i < o1 ? told.keys[i] : kEmptyKey;
            ps[q] = i < o1 ? told.starts[i] : kDupStart;
            const uint32_t c = c0n + threadIdx.x + q * NT;
            nkn[q] = c < c1n ? tcur.keys[c] : 0ull;
        }
    };
    bounds(blk.id());
    if (blk.id() < nreg) prefetch(blk.id());
    for (uint32_t r = blk.id(); r < nreg; r += blk.count()) {
        for (uint32_t k = threadIdx.x; k < kRegion; k += NT) L.hk[k] = kEmptyKey;
        const uint32_t c0 = c0n, c1 = c1n;
#pragma unroll
        for (int q = 0; q < kPer; q++) nk[q] = nkn[q];
        // the region's pair positions (key-value output), in flight while
        // the region is hashed and the previous generation is read
        PairPos pp[kPer];
#pragma unroll
        for (int q = 0; q < kPer; q++) pp[q] = c0 + threadIdx.x + q * NT < c1 ? pair_pos<KV>(s, nk[q]) : PairPos{};
        lds_barrier();
        // the tick's contacts that can continue a run or be a second copy
#pragma unroll
        for (int q = 0; q < kPer; q++) {
            const uint32_t c = c0 + threadIdx.x + q * NT;
            ns[q] = kNoSlot;
            if (c < c1 && (nk[q] & (kTagPrev | kTagBorder))) {
                ns[q] = table_put(L, s, tcur, nk[q] & kKeyBits, c, tick, bits);
                dups += ns[q] == kNoSlot;
                if (ns[q] == kNoSlot) ns[q] = kNoSlot - 1;   // nothing left to write
            }
        }
        for (uint32_t c = c0 + kRegion + threadIdx.x; c < c1; c += NT) {   // regions of more than kRegion contacts
            const unsigned long long k = tcur.keys[c];
            if (k & (kTagPrev | kTagBorder)) dups += table_put(L, s, tcur, k & kKeyBits, c, tick, bits) == kNoSlot;
        }
        lds_barrier();
        // the previous generation's entries of the region
        if (filter_old) {   // a table that grew since the previous tick: only this region's pairs
#pragma unroll
            for (int q = 0; q < kPer; q++) {
                const uint32_t i = o0 + threadIdx.x + q * NT;
                const bool mine = region_of(pk[q] & kKeyBits, bits) == r;
                table_old(L, s, i < o1 && mine, pk[q], ps[q], i, tick, last, st, sb, nst, bits);
            }
        } else {
#pragma unroll
            for (int q = 0; q < kPer; q++) {
                const uint32_t i = o0 + threadIdx.x + q * NT;
                table_old(L, s, i < o1, pk[q], ps[q], i, tick, last, st, sb, nst, bits);
            }
        }
        for (uint32_t b = o0 + kRegion; b < o1; b += NT) {   // regions of more than kRegion entries
            const uint32_t i = b + threadIdx.x;
            const unsigned long long k = i < o1 ? told.keys[i] : 0ull;
            const bool v = i < o1 && (!filter_old || region_of(k & kKeyBits, bits) == r);
            table_old(L, s, v, k, v ? told.starts[i] : kDupStart, i, tick, last, st, sb, nst, bits);
        }
        const uint32_t rn = r + blk.count();
        if (rn < nreg) prefetch(rn);
        lds_barrier();
        // starts (and boxes) of every contact of the region
#pragma unroll
        for (int q = 0; q < kPer; q++) {
            const uint32_t c = c0 + threadIdx.x + q * NT;
            if (c < c1 && ns[q] != kNoSlot - 1) table_write(L, s, tcur, nk[q], c, ns[q], tick, last, pp[q]);
        }
        for (uint32_t c = c0 + kRegion + threadIdx.x; c < c1; c += NT) {
            const unsigned long long k = tcur.keys[c];
            uint32_t q = kNoSlot;
            if (k & (kTagPrev | kTagBorder)) {
                q = table_find(L, k & kKeyBits, bits);
                if (q == kNoSlot || L.hc[q] != c) continue;   // a second copy (already marked)
            }
            table_write(L, s, tcur, k, c, q, tick, last, pair_pos<KV>(s, k));
        }
        lds_barrier();
    }
    if (nst) table_flush_stage<KV, NT>(s, st, sb, nst);
    // distinct pairs of the tick: contacts less second copies
    for (int d = 32; d >= 1; d >>= 1) dups += __shfl_xor(dups, d);
    if (lane_id() == 0 && dups) atomicAdd(&s.ctr->n_dup, (unsigned long long)dups);
}

template <bool KV, int NT, bool RESHAPED>
__global__ __launch_bounds__(NT) void k_table(DevState s) { table_body<KV, NT, RESHAPED>(s, this_blk()); }
template <bool KV, int NT, bool RESHAPED>
__global__ __launch_bounds__(NT) void kf_table(const Many m)
{
    const Member w = member_of(m);
    table_body<KV, NT, RESHAPED>(m.ss[w.i], w.blk);
}

// ------------------------------------------------------------ host side

inline uint32_t grid_for(uint64_t n, uint32_t cap = 8192)
{
    uint64_t g = (n + kBlock - 1) / kBlock;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (uint32_t)g;
}

inline uint64_t next_pow2(uint64_t v)
{
    uint64_t p = 1;
    while (p < v) p <<= 1;
    return p;
}

// Largest double s with sqrt(s) <= r (correctly rounded host sqrt).
double sqrt_threshold(double r)
{
    double s = r * r;
    while (std::sqrt(s) > r) s = std::nextafter(s, 0.0);
    for (;;) {
        double n = std::nextafter(s, INFINITY);
        if (std::sqrt(n) <= r)
            s = n;
        else
            break;
    }
    return s;
}

enum KernelId {
    K_FILL_COUNTS, K_RESET, K_ASSIGN, K_SCAN_REDUCE, K_SCAN_TOP, K_SCAN_OBJECTS, K_SCATTER, K_FILL_UCOUNT,
    K_LOOKUP, K_SCAN_UNITS, K_UNIT_SCATTER, K_JOIN, K_BUCKET_COUNT, K_SCAN_WRITE, K_BUCKET_SCATTER, K_BUCKET_FINE,
    K_TABLE, K_TABLE_RESHAPED, K_COLLECT, K_SPLIT, K_MERGE, K_SCHEMA_DONE, K_TICK_PARAMS, K_POS32, K_LOG_APPEND, K_GRID_CELLS,
    K_GRID_PYRAMID, K_GRID_TOP, K_GRID_FINE_ZERO, K_GRID_FINE, K_GRID_LEVEL, K_REGIONS, K_COUNT
};
const char *const kKernelNames[K_COUNT] = {
    "k_fill_u32", "k_reset_tick", "k_assign", "k_scan_reduce", "k_scan_top", "k_scan_objects", "k_scatter",
    "k_fill_ucount", "k_lookup", "k_scan_units", "k_unit_scatter", "k_join", "k_bucket_count", "k_scan_write",
    "k_bucket_scatter", "k_bucket_fine", "k_table", "k_table_reshaped", "k_collect", "k_split", "k_merge", "k_schema_done",
    "k_advance_tick", "k_pos32", "k_log_append", "k_grid_cells", "k_grid_pyramid", "k_grid_top",
    "k_grid_fine_zero", "k_grid_fine", "k_grid_level", "k_regions"};

// Per-kernel HIP-event timing on the context's stream (optional).
struct KernelProfiler {
    bool on = false;
    std::vector<hipEvent_t> pool;
    struct Pending {
        int id;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
    double total_ms[K_COUNT] = {};
    unsigned long long launches[K_COUNT] = {};
    hipEvent_t get()
    {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }
    void resolve()
    {
        for (const Pending &p : pending) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
                total_ms[p.id] += ms;
                launches[p.id]++;
            }
            pool.push_back(p.a);
            pool.push_back(p.b);
        }
        pending.clear();
        (void)hipGetLastError();
    }
    void clear()
    {
        resolve();
        for (hipEvent_t e : pool) (void)hipEventDestroy(e);
        pool.clear();
    }
};

}  // namespace

struct glint_ctx {
    glint_config cfg{};
    int device = 0;
    hipStream_t stream = nullptr;
    DevState s{};
    TableGen tab[2] = {};
    float4 *mbr[2] = {nullptr, nullptr};    // pair boxes of each generation (key-value output)
    KvEngine *kv = nullptr;                 // sorted key-value output (null when off)
    double *hist[2] = {nullptr, nullptr};   // positions of the last two ticks (generation parity)
    int cur_gen = 0;            // table generation filled at the current tick
    uint4 *cand[2] = {};        // candidate-list buffers: cand[cur_gen] is written at the current tick
    bool built = false;
    bool has_positions = false;
    bool partitioned = false;
    bool refined = false;
    bool evaluated = false;
    bool has_prev = false;
    uint32_t tick = 0;
    uint32_t prev_tick = 0;
    uint32_t scan_blocks = 0;
    uint32_t join_blocks = 1024;   // workgroups of the join
    uint32_t fused_join_blocks = 1024;   // its share of a fused join: about one per 1024 objects, 64 to 1024
    uint32_t table_grid = 4096;    // workgroups of the table update (persistent over the regions)
    uint32_t fine_grid = 1024;     // workgroups of the fine bucketing pass (persistent over the bins)
    double *own_xy = nullptr;
    const double *last_xy = nullptr;   // positions of the current tick (caller's or own_xy)
    bool resident = false;             // own_xy holds the resident positions (update ingestion)
    Counters *h_ctr = nullptr;  // pinned
    std::vector<void *> allocs;
    std::string err;
    uint32_t inject = 0;
    hipEvent_t ev[6] = {};
    glint_tick_stats last{};
    KernelProfiler prof;
    hipGraphExec_t graph_exec = nullptr;
    unsigned long long *stamps = nullptr;
    uint32_t stamp_cap = 0;
    double wall_khz = 0;
};

// Independent scenarios ticked together: one context each, grouped by
// kernel variant; every step of a tick is one fused launch per group on the
// batch's stream, optionally captured for graph_ticks ticks in one graph.
struct glint_batch {
    std::vector<glint_ctx *> ctx;
    std::vector<std::vector<uint32_t>> groups;   // contexts fused into one launch per step
    std::vector<DevState *> group_states;        // each group's member states (device memory)
    hipStream_t origin = nullptr;
    hipEvent_t after = nullptr;
    std::vector<hipEvent_t> ready;
    hipGraphExec_t exec = nullptr;
    uint32_t graph_ticks = 0;
    std::string err;
};

namespace {

std::string g_create_error;

enum InjectBits : uint32_t {
    kInjPartition = 1,
    kInjRefine = 2,
    kInjEvaluate = 4,
    kInjSchema = 8,
    kInjAlloc = 16,
};

glint_status hip_fail(glint_ctx *c, hipError_t e, const char *what)
{
    if (c) c->err = std::string(what) + ": " + hipGetErrorString(e);
    return GLINT_ERR_HIP;
}

#define HIP_TRY(c, call)                                      \
    do {                                                      \
        hipError_t e_ = (call);                               \
        if (e_ != hipSuccess) return hip_fail((c), e_, #call); \
    } while (0)

template <typename T>
glint_status dalloc(glint_ctx *c, T **p, uint64_t count)
{
    if (count == 0) count = 1;
    void *v = nullptr;
    hipError_t e = hipMalloc(&v, count * sizeof(T));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        c->err = "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed";
        return GLINT_ERR_OUT_OF_MEMORY;
    }
    c->allocs.push_back(v);
    *p = (T *)v;
    return GLINT_OK;
}

glint_status status_from_flags(glint_ctx *c, uint32_t f)
{
    if (!f) return GLINT_OK;
    if (f & kErrBadInput) {
        c->err = "a position is NaN or outside the root box";
        return GLINT_ERR_BAD_INPUT;
    }
    if (f & kErrGridEdge) {
        c->err = "a position lies on the high y edge of the root box, which has no cell in the 16384^2 grid "
                 "of the grid rebuild";
        return GLINT_ERR_BAD_INPUT;
    }
    if (f & kErrNodeCap) {
        c->err = "quadtree node pool exhausted (raise node_capacity)";
        return GLINT_ERR_NODE_CAPACITY;
    }
    if (f & kErrTableFull) {
        c->err = "meeting table full (raise num_meeting_buckets)";
        return GLINT_ERR_TABLE_FULL;
    }
    if (f & kErrStack) {
        c->err = "inconsistent quadtree (a node without its parent link)";
        return GLINT_ERR_HIP;
    }
    if (f & kErrUnitCap) {
        c->err = "border check units exceed unit_capacity";
        return GLINT_ERR_UNIT_CAPACITY;
    }
    if (f & kErrKvCap) {
        c->err = "key-value log exceeds kv_capacity (flush more often or raise kv_capacity)";
        return GLINT_ERR_OUTPUT_CAPACITY;
    }
    c->err = "meeting or contact output exceeds its capacity";
    return GLINT_ERR_OUTPUT_CAPACITY;
}

glint_status read_counters(glint_ctx *c)
{
    HIP_TRY(c, hipMemcpyAsync(c->h_ctr, c->s.ctr, sizeof(Counters), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return GLINT_OK;
}

void set_tick_params(glint_ctx *c, const double *xy)
{
    const int g = c->cur_gen;
    k_set_tick<<<1, 1, 0, c->stream>>>(c->s.tp, xy, c->tab[1 - g], c->tab[g], c->mbr[1 - g], c->mbr[g],
                                       c->hist[1 - g], c->hist[g], c->cand[1 - g], c->cand[g], c->s.sxb[1 - g],
                                       c->s.syb[1 - g], c->s.sxb[g], c->s.syb[g], c->s.sleafb[1 - g],
                                       c->s.sleafb[g], c->tick,
                                       c->has_prev ? c->prev_tick : c->tick, c->has_prev ? 1u : 0u);
}

// Launch through the profiler: HIP events around the kernel when enabled.
template <typename F>
inline void launch(glint_ctx *c, int id, F &&f)
{
    if (!c->prof.on) {
        f();
        return;
    }
    hipEvent_t a = c->prof.get(), b = c->prof.get();
    if (!a || !b) {
        f();
        return;
    }
    (void)hipEventRecord(a, c->stream);
    f();
    (void)hipEventRecord(b, c->stream);
    c->prof.pending.push_back({id, a, b});
}

glint_status build_tree(glint_ctx *c)
{
    DevState &s = c->s;
    const uint32_t gn = grid_for(s.n), gnode = grid_for(s.node_cap);
    k_init_tree<<<gnode, kBlock, 0, c->stream>>>(s);
    if (c->cfg.dynamic_schema == GLINT_SCHEMA_GRID_REBUILD) {   // every tick builds its own tree
        HIP_TRY(c, hipGetLastError());
        c->built = true;
        return GLINT_OK;
    }
    for (int level = 0; level < 64; level++) {
        k_fill_u32<<<gnode, kBlock, 0, c->stream>>>(s.count, 0u, s.node_cap);
        k_reset_tick<<<1, 1, 0, c->stream>>>(s);
        k_build_count<<<gn, kBlock, 0, c->stream>>>(s);
        k_build_collect<<<gnode, kBlock, 0, c->stream>>>(s);
        glint_status st = read_counters(c);
        if (st) return st;
        if (c->h_ctr->error) return status_from_flags(c, c->h_ctr->error);
        if (c->h_ctr->n_split == 0) break;
        k_split<<<grid_for(c->h_ctr->n_split), kBlock, 0, c->stream>>>(s, 1);
    }
    HIP_TRY(c, hipGetLastError());
    c->built = true;
    return GLINT_OK;
}

// Where a tick sequence is enqueued: one context on its own stream (each
// kernel timed when profiling is on), or a fused group of contexts on the
// group's stream, one launch per step for all of them.  The members of a
// group share every setting that picks a kernel variant (join tile, table
// workgroup, key-value output, geography, schema mode), so the sequence is
// decided by any one of them.
struct FusedGroup {
    std::vector<glint_ctx *> ctx;      // at most kMaxFused
    DevState *d_states = nullptr;      // the members' states in device memory
    hipStream_t stream = nullptr;
};

struct Seq {
    glint_ctx *one = nullptr;
    const FusedGroup *group = nullptr;
    const glint_ctx &rep() const { return one ? *one : *group->ctx[0]; }
};

// One step of the sequence: blocks(c) workgroups for context c (none: c
// takes no part); single(stream, grid, c) launches the one-context kernel,
// fused(stream, grid, many) the fused one.
template <class B, class S, class F>
void step(const Seq &q, int id, B blocks, S single, F fused)
{
    if (q.one) {
        glint_ctx *c = q.one;
        const uint32_t g = blocks(*c);
        if (g) launch(c, id, [&] { single(c->stream, g, *c); });
        return;
    }
    Many m{};
    m.ss = q.group->d_states;
    m.n = (uint32_t)q.group->ctx.size();
    uint32_t total = 0;
    for (uint32_t i = 0; i < m.n; i++) {
        m.start[i] = total;
        total += blocks(*q.group->ctx[i]);
    }
    m.start[m.n] = total;
    if (total) fused(q.group->stream, total, m);
}

// A step of a kernel whose only argument is the state: k_<name> and kf_<name>.
#define GLINT_STEP(q, id, threads, blocks_expr, name)                                                     \
    step(                                                                                                \
        q, id, [&](const glint_ctx &c) -> uint32_t { return (blocks_expr); },                            \
        [&](hipStream_t st_, uint32_t g_, glint_ctx &c) { k_##name<<<g_, threads, 0, st_>>>(c.s); },     \
        [&](hipStream_t st_, uint32_t g_, const Many &m_) { kf_##name<<<g_, threads, 0, st_>>>(m_); })

// Clearing of a per-node counter array, and the three kernels of its scan.
void step_fill(const Seq &q, int id, int which)
{
    step(
        q, id,
        [&](const glint_ctx &c) -> uint32_t {
            return which == kFillCoarse ? grid_for(kCoarseDim * kCoarseDim)
                                        : grid_for((which == kFillMeta ? 4ull : 1ull) * c.s.node_cap);
        },
        [&](hipStream_t st, uint32_t g, glint_ctx &c) {
            const DevState &s = c.s;
            if (which == kFillCount) k_fill_u32<<<g, kBlock, 0, st>>>(s.count, 0u, s.node_cap);
            if (which == kFillUcount) k_fill_u32<<<g, kBlock, 0, st>>>(s.ucount, 0u, s.node_cap);
            if (which == kFillMeta)
                k_fill_u32<<<g, kBlock, 0, st>>>(reinterpret_cast<uint32_t *>(s.meta), 0u, 4ull * s.node_cap);
            if (which == kFillCoarse) k_fill_u32<<<g, kBlock, 0, st>>>(s.g10, 0u, kCoarseDim * kCoarseDim);
        },
        [&](hipStream_t st, uint32_t g, const Many &m) {
            if (which == kFillCount) kf_fill<kFillCount><<<g, kBlock, 0, st>>>(m);
            if (which == kFillUcount) kf_fill<kFillUcount><<<g, kBlock, 0, st>>>(m);
            if (which == kFillMeta) kf_fill<kFillMeta><<<g, kBlock, 0, st>>>(m);
            if (which == kFillCoarse) kf_fill<kFillCoarse><<<g, kBlock, 0, st>>>(m);
        });
}

template <int A>
const uint32_t *scan_input(const DevState &s)
{
    return A == kScanCount ? s.count : (A == kScanUcount ? s.ucount : s.bcount);
}

template <int A>
void step_scan_totals(const Seq &q)
{
    auto nb = [](const glint_ctx &c) -> uint32_t { return scan_blocks_for(scan_length(c.s, A)); };
    step(
        q, K_SCAN_REDUCE, nb,
        [&](hipStream_t st, uint32_t g, glint_ctx &c) {
            k_scan_reduce<<<g, kBlock, 0, st>>>(scan_input<A>(c.s), scan_length(c.s, A), c.s.block_sums);
        },
        [&](hipStream_t st, uint32_t g, const Many &m) { kf_scan_reduce<A><<<g, kBlock, 0, st>>>(m); });
    step(
        q, K_SCAN_TOP, [](const glint_ctx &) -> uint32_t { return 1; },
        [&](hipStream_t st, uint32_t, glint_ctx &c) { k_scan_top<<<1, kBlock, 0, st>>>(c.s.block_sums, nb(c)); },
        [&](hipStream_t st, uint32_t g, const Many &m) { kf_scan_top<A><<<g, kBlock, 0, st>>>(m); });
}

// The tree of this tick's positions (GLINT_SCHEMA_GRID_REBUILD), before the
// assignment: a fixed sequence of launches.
void enqueue_grid_rebuild(const Seq &q)
{
    step_fill(q, K_FILL_COUNTS, kFillCoarse);
    step_fill(q, K_FILL_COUNTS, kFillMeta);
    GLINT_STEP(q, K_GRID_CELLS, kBlock, grid_for(c.s.n), grid_cells);
    GLINT_STEP(q, K_GRID_PYRAMID, kBlock, 1024u, grid_pyramid);
    GLINT_STEP(q, K_GRID_TOP, 1024, 1u, grid_top);
    GLINT_STEP(q, K_GRID_FINE_ZERO, kBlock, grid_for((uint64_t)c.s.heavy_cap * kFineCells, 2048), grid_fine_zero);
    GLINT_STEP(q, K_GRID_FINE, kBlock, grid_for(c.s.n), grid_fine);
    for (uint32_t d = 0; d <= kGridBits; d++)
        step(
            q, K_GRID_LEVEL,
            [&](const glint_ctx &c) -> uint32_t {
                return grid_for(std::min<uint64_t>(1ull << (2 * d), c.s.glist_cap), 1024);
            },
            [&](hipStream_t st, uint32_t g, glint_ctx &c) { k_grid_level<<<g, kBlock, 0, st>>>(c.s, d); },
            [&](hipStream_t st, uint32_t g, const Many &m) { kf_grid_level<<<g, kBlock, 0, st>>>(m, d); });
}

void enqueue_partition(const Seq &q)
{
    step_fill(q, K_FILL_COUNTS, kFillCount);
    GLINT_STEP(q, K_RESET, 1, 1u, reset_tick);
    if (q.rep().cfg.dynamic_schema == GLINT_SCHEMA_GRID_REBUILD) enqueue_grid_rebuild(q);
    if (q.rep().s.incremental) GLINT_STEP(q, K_REGIONS, kBlock, grid_for(c.s.node_cap), regions);
    GLINT_STEP(q, K_ASSIGN, kBlock, grid_for(c.s.n), assign);
    step_scan_totals<kScanCount>(q);
    GLINT_STEP(q, K_SCAN_OBJECTS, kBlock, c.scan_blocks, scan_objects);
    if (q.rep().s.geography)
        GLINT_STEP(q, K_SCATTER, kBlock, grid_for(c.s.n), scatter<true>);
    else
        GLINT_STEP(q, K_SCATTER, kBlock, grid_for(c.s.n), scatter<false>);
}

void enqueue_lookup(const Seq &q)
{
    step_fill(q, K_FILL_UCOUNT, kFillUcount);
    if (q.rep().s.geography)
        GLINT_STEP(q, K_LOOKUP, kBlock, grid_for(c.s.n), lookup<true>);
    else
        GLINT_STEP(q, K_LOOKUP, kBlock, grid_for(c.s.n), lookup<false>);
    step_scan_totals<kScanUcount>(q);
    GLINT_STEP(q, K_SCAN_UNITS, kBlock, c.scan_blocks, scan_units);
    GLINT_STEP(q, K_UNIT_SCATTER, kBlock, grid_for(c.s.unit_cap), unit_scatter);
}

void enqueue_refine(const Seq &q)
{
    const glint_ctx &r = q.rep();
    // a context alone gets the whole device; in a fused group, a share by size
    auto jb = [&](const glint_ctx &c) -> uint32_t { return q.one ? c.join_blocks : c.fused_join_blocks; };
    if (r.s.geography)
        step(
            q, K_JOIN, jb, [&](hipStream_t st, uint32_t g, glint_ctx &c) { k_join<256, true><<<g, kBlock, 0, st>>>(c.s); },
            [&](hipStream_t st, uint32_t g, const Many &m) { kf_join<256, true><<<g, kBlock, 0, st>>>(m); });
    else if (r.cfg.join_tile == 512)
        step(
            q, K_JOIN, jb, [&](hipStream_t st, uint32_t g, glint_ctx &c) { k_join<512><<<g, kBlock, 0, st>>>(c.s); },
            [&](hipStream_t st, uint32_t g, const Many &m) { kf_join<512><<<g, kBlock, 0, st>>>(m); });
    else
        step(
            q, K_JOIN, jb, [&](hipStream_t st, uint32_t g, glint_ctx &c) { k_join<kTile><<<g, kBlock, 0, st>>>(c.s); },
            [&](hipStream_t st, uint32_t g, const Many &m) { kf_join<kTile><<<g, kBlock, 0, st>>>(m); });
    // group the tick's contacts by table region
    // (the table's shape is decided on the device from the tick's stored
    // contacts; launch sizes cover the largest shape)
    GLINT_STEP(q, K_BUCKET_COUNT, kBlock, c.s.bucket_blocks, bucket_count);
    step_scan_totals<kScanBins>(q);
    step(
        q, K_SCAN_WRITE, [](const glint_ctx &c) -> uint32_t { return scan_blocks_for(scan_length(c.s, kScanBins)); },
        [&](hipStream_t st, uint32_t g, glint_ctx &c) {
            k_scan_write<<<g, kBlock, 0, st>>>(c.s.bcount, scan_length(c.s, kScanBins), c.s.block_sums, c.s.boff);
        },
        [&](hipStream_t st, uint32_t g, const Many &m) { kf_scan_write_bins<<<g, kBlock, 0, st>>>(m); });
    GLINT_STEP(q, K_BUCKET_SCATTER, kBucketThreads, c.s.bucket_blocks, bucket_scatter);
    GLINT_STEP(q, K_BUCKET_FINE, kBucketThreads, c.fine_grid, bucket_fine);
}

// Location updates scattered into the resident positions; pids out of
// range flag bad input.
__global__ __launch_bounds__(kBlock) void k_apply_updates(DevState s, double *xy, const uint32_t *pids,
                                                          const double *upd, uint32_t k)
{
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < k; i += gridDim.x * kBlock) {
        const uint32_t p = pids[i];
        if (p >= s.n) {
            atomicOr(&s.ctr->error, kErrBadInput);
            continue;
        }
        xy[2 * (size_t)p] = upd[2 * (size_t)i];
        xy[2 * (size_t)p + 1] = upd[2 * (size_t)i + 1];
    }
}

// Float copies of this tick's positions and of the tick before's (pid
// order) for the pair boxes.
template <class P>
__device__ __forceinline__ void pos32_body(const DevState s, const P blk)
{
    const double *xy = s.tp->xy;
    const double *xp = s.tp->has_prev ? s.tp->xy_prev : xy;
    for (uint32_t i = blk.id() * kBlock + threadIdx.x; i < s.n; i += blk.count() * kBlock)
        s.pos32[i] = make_float4((float)xy[2 * (size_t)i], (float)xy[2 * (size_t)i + 1], (float)xp[2 * (size_t)i],
                                 (float)xp[2 * (size_t)i + 1]);
}

__global__ __launch_bounds__(kBlock) void k_pos32(DevState s) { pos32_body(s, this_blk()); }
__global__ __launch_bounds__(kBlock) void kf_pos32(const Many m)
{
    const Member w = member_of(m);
    pos32_body(m.ss[w.i], w.blk);
}

template <bool KV, int NT>
void step_table(const Seq &q, bool reshaped)
{
    auto g = [](const glint_ctx &c) -> uint32_t { return c.table_grid; };
    if (reshaped)
        step(
            q, K_TABLE_RESHAPED, g,
            [&](hipStream_t st, uint32_t n, glint_ctx &c) { k_table<KV, NT, true><<<n, NT, 0, st>>>(c.s); },
            [&](hipStream_t st, uint32_t n, const Many &m) { kf_table<KV, NT, true><<<n, NT, 0, st>>>(m); });
    else
        step(
            q, K_TABLE, g,
            [&](hipStream_t st, uint32_t n, glint_ctx &c) { k_table<KV, NT, false><<<n, NT, 0, st>>>(c.s); },
            [&](hipStream_t st, uint32_t n, const Many &m) { kf_table<KV, NT, false><<<n, NT, 0, st>>>(m); });
}

void enqueue_evaluate(const Seq &q)
{
    const glint_ctx &r = q.rep();
    if (r.kv) GLINT_STEP(q, K_POS32, kBlock, grid_for(c.s.n), pos32);
    // persistent over the regions
    const uint32_t nt = r.cfg.table_threads ? r.cfg.table_threads : 512;
    // both shape variants; the one that does not apply to the tick returns at once
    for (int reshaped = 0; reshaped < 2; reshaped++) {
        if (r.kv)
            nt == 1024 ? step_table<true, 1024>(q, reshaped) : nt == 512 ? step_table<true, 512>(q, reshaped)
                                                                         : step_table<true, 256>(q, reshaped);
        else
            nt == 1024 ? step_table<false, 1024>(q, reshaped) : nt == 512 ? step_table<false, 512>(q, reshaped)
                                                                          : step_table<false, 256>(q, reshaped);
    }
    GLINT_STEP(q, K_LOG_APPEND, kBlock, c.s.mlog ? 1024u : 0u, log_append);
}

void enqueue_schema(const Seq &q)
{
    if (q.rep().cfg.dynamic_schema == GLINT_SCHEMA_DYNAMIC) {
        GLINT_STEP(q, K_COLLECT, kBlock, grid_for(c.s.node_cap), collect);
        step(
            q, K_SPLIT, [](const glint_ctx &c) -> uint32_t { return grid_for(c.s.node_cap); },
            [&](hipStream_t st, uint32_t g, glint_ctx &c) { k_split<<<g, kBlock, 0, st>>>(c.s, 0); },
            [&](hipStream_t st, uint32_t g, const Many &m) { kf_split<<<g, kBlock, 0, st>>>(m, 0); });
        GLINT_STEP(q, K_MERGE, kBlock, grid_for(c.s.node_cap), merge);
    }
    GLINT_STEP(q, K_SCHEMA_DONE, 1, 1u, schema_done);
}

void enqueue_partition(glint_ctx *c) { enqueue_partition(Seq{c, nullptr}); }
void enqueue_lookup(glint_ctx *c) { enqueue_lookup(Seq{c, nullptr}); }
void enqueue_refine(glint_ctx *c) { enqueue_refine(Seq{c, nullptr}); }
void enqueue_evaluate(glint_ctx *c) { enqueue_evaluate(Seq{c, nullptr}); }
void enqueue_schema(glint_ctx *c) { enqueue_schema(Seq{c, nullptr}); }

void fill_stats(glint_ctx *c)
{
    const Counters &h = *c->h_ctr;
    glint_tick_stats &t = c->last;
    t.tick = c->tick;
    t.num_leaves = h.n_leaves;
    t.num_nodes = c->cfg.dynamic_schema == GLINT_SCHEMA_GRID_REBUILD
                      ? std::min(h.g_nodes, c->s.node_cap)
                      : (uint32_t)(c->s.node_cap - (uint32_t)std::max(h.free_top, 0));
    t.num_border = h.n_border;
    t.max_leaf_count = h.max_leaf;
    t.splits = h.last_splits;
    t.merges = h.last_merges;
    t.error_flags = h.error;
    t.num_units = h.n_units;
    t.num_contacts = h.n_contacts;
    t.num_active = h.n_contacts - h.n_dup;
    t.num_stored = h.n_stored;
    t.num_meetings = h.n_meet;
    t.num_rechecks = h.n_rechecks;
    t.table_slots = (uint64_t)kRegion << h.tab_bits;
    t.ref_drop_leaves = h.n_drop_leaves;
    t.ref_drop_objects = h.n_drop_objs;
    t.num_tests = h.n_tests;
    t.num_evaluated = h.n_evals;
}

}  // namespace

// ============================================== allocation-free tick runs
//
// glint_run_ticks_noalloc keeps its per-tick clock stamps and their
// conversion to milliseconds in this file-scope device array instead of a
// buffer the library allocates.  Every GPU holds its own copy until the
// process ends; all calls on one GPU share it, so calls that overlap on one
// GPU race (each call writes every entry it reads, nothing needs clearing).
// Host code never passes the array's name to a kernel or copy: its device
// address on each GPU is resolved once with hipGetSymbolAddress (at context
// creation, with that device current) and that pointer is used.
constexpr uint32_t kNoAllocMaxTicks = 4096;
constexpr uint32_t kNoAllocWords = 2 * kNoAllocMaxTicks + kNoAllocMaxTicks / 2;   // stamps, then floats
static __device__ __align__(256) unsigned long long scratch[kNoAllocWords];

namespace {

constexpr int kMaxDevices = 64;
unsigned long long *g_scratch[kMaxDevices] = {};   // scratch's address on each device (null: not resolved)
std::mutex g_scratch_mu;

// Resolves scratch's address on `device`, which must be current.
glint_status resolve_scratch(glint_ctx *c, int device)
{
    if (device < 0 || device >= kMaxDevices) {
        c->err = "device ordinal beyond the scratch table";
        return GLINT_ERR_INVALID_ARGUMENT;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[device]) return GLINT_OK;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return hip_fail(c, e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
    g_scratch[device] = static_cast<unsigned long long *>(ptr);
    return GLINT_OK;
}

unsigned long long *scratch_on(int device)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return device >= 0 && device < kMaxDevices ? g_scratch[device] : nullptr;
}

// Tick k's latency in milliseconds from its two stamps.
__global__ __launch_bounds__(kBlock) void k_stamps_to_ms(const unsigned long long *stamps, float *ms, uint32_t n,
                                                         double khz)
{
    for (uint32_t k = blockIdx.x * kBlock + threadIdx.x; k < n; k += gridDim.x * kBlock)
        ms[k] = (float)((double)(stamps[2 * k + 1] - stamps[2 * k]) / khz);
}

__global__ __launch_bounds__(kBlock) void k_longitude_scale(const double *lat, double *out, uint64_t n)
{
    for (uint64_t k = (uint64_t)blockIdx.x * kBlock + threadIdx.x; k < n; k += (uint64_t)gridDim.x * kBlock)
        out[k] = deg_per_meter_lon(lat[k]);
}

}  // namespace

// ====================================================================== C ABI

extern "C" {

const char *glint_status_string(glint_status s)
{
    switch (s) {
    case GLINT_OK: return "ok";
    case GLINT_ERR_INVALID_ARGUMENT: return "invalid argument";
    case GLINT_ERR_OUT_OF_MEMORY: return "out of device memory";
    case GLINT_ERR_HIP: return "HIP runtime error";
    case GLINT_ERR_BAD_INPUT: return "position is NaN or outside the root box";
    case GLINT_ERR_TABLE_FULL: return "meeting table full";
    case GLINT_ERR_NODE_CAPACITY: return "quadtree node capacity exceeded";
    case GLINT_ERR_UNIT_CAPACITY: return "check-unit capacity exceeded";
    case GLINT_ERR_OUTPUT_CAPACITY: return "output capacity exceeded";
    case GLINT_ERR_STATE: return "call out of order";
    case GLINT_ERR_INJECTED: return "injected failure";
    }
    return "unknown status";
}

const char *glint_last_error_message(const glint_ctx *ctx)
{
    return ctx ? ctx->err.c_str() : g_create_error.c_str();
}

double glint_reach_threshold(double r)
{
    return r > 0 ? sqrt_threshold(r) : 0.0;
}

glint_status glint_longitude_scale_device(const double *d_latitudes, double *d_scale, uint64_t n, uint32_t device)
{
    if (n && (!d_latitudes || !d_scale)) return GLINT_ERR_INVALID_ARGUMENT;
    if (n == 0) return GLINT_OK;
    if (hipSetDevice((int)device) != hipSuccess) return GLINT_ERR_HIP;
    k_longitude_scale<<<grid_for(n), kBlock>>>(d_latitudes, d_scale, n);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) return GLINT_ERR_HIP;
    return GLINT_OK;
}

void glint_config_default(glint_config *cfg)
{
    if (!cfg) return;
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->grid_capacity = 50;
    cfg->min_meet_time = 5;
    cfg->max_meet_time = 4096;
    cfg->schema_update_delay = 1;
    cfg->dynamic_schema = 1;
    cfg->reach_distance = 2.0;
    cfg->guard_epsilon = 1e-3;
    cfg->kv_split_num = 10;        // config.h:60
    cfg->kv_area_limit = 100.0;    // the "not cut" entry of process.cpp:408-410
}

glint_status glint_create(const glint_config *cfg, glint_ctx **out)
{
    if (!cfg || !out) return GLINT_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    g_create_error.clear();
    if (cfg->num_objects == 0 || cfg->grid_capacity == 0 || !(cfg->reach_distance > 0) ||
        !(cfg->high_x > cfg->low_x) || !(cfg->high_y > cfg->low_y) || cfg->candidate_mode > 1 ||
        cfg->precision > 1 || cfg->num_objects >= 0x40000000u || cfg->dynamic_schema > GLINT_SCHEMA_GRID_REBUILD ||
        (cfg->join_tile != 0 && cfg->join_tile != 256 && cfg->join_tile != 512) ||
        (cfg->table_threads != 0 && cfg->table_threads != 256 && cfg->table_threads != 512 &&
         cfg->table_threads != 1024)) {
        g_create_error = "invalid configuration";
        return GLINT_ERR_INVALID_ARGUMENT;
    }
    if (cfg->precision == GLINT_PRECISION_FP32_GUARD && !(cfg->guard_epsilon > 0)) {
        g_create_error = "guard_epsilon must be positive";
        return GLINT_ERR_INVALID_ARGUMENT;
    }
    if (cfg->geography > 1 ||
        (cfg->geography && (cfg->precision != GLINT_PRECISION_FP64 ||
                            cfg->dynamic_schema == GLINT_SCHEMA_GRID_REBUILD || cfg->low_y <= -90.0 ||
                            cfg->high_y >= 90.0))) {
        g_create_error = "geography needs FP64 precision, a schema other than the grid rebuild and latitudes "
                         "inside (-90, 90)";
        return GLINT_ERR_INVALID_ARGUMENT;
    }
    glint_ctx *c = new (std::nothrow) glint_ctx();
    if (!c) return GLINT_ERR_OUT_OF_MEMORY;
    c->cfg = *cfg;
    if (c->cfg.schema_update_delay == 0) c->cfg.schema_update_delay = 1;
    c->device = (int)cfg->device;
    auto fail = [&](glint_status st) {
        g_create_error = c->err;
        glint_destroy(c);
        return st;
    };
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) {
        c->err = std::string("hipSetDevice: ") + hipGetErrorString(e);
        return fail(GLINT_ERR_HIP);
    }
    {
        int khz = 0;
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device) == hipSuccess && khz > 0)
            c->wall_khz = khz;
    }
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        c->err = std::string("hipStreamCreate: ") + hipGetErrorString(e);
        return fail(GLINT_ERR_HIP);
    }
    {
        const glint_status sst = resolve_scratch(c, c->device);
        if (sst) return fail(sst);
    }
    for (auto &ev : c->ev) {
        if (hipEventCreate(&ev) != hipSuccess) {
            c->err = "hipEventCreate failed";
            return fail(GLINT_ERR_HIP);
        }
    }
    if (hipHostMalloc((void **)&c->h_ctr, sizeof(Counters), hipHostMallocDefault) != hipSuccess) {
        c->err = "hipHostMalloc failed";
        return fail(GLINT_ERR_OUT_OF_MEMORY);
    }
    std::memset(c->h_ctr, 0, sizeof(Counters));

    const uint64_t N = cfg->num_objects;
    DevState &s = c->s;
    s.n = (uint32_t)N;
    s.cap = cfg->grid_capacity;
    s.min_meet = cfg->min_meet_time;
    s.complete = cfg->candidate_mode;
    s.precision = cfg->precision;
    // a contact first seen by the inside join of its leaf and not within reach
    // at the tick before needs no entry: it can only continue a run (found
    // again from the leaves of the tick before) or end unemitted, unless
    // min_meet_time is 0 or every contact is kept for glint_contacts
    s.store_all = (cfg->record_contacts || cfg->min_meet_time == 0 || cfg->geography) ? 1u : 0u;
    // with geography the scaled pair distance depends on which object is p1,
    // so a pair's state at the tick before cannot be told from positions
    // alone: every contact is stored and looked up
    s.geography = cfg->geography;
    s.incremental = cfg->dynamic_schema != GLINT_SCHEMA_GRID_REBUILD;
    s.xbuf = cfg->geography ? cfg->reach_distance * deg_per_meter_lon_host(cfg->low_y) : cfg->reach_distance;
    s.ybuf = cfg->geography ? cfg->reach_distance * kDegPerMeterLat : cfg->reach_distance;
    s.delay = c->cfg.schema_update_delay;
    s.lx = cfg->low_x;
    s.ly = cfg->low_y;
    s.hx = cfg->high_x;
    s.hy = cfg->high_y;
    s.r = cfg->reach_distance;
    s.t64 = sqrt_threshold(cfg->reach_distance);
    {
        double r = cfg->reach_distance, eps = cfg->guard_epsilon;
        double lo = r > eps ? (r - eps) * (r - eps) : 0.0, hi = (r + eps) * (r + eps);
        float fl = (float)lo, fh = (float)hi;
        if ((double)fl > lo) fl = std::nextafter(fl, 0.0f);
        if ((double)fh < hi) fh = std::nextafter(fh, INFINITY);
        s.t32_in = fl;
        s.t32_out = fh;
        // Largest leaf side whose FP32 offsets keep |d32 - d64| well inside
        // the guard band: offsets carry <= u*(W + r) error each (u = 2^-24),
        // the FP32 difference, squares and sum (the sum fused with one square)
        // add <= (2+sqrt2)*u*(r+eps);
        // the bound is 2*sqrt2*u*(W+r) + (2+sqrt2)*u*(r+eps) < eps/2.
        const double u = std::ldexp(1.0, -24);
        double w = (eps / 2 - (2 + std::sqrt(2.0)) * u * (r + eps)) / (2 * std::sqrt(2.0) * u * (1 + u)) - r;
        s.w32_max = w > 0 ? w : 0.0;
        s.pf_lim = s.w32_max + r;
    }

    uint64_t node_cap = cfg->node_capacity ? cfg->node_capacity
                                           : std::max<uint64_t>(65536, 16 * (N / cfg->grid_capacity + 1) + 1024);
    if (node_cap > 0x7FFFFFF0ull) node_cap = 0x7FFFFFF0ull;
    s.node_cap = (uint32_t)node_cap;
    uint64_t slots = cfg->num_meeting_buckets ? next_pow2(cfg->num_meeting_buckets)
                                              : next_pow2(std::max<uint64_t>(4 * N, 1024));
    if (slots < (uint64_t)kRegion) slots = kRegion;
    if (slots > (1ull << 32)) {
        c->err = "num_meeting_buckets above 2^32";
        return fail(GLINT_ERR_INVALID_ARGUMENT);
    }
    s.unit_cap = cfg->unit_capacity ? cfg->unit_capacity : std::max<uint64_t>(4 * N, 1024);
    s.meet_cap = cfg->meeting_capacity ? cfg->meeting_capacity : std::max<uint64_t>(2 * N, 1u << 20);
    c->fused_join_blocks =
        (uint32_t)std::min<uint64_t>(1024, std::max<uint64_t>(64, ((N + 1023) / 1024 + 63) / 64 * 64));
    // contacts per tick plus the unused tails of the join's per-wave chunks;
    // the chunk is about a quarter of the contacts a wave can expect, so that
    // sparse scenes do not fill the list (and every pass over it) with tails
    const uint64_t contacts = cfg->contact_capacity ? cfg->contact_capacity : 16 * N + (1u << 20);
    const uint64_t waves = (uint64_t)c->join_blocks * (kBlock / 64);
    s.chunk = 64;
    while (s.chunk < kChunk && (uint64_t)s.chunk * 8 <= contacts / waves) s.chunk *= 2;
    s.clist_cap = contacts + waves * kChunk;
    {
        // the configured slots fix the fewest regions a tick uses; a tick
        // that stores more contacts than they hold at kRegionFill per region
        // uses more, up to the most its contact list can fill
        uint32_t bits = 0;
        while ((1ull << bits) < slots / kRegion) bits++;
        uint32_t most = 0;
        while (most < kMaxRegionBits && ((uint64_t)kRegionFill << most) < s.clist_cap) most++;
        s.rbits_min = std::min<uint32_t>(bits, kMaxRegionBits);
        s.rbits_max = std::max(s.rbits_min, most);
        const uint32_t lo = s.rbits_max > 10 ? s.rbits_max - 10 : 0;
        s.coarse_max = 1u << (s.rbits_max - lo);
        s.rs_total = 1u << s.rbits_max;
        // coarse-pass workgroups: one per 4 tiles of the configured
        // contacts, a power of two from 16 to 1024
        uint32_t bb = 16;
        while (bb < 1024 && (uint64_t)bb * 4 * kBucketTile < contacts) bb *= 2;
        s.bucket_blocks = bb;
        // the persistent passes get the workgroups of the configured shape
        // (at least 256 table regions and 64 bins when the shape can grow
        // that far); a tick whose table grows past it loops over the rest
        c->table_grid = std::min<uint32_t>(
            std::max<uint32_t>(1u << s.rbits_min, std::min<uint32_t>(1u << s.rbits_max, 256)), 4096);
        // (a shape of b bits has min(2^b, 1024) coarse bins)
        const uint32_t fine_min = 1u << std::min<uint32_t>(s.rbits_min, 10);
        c->fine_grid = std::min<uint32_t>(s.coarse_max, std::max<uint32_t>(fine_min, 64));
    }
    c->scan_blocks = (uint32_t)((node_cap + kScanTile - 1) / kScanTile);
    const uint64_t nbc = (uint64_t)(s.coarse_max + 1) * s.bucket_blocks;
    const uint32_t sum_blocks = std::max<uint32_t>(c->scan_blocks, (uint32_t)((nbc + kScanTile - 1) / kScanTile));

    glint_status st = GLINT_OK;
#define ALLOC(p, n)                          \
    do {                                     \
        if ((st = dalloc(c, &(p), (n)))) return fail(st); \
    } while (0)
    ALLOC(s.tp, 1);
    ALLOC(s.nodes, node_cap);
    ALLOC(s.meta, node_cap);
    ALLOC(s.free_stack, node_cap);
    ALLOC(s.count, node_cap);
    ALLOC(s.offset, node_cap);
    ALLOC(s.block_sums, sum_blocks);
    ALLOC(s.bcount, nbc);
    ALLOC(s.boff, nbc);
    ALLOC(s.clist2, s.clist_cap);
    ALLOC(s.split_list, node_cap);
    ALLOC(s.merge_list, node_cap);
    ALLOC(s.leaf_list, node_cap);
    ALLOC(s.sleafb[0], N);
    ALLOC(s.sleafb[1], N);
    ALLOC(s.parent, node_cap);
    ALLOC(s.region, node_cap);
    if (!s.store_all) {
        ALLOC(s.cand, N);
        ALLOC(s.prev_cand, N);
        c->cand[0] = s.cand;
        c->cand[1] = s.prev_cand;
        // empty lists stamped with no tick
        k_fill_u32<<<grid_for(4 * N), kBlock, 0, c->stream>>>(reinterpret_cast<uint32_t *>(s.cand), kNone, 4 * N);
        k_fill_u32<<<grid_for(4 * N), kBlock, 0, c->stream>>>(reinterpret_cast<uint32_t *>(s.prev_cand), kNone,
                                                             4 * N);
    }
    ALLOC(s.border, N);
    ALLOC(s.units, s.unit_cap);
    ALLOC(s.urank, s.unit_cap);
    ALLOC(s.unit_pid, s.unit_cap);
    ALLOC(s.uaux, s.unit_cap);
    ALLOC(s.unit_lf, s.unit_cap);
    ALLOC(s.ucount, node_cap);
    ALLOC(s.ustart, node_cap);
    if (cfg->dynamic_schema == GLINT_SCHEMA_GRID_REBUILD) {
        // blocks over capacity at one depth are disjoint: at most N/(cap+1)
        s.heavy_cap = (uint32_t)(N / ((uint64_t)cfg->grid_capacity + 1) + 1);
        s.glist_cap = s.heavy_cap;
        ALLOC(s.gcell, N);
        ALLOC(s.g10, (uint64_t)kCoarseDim * kCoarseDim);
        ALLOC(s.gpyr, (uint64_t)pyr_off(kCoarseDepth));
        ALLOC(s.gheavy, (uint64_t)kCoarseDim * kCoarseDim);
        ALLOC(s.gfine, (uint64_t)s.heavy_cap * kFineCells);
        ALLOC(s.glist[0], s.glist_cap);
        ALLOC(s.glist[1], s.glist_cap);
    }
    ALLOC(s.sxb[0], N);
    ALLOC(s.syb[0], N);
    ALLOC(s.sxb[1], N);
    ALLOC(s.syb[1], N);
    ALLOC(s.ord_xy, N);
    ALLOC(c->hist[0], 2 * N);
    ALLOC(c->hist[1], 2 * N);
    ALLOC(s.spid, N);
    ALLOC(s.sprev, N);
    if (cfg->geography) ALLOC(s.sgeo, N);
    ALLOC(s.order, N);
    ALLOC(s.ord_leaf, N);
    ALLOC(s.ord_rank, N);
    ALLOC(s.ord_pleaf, N);
    ALLOC(s.splf, N);
    if (cfg->precision == GLINT_PRECISION_FP32_GUARD) {
        ALLOC(s.fx, N);
        ALLOC(s.fy, N);
        ALLOC(s.pfx, N);
        ALLOC(s.pfy, N);
    }
    ALLOC(s.meet, s.meet_cap);
    ALLOC(s.ctr, 1);
    for (int g = 0; g < 2; g++) {
        ALLOC(c->tab[g].keys, s.clist_cap);
        ALLOC(c->tab[g].starts, s.clist_cap);
        ALLOC(c->tab[g].rstart, (uint64_t)s.rs_total + 2);
    }
    ALLOC(c->own_xy, 2 * N);
    if (cfg->meeting_log_capacity) {
        ALLOC(s.mlog, cfg->meeting_log_capacity);
        s.mlog_cap = cfg->meeting_log_capacity;
    }
    if (cfg->kv_capacity) {
        ALLOC(c->mbr[0], s.clist_cap);
        ALLOC(c->mbr[1], s.clist_cap);
        ALLOC(s.pos32, N);
        KvParams kp{};
        kp.num_objects = (uint32_t)N;
        kp.split_num = cfg->kv_split_num;
        kp.max_meet_time = cfg->max_meet_time;
        kp.first_end_time_min = cfg->start_time + cfg->min_meet_time;   // process.cpp:404
        kp.area_limit = cfg->kv_area_limit;
        kp.map[0] = cfg->low_x;
        kp.map[1] = cfg->low_y;
        kp.map[2] = cfg->high_x;
        kp.map[3] = cfg->high_y;
        kp.key_capacity = cfg->kv_capacity;
        if ((st = kv_create(kp, c->stream, &c->kv, c->err))) return fail(st);
        s.kvlog = kv_log(c->kv);
        s.kvlog_cap = kv_log_capacity(c->kv);
    }
#undef ALLOC
    e = hipMemsetAsync(s.ctr, 0, sizeof(Counters), c->stream);
    if (e != hipSuccess) {
        c->err = std::string("hipMemsetAsync: ") + hipGetErrorString(e);
        return fail(GLINT_ERR_HIP);
    }
    // both generations start empty
    for (int g = 0; g < 2; g++)
        k_fill_u32<<<grid_for((uint64_t)s.rs_total + 2), kBlock, 0, c->stream>>>(c->tab[g].rstart, 0u,
                                                                               (uint64_t)s.rs_total + 2);
    e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        c->err = std::string("initialisation: ") + hipGetErrorString(e);
        return fail(GLINT_ERR_HIP);
    }
    c->tick = cfg->start_time;
    *out = c;
    return GLINT_OK;
}

glint_status glint_destroy(glint_ctx *c)
{
    if (!c) return GLINT_ERR_INVALID_ARGUMENT;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (void *p : c->allocs) (void)hipFree(p);
    for (auto &ev : c->ev)
        if (ev) (void)hipEventDestroy(ev);
    c->prof.clear();
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    if (c->stamps) (void)hipFree(c->stamps);
    kv_destroy(c->kv);
    if (c->h_ctr) (void)hipHostFree(c->h_ctr);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return GLINT_OK;
}

static glint_status begin_tick(glint_ctx *c, const double *d_xy, uint32_t tick)
{
    if (c->has_positions && !c->evaluated) {
        c->err = "previous tick was not evaluated";
        return GLINT_ERR_STATE;
    }
    if (c->has_prev && tick <= c->prev_tick) {
        c->err = "ticks must increase";
        return GLINT_ERR_STATE;
    }
    c->tick = tick;
    set_tick_params(c, d_xy);
    c->last_xy = d_xy;
    c->resident = d_xy == c->own_xy && c->resident;
    c->has_positions = true;
    c->partitioned = c->refined = c->evaluated = false;
    return GLINT_OK;
}

glint_status glint_set_positions(glint_ctx *c, const double *xy, uint32_t n, uint32_t tick)
{
    if (!c || !xy || n != c->s.n) return GLINT_ERR_INVALID_ARGUMENT;
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipMemcpyAsync(c->own_xy, xy, (size_t)n * 16, hipMemcpyHostToDevice, c->stream));
    return begin_tick(c, c->own_xy, tick);
}

glint_status glint_set_positions_device(glint_ctx *c, const double *d_xy, uint32_t n, uint32_t tick)
{
    if (!c || !d_xy || n != c->s.n) return GLINT_ERR_INVALID_ARGUMENT;
    HIP_TRY(c, hipSetDevice(c->device));
    return begin_tick(c, d_xy, tick);
}

glint_status glint_ingest_updates(glint_ctx *c, const uint32_t *d_pids, const double *d_xy, uint32_t k,
                                  uint32_t tick)
{
    if (!c || (k && (!d_pids || !d_xy))) return GLINT_ERR_INVALID_ARGUMENT;
    if (!c->last_xy) {
        c->err = "ingestion needs a first tick with every position";
        return GLINT_ERR_STATE;
    }
    if (c->has_positions && !c->evaluated) {
        c->err = "previous tick was not evaluated";
        return GLINT_ERR_STATE;
    }
    HIP_TRY(c, hipSetDevice(c->device));
    if (!c->resident) {
        // the last tick's positions become the resident copy
        if (c->last_xy != c->own_xy)
            HIP_TRY(c, hipMemcpyAsync(c->own_xy, c->last_xy, (size_t)c->s.n * 16, hipMemcpyDeviceToDevice,
                                      c->stream));
    }
    if (k) k_apply_updates<<<grid_for(k), kBlock, 0, c->stream>>>(c->s, c->own_xy, d_pids, d_xy, k);
    HIP_TRY(c, hipGetLastError());
    glint_status st = begin_tick(c, c->own_xy, tick);
    c->resident = true;
    return st;
}

glint_status glint_ctx_view(glint_ctx *c, CtxView *v)
{
    if (!c || !v) return GLINT_ERR_INVALID_ARGUMENT;
    v->stream = c->stream;
    v->device = c->device;
    v->num_objects = c->s.n;
    v->tick = c->prev_tick;
    v->min_meet = c->s.min_meet;
    v->evaluated = c->evaluated && c->has_prev;
    // after evaluate the current tick's pairs are the previous generation
    const TableGen &tab = c->tab[c->evaluated ? 1 - c->cur_gen : c->cur_gen];
    v->keys = tab.keys;
    v->starts = tab.starts;
    v->count = tab.rstart + c->s.rs_total;
    v->cap = c->s.clist_cap;
    v->mlog = c->s.mlog;
    v->mlog_cap = c->s.mlog_cap;
    return GLINT_OK;
}

glint_status glint_partition(glint_ctx *c)
{
    if (!c) return GLINT_ERR_INVALID_ARGUMENT;
    if (!c->has_positions || c->partitioned) {
        c->err = "partition needs new positions";
        return GLINT_ERR_STATE;
    }
    if (c->inject & kInjPartition) {
        c->inject &= ~kInjPartition;
        c->err = "injected partition failure";
        return GLINT_ERR_INJECTED;
    }
    HIP_TRY(c, hipSetDevice(c->device));
    if (!c->built) {
        glint_status st = build_tree(c);
        if (st) return st;
    }
    HIP_TRY(c, hipEventRecord(c->ev[0], c->stream));
    enqueue_partition(c);
    HIP_TRY(c, hipEventRecord(c->ev[1], c->stream));
    enqueue_lookup(c);
    HIP_TRY(c, hipEventRecord(c->ev[2], c->stream));
    HIP_TRY(c, hipGetLastError());
    c->partitioned = true;
    return GLINT_OK;
}

glint_status glint_refine(glint_ctx *c)
{
    if (!c) return GLINT_ERR_INVALID_ARGUMENT;
    if (!c->partitioned || c->refined) {
        c->err = "refine needs a partitioned tick";
        return GLINT_ERR_STATE;
    }
    if (c->inject & kInjRefine) {
        c->inject &= ~kInjRefine;
        c->err = "injected refine failure";
        return GLINT_ERR_INJECTED;
    }
    HIP_TRY(c, hipSetDevice(c->device));
    enqueue_refine(c);
    HIP_TRY(c, hipEventRecord(c->ev[3], c->stream));
    HIP_TRY(c, hipGetLastError());
    c->refined = true;
    return GLINT_OK;
}

glint_status glint_evaluate(glint_ctx *c, glint_meeting *out, uint64_t cap, uint64_t *n_out)
{
    if (!c) return GLINT_ERR_INVALID_ARGUMENT;
    if (!c->refined || c->evaluated) {
        c->err = "evaluate needs a refined tick";
        return GLINT_ERR_STATE;
    }
    if (c->inject & kInjEvaluate) {
        c->inject &= ~kInjEvaluate;
        c->err = "injected evaluate failure";
        return GLINT_ERR_INJECTED;
    }
    HIP_TRY(c, hipSetDevice(c->device));
    enqueue_evaluate(c);
    HIP_TRY(c, hipEventRecord(c->ev[4], c->stream));
    HIP_TRY(c, hipGetLastError());
    c->evaluated = true;
    c->cur_gen = 1 - c->cur_gen;   // this tick's pairs become the previous generation
    c->prev_tick = c->tick;
    c->has_prev = true;
    if (!out && !n_out) return GLINT_OK;
    glint_status st = read_counters(c);
    if (st) return st;
    uint64_t n = c->h_ctr->n_meet;
    if (n_out) *n_out = n;
    if (c->h_ctr->error) return status_from_flags(c, c->h_ctr->error);
    if (!out) return GLINT_OK;
    if (n > cap) {
        c->err = "meeting output buffer too small";
        return GLINT_ERR_OUTPUT_CAPACITY;
    }
    if (n) {
        HIP_TRY(c, hipMemcpy(out, c->s.meet, n * sizeof(glint_meeting), hipMemcpyDeviceToHost));
        std::sort(out, out + n, [](const glint_meeting &a, const glint_meeting &b) {
            if (a.pid_a != b.pid_a) return a.pid_a < b.pid_a;
            if (a.pid_b != b.pid_b) return a.pid_b < b.pid_b;
            return a.start < b.start;
        });
    }
    return GLINT_OK;
}

glint_status glint_update_schema(glint_ctx *c)
{
    if (!c) return GLINT_ERR_INVALID_ARGUMENT;
    if (!c->evaluated) {
        c->err = "update_schema needs an evaluated tick";
        return GLINT_ERR_STATE;
    }
    if (c->inject & kInjSchema) {
        c->inject &= ~kInjSchema;
        c->err = "injected schema failure";
        return GLINT_ERR_INJECTED;
    }
    HIP_TRY(c, hipSetDevice(c->device));
    enqueue_schema(c);
    HIP_TRY(c, hipEventRecord(c->ev[5], c->stream));
    HIP_TRY(c, hipGetLastError());
    return GLINT_OK;
}

glint_status glint_process_tick(glint_ctx *c, const double *xy, uint32_t tick)
{
    glint_status st;
    if ((st = glint_set_positions(c, xy, c ? c->s.n : 0, tick))) return st;
    if ((st = glint_partition(c))) return st;
    if ((st = glint_refine(c))) return st;
    if ((st = glint_evaluate(c, nullptr, 0, nullptr))) return st;
    return glint_update_schema(c);
}

static void enqueue_tick_body(glint_ctx *c, uint64_t stride)
{
    launch(c, K_TICK_PARAMS, [&] { k_advance_tick<<<1, 1, 0, c->stream>>>(c->s.tp, stride); });
    enqueue_partition(c);
    enqueue_lookup(c);
    enqueue_refine(c);
    enqueue_evaluate(c);
    enqueue_schema(c);
    k_stamp<<<1, 1, 0, c->stream>>>(c->s.tp);
}

// One tick of every member of a fused group: one launch per step.
static void enqueue_tick_body(const FusedGroup &g)
{
    const Seq q{nullptr, &g};
    const uint32_t n = (uint32_t)g.ctx.size();
    Many m{};
    m.ss = g.d_states;
    m.n = n;
    for (uint32_t i = 0; i <= n; i++) m.start[i] = i;   // one workgroup each
    kf_advance_tick<<<n, 1, 0, g.stream>>>(m);
    enqueue_partition(q);
    enqueue_lookup(q);
    enqueue_refine(q);
    enqueue_evaluate(q);
    enqueue_schema(q);
    kf_stamp<<<n, 1, 0, g.stream>>>(m);
}

// The settings that pick kernel variants; contexts alike in all of them can
// share fused launches.
static std::vector<uint64_t> variant_key(const glint_ctx *c)
{
    const uint64_t tile = c->s.geography ? 256 : (c->cfg.join_tile == 512 ? 512 : kTile);
    const uint64_t nt = c->cfg.table_threads ? c->cfg.table_threads : 512;
    return {tile, nt, c->kv ? 1ull : 0ull, c->s.geography, c->cfg.dynamic_schema};
}

// Checks, stamp buffer, the pre-run state (k_begin_run) and, on the first
// run, the tree build from the first snapshot.
// stamps: device buffer of 2*nticks stamps, or null for the context's own
// (allocated, or grown, here).
static glint_status run_begin(glint_ctx *c, const double *d_trace, uint32_t first_tick, uint32_t nticks,
                              unsigned long long *stamps = nullptr)
{
    if (c->has_positions && !c->evaluated) {
        c->err = "previous tick was not evaluated";
        return GLINT_ERR_STATE;
    }
    if (c->has_prev && first_tick <= c->prev_tick) {
        c->err = "ticks must increase";
        return GLINT_ERR_STATE;
    }
    HIP_TRY(c, hipSetDevice(c->device));
    if (!stamps && c->stamp_cap < 2 * nticks) {
        if (c->stamps) HIP_TRY(c, hipFree(c->stamps));
        c->stamps = nullptr;
        c->stamp_cap = 0;
        HIP_TRY(c, hipMalloc((void **)&c->stamps, (size_t)2 * nticks * sizeof(unsigned long long)));
        c->stamp_cap = 2 * nticks;
    }
    if (!stamps) stamps = c->stamps;
    c->tick = first_tick;
    const int g = c->cur_gen;
    // one advance short of the first tick: tables and histories pre-swapped
    k_begin_run<<<1, 1, 0, c->stream>>>(c->s.tp, d_trace, c->tab[g], c->tab[1 - g], c->mbr[g], c->mbr[1 - g],
                                        c->hist[g], c->hist[1 - g], c->cand[g], c->cand[1 - g], c->s.sxb[g],
                                        c->s.syb[g], c->s.sxb[1 - g], c->s.syb[1 - g], c->s.sleafb[g],
                                        c->s.sleafb[1 - g], c->prev_tick, first_tick, c->has_prev ? 1u : 0u, stamps, 2 * nticks,
                                        nticks);
    HIP_TRY(c, hipGetLastError());
    if (!c->built) {
        glint_status st = build_tree(c);
        if (st) return st;
    }
    return GLINT_OK;
}

// Host view of the state after nticks enqueued ticks.
static void run_end_state(glint_ctx *c, uint32_t first_tick, uint32_t nticks)
{
    for (uint32_t k = 0; k < nticks; k++) c->cur_gen = 1 - c->cur_gen;
    c->prev_tick = first_tick + nticks - 1;
    c->has_prev = true;
    c->tick = first_tick + nticks - 1;
    c->has_positions = true;
    c->partitioned = c->refined = c->evaluated = true;
}

// Counters, per-tick latencies from the device stamps and the status.
static glint_status run_finish(glint_ctx *c, uint32_t nticks, float *tick_ms)
{
    glint_status st = read_counters(c);
    if (st) return st;
    if (tick_ms) {
        std::vector<unsigned long long> h(2 * (size_t)nticks);
        HIP_TRY(c, hipMemcpy(h.data(), c->stamps, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
        const double khz = c->wall_khz > 0 ? c->wall_khz : 100000.0;
        for (uint32_t k = 0; k < nticks; k++) tick_ms[k] = (float)((double)(h[2 * k + 1] - h[2 * k]) / khz);
    }
    fill_stats(c);
    return status_from_flags(c, c->h_ctr->error);
}

glint_status glint_run_ticks(glint_ctx *c, const double *d_trace, uint32_t first_tick, uint32_t nticks,
                             float *tick_ms)
{
    if (!c || !d_trace || nticks == 0) return GLINT_ERR_INVALID_ARGUMENT;
    glint_status st = run_begin(c, d_trace, first_tick, nticks);
    if (st) return st;
    const uint64_t stride = 2ull * c->s.n;
    const uint32_t G = c->cfg.graph_ticks;
    uint32_t done = 0;
    if (G && !c->prof.on && nticks >= G) {
        if (!c->graph_exec) {
            // G consecutive ticks captured once; every tick reads its parameters
            // from device memory, so the graph is replayed unchanged.  Every
            // exit after the capture began ends it, so the stream stays usable.
            hipGraph_t graph = nullptr;
            HIP_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            for (uint32_t k = 0; k < G; k++) enqueue_tick_body(c, stride);
            hipError_t e = hipGetLastError();
            const hipError_t ee = hipStreamEndCapture(c->stream, &graph);
            if (e == hipSuccess) e = ee;
            if (e == hipSuccess) e = hipGraphInstantiate(&c->graph_exec, graph, nullptr, nullptr, 0);
            if (graph) (void)hipGraphDestroy(graph);
            if (e != hipSuccess) {
                c->graph_exec = nullptr;
                return hip_fail(c, e, "graph capture of the tick sequence");
            }
        }
        for (; done + G <= nticks; done += G) HIP_TRY(c, hipGraphLaunch(c->graph_exec, c->stream));
    }
    for (; done < nticks; done++) enqueue_tick_body(c, stride);
    HIP_TRY(c, hipGetLastError());
    run_end_state(c, first_tick, nticks);
    return run_finish(c, nticks, tick_ms);
}

glint_status glint_run_ticks_noalloc(glint_ctx *c, const double *d_trace, uint32_t first_tick, uint32_t nticks,
                                     float *tick_ms)
{
    if (!c || !d_trace || nticks == 0) return GLINT_ERR_INVALID_ARGUMENT;
    if (nticks > kNoAllocMaxTicks) {
        c->err = "glint_run_ticks_noalloc runs at most " + std::to_string(kNoAllocMaxTicks) +
                 " ticks per call (its stamps live in a fixed device array)";
        return GLINT_ERR_INVALID_ARGUMENT;
    }
    if (c->prof.on) {
        c->err = "glint_run_ticks_noalloc: kernel profiling creates events; turn it off first";
        return GLINT_ERR_STATE;
    }
    unsigned long long *stamps = scratch_on(c->device);   // resolved when the context was created
    if (!stamps) {
        c->err = "scratch array not resolved on this device";
        return GLINT_ERR_STATE;
    }
    float *ms = reinterpret_cast<float *>(stamps + 2 * kNoAllocMaxTicks);
    glint_status st = run_begin(c, d_trace, first_tick, nticks, stamps);
    if (st) return st;
    const uint64_t stride = 2ull * c->s.n;
    for (uint32_t k = 0; k < nticks; k++) enqueue_tick_body(c, stride);   // plain launches, no graph
    HIP_TRY(c, hipGetLastError());
    run_end_state(c, first_tick, nticks);
    if (tick_ms) {
        const double khz = c->wall_khz > 0 ? c->wall_khz : 100000.0;
        k_stamps_to_ms<<<grid_for(nticks), kBlock, 0, c->stream>>>(stamps, ms, nticks, khz);
        HIP_TRY(c, hipGetLastError());
        HIP_TRY(c, hipMemcpyAsync(tick_ms, ms, (size_t)nticks * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    }
    st = read_counters(c);   // synchronises the stream
    if (st) return st;
    fill_stats(c);
    return status_from_flags(c, c->h_ctr->error);
}

glint_status glint_synchronize(glint_ctx *c)
{
    if (!c) return GLINT_ERR_INVALID_ARGUMENT;
    glint_status st = read_counters(c);
    if (st) return st;
    return status_from_flags(c, c->h_ctr->error);
}

glint_status glint_stats(glint_ctx *c, glint_tick_stats *out)
{
    if (!c || !out) return GLINT_ERR_INVALID_ARGUMENT;
    glint_status st = read_counters(c);
    if (st) return st;
    fill_stats(c);
    float ms;
    glint_tick_stats &t = c->last;
    t.ms_partition = t.ms_lookup = t.ms_refine = t.ms_evaluate = t.ms_schema = 0;
    if (c->evaluated && c->partitioned) {
        if (hipEventElapsedTime(&ms, c->ev[0], c->ev[1]) == hipSuccess) t.ms_partition = ms;
        if (hipEventElapsedTime(&ms, c->ev[1], c->ev[2]) == hipSuccess) t.ms_lookup = ms;
        if (hipEventElapsedTime(&ms, c->ev[2], c->ev[3]) == hipSuccess) t.ms_refine = ms;
        if (hipEventElapsedTime(&ms, c->ev[3], c->ev[4]) == hipSuccess) t.ms_evaluate = ms;
        if (hipEventElapsedTime(&ms, c->ev[4], c->ev[5]) == hipSuccess) t.ms_schema = ms;
    }
    (void)hipGetLastError();
    *out = t;
    return status_from_flags(c, c->h_ctr->error);
}

glint_status glint_contacts(glint_ctx *c, uint64_t *out, uint64_t cap, uint64_t *n_out)
{
    if (!c || !n_out) return GLINT_ERR_INVALID_ARGUMENT;
    if (!c->cfg.record_contacts) {
        c->err = "contacts are kept only with record_contacts";
        return GLINT_ERR_STATE;
    }
    glint_status st = read_counters(c);
    if (st) return st;
    if (c->h_ctr->error) return status_from_flags(c, c->h_ctr->error);
    // the tick's bucketed contact list holds exactly n_contacts entries
    const uint64_t n = std::min<uint64_t>(c->h_ctr->n_contacts, c->s.clist_cap);
    // after evaluate the tick's list is the previous generation
    const unsigned long long *list = c->tab[c->evaluated ? 1 - c->cur_gen : c->cur_gen].keys;
    std::vector<uint64_t> v(n);
    if (n) HIP_TRY(c, hipMemcpy(v.data(), list, n * 8, hipMemcpyDeviceToHost));
    for (auto &k : v) k &= kKeyBits;
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    *n_out = v.size();
    if (!out) return GLINT_OK;
    if (v.size() > cap) return GLINT_ERR_OUTPUT_CAPACITY;
    std::copy(v.begin(), v.end(), out);
    return GLINT_OK;
}

glint_status glint_leaves(glint_ctx *c, double *boxes, uint32_t *levels, uint32_t *counts, uint64_t cap,
                          uint64_t *n_out)
{
    if (!c || !n_out) return GLINT_ERR_INVALID_ARGUMENT;
    if (!c->built) {
        *n_out = 0;
        return GLINT_OK;
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    const uint32_t nc = c->s.node_cap;
    std::vector<Node> nodes(nc);
    std::vector<NodeMeta> meta(nc);
    std::vector<uint32_t> cnt(nc);
    HIP_TRY(c, hipMemcpy(nodes.data(), c->s.nodes, nc * sizeof(Node), hipMemcpyDeviceToHost));
    HIP_TRY(c, hipMemcpy(meta.data(), c->s.meta, nc * sizeof(NodeMeta), hipMemcpyDeviceToHost));
    HIP_TRY(c, hipMemcpy(cnt.data(), c->s.count, nc * sizeof(uint32_t), hipMemcpyDeviceToHost));
    std::vector<uint32_t> stack{0}, order;
    while (!stack.empty()) {
        uint32_t n = stack.back();
        stack.pop_back();
        if (nodes[n].child[0] == kLeafTag) {
            order.push_back(n);
        } else {
            for (int i = 3; i >= 0; i--) stack.push_back(nodes[n].child[i]);
        }
    }
    *n_out = order.size();
    for (size_t k = 0; k < order.size() && k < cap; k++) {
        const Node &nd = nodes[order[k]];
        if (boxes) {
            boxes[4 * k] = nd.lx;
            boxes[4 * k + 1] = nd.ly;
            boxes[4 * k + 2] = nd.hx;
            boxes[4 * k + 3] = nd.hy;
        }
        if (levels) levels[k] = (uint32_t)meta[order[k]].level;
        if (counts) counts[k] = cnt[order[k]];
    }
    return order.size() > cap && (boxes || levels || counts) ? GLINT_ERR_OUTPUT_CAPACITY : GLINT_OK;
}

glint_status glint_active_pairs(glint_ctx *c, uint64_t *keys, uint32_t *starts, uint64_t cap, uint64_t *n_out)
{
    if (!c || !n_out) return GLINT_ERR_INVALID_ARGUMENT;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    // after evaluate the current tick's pairs are the previous generation
    const TableGen &tab = c->tab[c->evaluated ? 1 - c->cur_gen : c->cur_gen];
    uint32_t n = 0;
    HIP_TRY(c, hipMemcpy(&n, tab.rstart + c->s.rs_total, 4, hipMemcpyDeviceToHost));
    std::vector<unsigned long long> hk(n);
    std::vector<uint32_t> hs(n);
    if (n) {
        HIP_TRY(c, hipMemcpy(hk.data(), tab.keys, (size_t)n * 8, hipMemcpyDeviceToHost));
        HIP_TRY(c, hipMemcpy(hs.data(), tab.starts, (size_t)n * 4, hipMemcpyDeviceToHost));
    }
    // without store_all the generation also holds some pairs first seen at
    // the last tick (border and tagged ones); they are listed only with it
    const uint32_t last = c->prev_tick;
    std::vector<std::pair<uint64_t, uint32_t>> v;
    for (uint32_t i = 0; i < n; i++)
        if (hs[i] != kDupStart && (c->s.store_all || hs[i] != last)) v.emplace_back(hk[i] & kKeyBits, hs[i]);
    std::sort(v.begin(), v.end());
    *n_out = v.size();
    if (!keys && !starts) return GLINT_OK;
    if (v.size() > cap) return GLINT_ERR_OUTPUT_CAPACITY;
    for (size_t k = 0; k < v.size(); k++) {
        if (keys) keys[k] = v[k].first;
        if (starts) starts[k] = v[k].second;
    }
    return GLINT_OK;
}

glint_status glint_device_meetings(glint_ctx *c, const glint_meeting **d_out, uint64_t *n_out)
{
    if (!c || !d_out || !n_out) return GLINT_ERR_INVALID_ARGUMENT;
    glint_status st = read_counters(c);
    if (st) return st;
    *d_out = c->s.meet;
    *n_out = std::min<uint64_t>(c->h_ctr->n_meet, c->s.meet_cap);
    return status_from_flags(c, c->h_ctr->error);
}

glint_status glint_set_kernel_profiling(glint_ctx *c, int enable)
{
    if (!c) return GLINT_ERR_INVALID_ARGUMENT;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->prof.resolve();
    c->prof.on = enable != 0;
    return GLINT_OK;
}

glint_status glint_kernel_times(glint_ctx *c, const char **names, double *total_ms, uint64_t *launches,
                                uint32_t cap, uint32_t *n_out, int reset)
{
    if (!c || !n_out) return GLINT_ERR_INVALID_ARGUMENT;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->prof.resolve();
    *n_out = K_COUNT;
    for (uint32_t k = 0; k < K_COUNT && k < cap; k++) {
        if (names) names[k] = kKernelNames[k];
        if (total_ms) total_ms[k] = c->prof.total_ms[k];
        if (launches) launches[k] = c->prof.launches[k];
    }
    if (reset) {
        for (int k = 0; k < K_COUNT; k++) {
            c->prof.total_ms[k] = 0;
            c->prof.launches[k] = 0;
        }
    }
    return GLINT_OK;
}

glint_status glint_totals_get(glint_ctx *c, glint_totals *out, int reset)
{
    if (!c || !out) return GLINT_ERR_INVALID_ARGUMENT;
    glint_status st = read_counters(c);
    if (st) return st;
    const Counters &h = *c->h_ctr;
    out->ticks = h.acc_ticks;
    out->contacts = h.acc_contacts;
    out->contacts_border = h.acc_contacts_border;
    out->distinct_pairs = h.acc_distinct;
    out->meetings = h.acc_meet;
    out->units = h.acc_units;
    out->border_objects = h.acc_border;
    out->rechecks = h.acc_rechecks;
    out->max_contacts_tick = h.max_contacts;
    out->max_distinct_tick = h.max_distinct;
    out->stored = h.acc_stored;
    out->max_stored_tick = h.max_stored;
    out->max_meetings_tick = h.max_meet;
    out->ref_drop_leaves = h.acc_drop_leaves;
    out->ref_drop_objects = h.acc_drop_objs;
    out->max_ref_drop_objects_tick = h.max_drop_objs;
    out->max_leaf_count = h.max_leaf_run;
    out->tests = h.acc_tests;
    out->evaluated = h.acc_evals;
    if (reset) {
        size_t off = offsetof(Counters, acc_ticks);
        HIP_TRY(c, hipMemsetAsync((char *)c->s.ctr + off, 0, sizeof(Counters) - off, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
    }
    return status_from_flags(c, h.error);
}

glint_status glint_batch_destroy(glint_batch *b)
{
    if (!b) return GLINT_ERR_INVALID_ARGUMENT;
    if (b->origin) (void)hipStreamSynchronize(b->origin);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    for (hipEvent_t e : b->ready)
        if (e) (void)hipEventDestroy(e);
    if (b->after) (void)hipEventDestroy(b->after);
    for (DevState *d : b->group_states)
        if (d) (void)hipFree(d);
    for (glint_ctx *c : b->ctx) glint_destroy(c);
    if (b->origin) (void)hipStreamDestroy(b->origin);
    delete b;
    return GLINT_OK;
}

glint_status glint_batch_create(const glint_config *cfgs, uint32_t n, uint32_t graph_ticks, glint_batch **out)
{
    if (!cfgs || !out || n == 0) return GLINT_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    glint_batch *b = new (std::nothrow) glint_batch();
    if (!b) return GLINT_ERR_OUT_OF_MEMORY;
    b->graph_ticks = graph_ticks;
    auto fail = [&](glint_status st, const std::string &why) {
        g_create_error = why;
        glint_batch_destroy(b);
        return st;
    };
    for (uint32_t i = 0; i < n; i++) {
        if (cfgs[i].device != cfgs[0].device) return fail(GLINT_ERR_INVALID_ARGUMENT, "one device per batch");
        glint_config c = cfgs[i];
        c.graph_ticks = 0;   // the batch captures its own graph
        glint_ctx *ctx = nullptr;
        glint_status st = glint_create(&c, &ctx);
        if (st) return fail(st, "scenario " + std::to_string(i) + ": " + g_create_error);
        b->ctx.push_back(ctx);
    }
    // contexts alike in every kernel variant share fused launches, at most
    // kMaxFused to a group, in the order given
    std::vector<std::vector<uint64_t>> keys;
    for (uint32_t i = 0; i < n; i++) {
        const std::vector<uint64_t> k = variant_key(b->ctx[i]);
        size_t gi = 0;
        while (gi < keys.size() && (keys[gi] != k || b->groups[gi].size() == kMaxFused)) gi++;
        if (gi == keys.size()) {
            keys.push_back(k);
            b->groups.emplace_back();
        }
        b->groups[gi].push_back(i);
    }
    if (hipSetDevice((int)cfgs[0].device) != hipSuccess ||
        hipStreamCreateWithFlags(&b->origin, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->after, hipEventDisableTiming) != hipSuccess)
        return fail(GLINT_ERR_HIP, "batch stream or event creation failed");
    b->group_states.assign(b->groups.size(), nullptr);
    for (size_t gi = 0; gi < b->groups.size(); gi++) {
        std::vector<DevState> h;
        for (uint32_t i : b->groups[gi]) h.push_back(b->ctx[i]->s);
        if (hipMalloc((void **)&b->group_states[gi], h.size() * sizeof(DevState)) != hipSuccess ||
            hipMemcpy(b->group_states[gi], h.data(), h.size() * sizeof(DevState), hipMemcpyHostToDevice) !=
                hipSuccess)
            return fail(GLINT_ERR_HIP, "batch state upload failed");
    }
    b->ready.assign(n, nullptr);
    for (uint32_t i = 0; i < n; i++)
        if (hipEventCreateWithFlags(&b->ready[i], hipEventDisableTiming) != hipSuccess)
            return fail(GLINT_ERR_HIP, "batch event creation failed");
    *out = b;
    return GLINT_OK;
}

glint_status glint_batch_context(glint_batch *b, uint32_t i, glint_ctx **out)
{
    if (!b || !out || i >= b->ctx.size()) return GLINT_ERR_INVALID_ARGUMENT;
    *out = b->ctx[i];
    return GLINT_OK;
}

const char *glint_batch_last_error(const glint_batch *b) { return b ? b->err.c_str() : g_create_error.c_str(); }

glint_status glint_batch_groups(const glint_batch *b, uint32_t *groups)
{
    if (!b || !groups) return GLINT_ERR_INVALID_ARGUMENT;
    *groups = (uint32_t)b->groups.size();
    return GLINT_OK;
}

glint_status glint_batch_run_ticks(glint_batch *b, const double *const *d_traces, uint32_t first_tick,
                                   uint32_t nticks, float *tick_ms)
{
    if (!b || !d_traces || nticks == 0) return GLINT_ERR_INVALID_ARGUMENT;
    const uint32_t n = (uint32_t)b->ctx.size();
    auto ctx_fail = [&](uint32_t i, glint_status st) {
        b->err = "scenario " + std::to_string(i) + ": " + b->ctx[i]->err;
        return st;
    };
#define BATCH_TRY(call)                                                        \
    do {                                                                       \
        hipError_t e_ = (call);                                                \
        if (e_ != hipSuccess) {                                                \
            b->err = std::string(#call) + ": " + hipGetErrorString(e_);        \
            return GLINT_ERR_HIP;                                              \
        }                                                                      \
    } while (0)
    for (uint32_t i = 0; i < n; i++) {
        if (!d_traces[i]) return GLINT_ERR_INVALID_ARGUMENT;
        if (b->ctx[i]->prof.on) {
            // fused launches cannot be timed per scenario
            b->err = "scenario " + std::to_string(i) + ": kernel profiling is on; a batch runs without it";
            return GLINT_ERR_STATE;
        }
    }
    for (uint32_t i = 0; i < n; i++) {
        glint_status st = run_begin(b->ctx[i], d_traces[i], first_tick, nticks);
        if (st) return ctx_fail(i, st);
    }
    BATCH_TRY(hipSetDevice(b->ctx[0]->device));
    std::vector<FusedGroup> groups(b->groups.size());
    for (size_t gi = 0; gi < groups.size(); gi++) {
        for (uint32_t i : b->groups[gi]) groups[gi].ctx.push_back(b->ctx[i]);
        groups[gi].d_states = b->group_states[gi];
        groups[gi].stream = b->origin;
    }
    auto enqueue_ticks = [&](uint32_t k) {
        for (; k > 0; k--)
            for (const FusedGroup &g : groups) enqueue_tick_body(g);
    };
    // the batch's stream runs after every scenario's pre-run launches, and
    // the scenarios' streams continue after it
    for (uint32_t i = 0; i < n; i++) {
        BATCH_TRY(hipEventRecord(b->ready[i], b->ctx[i]->stream));
        BATCH_TRY(hipStreamWaitEvent(b->origin, b->ready[i], 0));
    }
    const uint32_t G = b->graph_ticks;
    uint32_t done = 0;
    if (G && nticks >= G) {
        if (!b->exec) {
            // G ticks of every group captured once; every exit after the
            // capture began ends it
            hipGraph_t graph = nullptr;
            BATCH_TRY(hipStreamBeginCapture(b->origin, hipStreamCaptureModeThreadLocal));
            enqueue_ticks(G);
            hipError_t e = hipGetLastError();
            const hipError_t ee = hipStreamEndCapture(b->origin, &graph);
            if (e == hipSuccess) e = ee;
            if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0);
            if (graph) (void)hipGraphDestroy(graph);
            if (e != hipSuccess) {
                b->exec = nullptr;
                b->err = std::string("graph capture of the batch: ") + hipGetErrorString(e);
                return GLINT_ERR_HIP;
            }
        }
        for (; done + G <= nticks; done += G) BATCH_TRY(hipGraphLaunch(b->exec, b->origin));
    }
    enqueue_ticks(nticks - done);
    BATCH_TRY(hipGetLastError());
    BATCH_TRY(hipEventRecord(b->after, b->origin));
    for (uint32_t i = 0; i < n; i++) BATCH_TRY(hipStreamWaitEvent(b->ctx[i]->stream, b->after, 0));
#undef BATCH_TRY
    glint_status first_err = GLINT_OK;
    std::vector<std::vector<unsigned long long>> stamps(n);
    for (uint32_t i = 0; i < n; i++) {
        glint_ctx *c = b->ctx[i];
        run_end_state(c, first_tick, nticks);
        glint_status st = run_finish(c, nticks, nullptr);
        if (st && !first_err) first_err = ctx_fail(i, st);
        if (tick_ms) {
            stamps[i].resize(2 * (size_t)nticks);
            if (hipMemcpy(stamps[i].data(), c->stamps, 2 * (size_t)nticks * sizeof(unsigned long long),
                          hipMemcpyDeviceToHost) != hipSuccess && !first_err) {
                b->err = "stamp copy failed";
                first_err = GLINT_ERR_HIP;
            }
        }
    }
    if (tick_ms && !first_err) {
        // each scenario's tick times: device wall clock from the start of its
        // group's tick to its end
        const double khz = b->ctx[0]->wall_khz > 0 ? b->ctx[0]->wall_khz : 100000.0;
        for (uint32_t i = 0; i < n; i++)
            for (uint32_t k = 0; k < nticks; k++)
                tick_ms[(size_t)i * nticks + k] =
                    (float)((double)(stamps[i][2 * k + 1] - stamps[i][2 * k]) / khz);
    }
    return first_err;
}

glint_status glint_meeting_log_info(glint_ctx *c, uint64_t *total, uint64_t *first_retained)
{
    if (!c || !total || !first_retained) return GLINT_ERR_INVALID_ARGUMENT;
    if (!c->s.mlog) {
        c->err = "meeting log is off (meeting_log_capacity = 0)";
        return GLINT_ERR_STATE;
    }
    HIP_TRY(c, hipSetDevice(c->device));
    glint_status st = read_counters(c);
    if (st) return st;
    *total = c->h_ctr->n_mlog;
    *first_retained = *total > c->s.mlog_cap ? *total - c->s.mlog_cap : 0;
    return GLINT_OK;
}

glint_status glint_meeting_log_read(glint_ctx *c, uint64_t from, glint_meeting *out, uint64_t cap,
                                    uint64_t *n_out)
{
    if (!c || !n_out || (cap && !out)) return GLINT_ERR_INVALID_ARGUMENT;
    uint64_t total = 0, first = 0;
    glint_status st = glint_meeting_log_info(c, &total, &first);
    if (st) return st;
    if (from < first || from > total) {
        c->err = "meeting log range no longer retained (or past its end)";
        return GLINT_ERR_INVALID_ARGUMENT;
    }
    const uint64_t n = std::min<uint64_t>(total - from, cap);
    const uint64_t L = c->s.mlog_cap, a = from % L;
    const uint64_t n1 = std::min<uint64_t>(n, L - a);
    if (n1) HIP_TRY(c, hipMemcpy(out, c->s.mlog + a, n1 * sizeof(glint_meeting), hipMemcpyDeviceToHost));
    if (n > n1) HIP_TRY(c, hipMemcpy(out + n1, c->s.mlog, (n - n1) * sizeof(glint_meeting), hipMemcpyDeviceToHost));
    *n_out = n;
    return GLINT_OK;
}

}  // extern "C"

// ---------------------------------------------------------- snapshot
//
// The state the next tick reads, at a tick boundary: the tree (nodes, their
// metadata, the free-node stack, parent links, leaf counts), the last tick's
// objects in leaf order (ids, leaves, positions of the last two ticks), the
// candidate lists of the last two ticks, the position histories, the last
// tick's generation of the meeting table (its
// stored contacts, start ticks and region offsets; with the key-value output
// their boxes), the counters, the meeting log and the key-value log.  The
// reference dumps its meeting buckets and metadata to disk and loads them
// back into the table (workbench.cpp:446-470,574-595; geometry.cu:543-565).

namespace {

struct SnapHeader {
    uint64_t magic, bytes;
    uint32_t n, node_cap, rs_total, precision, store_all, kv, dynamic_schema, cur_gen;
    uint32_t built, has_prev, tick, prev_tick, kv_end_time_min, nsections;
    uint64_t clist_cap, mlog_cap, kvlog_cap, gen_count, mlog_count, kvlog_count;
};
constexpr uint64_t kSnapMagic = 0x31746e696c67ull;   // "glint1"

struct Section {
    void *dev;
    uint64_t bytes;
};

// The device ranges of a snapshot, in order (gen, mlog and kvlog counts as given).
std::vector<Section> state_sections(glint_ctx *c, uint64_t gen_count, uint64_t mlog_count, uint64_t kvlog_count)
{
    DevState &s = c->s;
    const uint64_t N = s.n, NC = s.node_cap;
    const int g = 1 - c->cur_gen;   // after evaluate the last tick's pairs are in the other generation
    std::vector<Section> v = {
        {s.nodes, NC * sizeof(Node)},      {s.meta, NC * sizeof(NodeMeta)}, {s.free_stack, NC * 4},
        {s.parent, NC * 4},                {s.count, NC * 4},               {s.spid, N * 4},
        {s.sleafb[0], N * 4},              {s.sleafb[1], N * 4},            {s.sxb[0], N * 8},
        {s.sxb[1], N * 8},                 {s.syb[0], N * 8},               {s.syb[1], N * 8},
        {c->hist[0], N * 16},              {c->hist[1], N * 16},            {s.ctr, sizeof(Counters)},
        {c->tab[g].keys, gen_count * 8},   {c->tab[g].starts, gen_count * 4},
        {c->tab[g].rstart, ((uint64_t)s.rs_total + 2) * 4},
    };
    if (s.cand) {
        v.push_back({s.cand, N * sizeof(uint4)});
        v.push_back({s.prev_cand, N * sizeof(uint4)});
    }
    if (c->kv) {
        v.push_back({c->mbr[g], gen_count * sizeof(float4)});
        v.push_back({s.kvlog, kvlog_count * sizeof(KvLogEntry)});
    }
    if (s.mlog) v.push_back({s.mlog, mlog_count * sizeof(glint_meeting)});
    return v;
}

glint_status snapshot_header(glint_ctx *c, SnapHeader &h)
{
    if (!c->evaluated || !c->has_prev) {
        c->err = "a snapshot is taken after an evaluated tick (at a tick boundary)";
        return GLINT_ERR_STATE;
    }
    glint_status st = read_counters(c);
    if (st) return st;
    DevState &s = c->s;
    const int g = 1 - c->cur_gen;
    uint32_t cnt = 0;
    HIP_TRY(c, hipMemcpy(&cnt, c->tab[g].rstart + s.rs_total, 4, hipMemcpyDeviceToHost));
    h = SnapHeader{};
    h.magic = kSnapMagic;
    h.n = s.n;
    h.node_cap = s.node_cap;
    h.rs_total = s.rs_total;
    h.precision = s.precision;
    h.store_all = s.store_all;
    h.kv = c->kv != nullptr;
    h.dynamic_schema = c->cfg.dynamic_schema;
    h.cur_gen = (uint32_t)c->cur_gen;
    h.built = c->built;
    h.has_prev = c->has_prev;
    h.tick = c->tick;
    h.prev_tick = c->prev_tick;
    h.kv_end_time_min = c->kv ? kv_end_time_min(c->kv) : 0;
    h.clist_cap = s.clist_cap;
    h.mlog_cap = s.mlog_cap;
    h.kvlog_cap = s.kvlog_cap;
    h.gen_count = cnt;
    h.mlog_count = s.mlog ? std::min<uint64_t>(c->h_ctr->n_mlog, s.mlog_cap) : 0;
    h.kvlog_count = c->kv ? std::min<uint64_t>(c->h_ctr->n_kvlog, s.kvlog_cap) : 0;
    const std::vector<Section> v = state_sections(c, h.gen_count, h.mlog_count, h.kvlog_count);
    h.nsections = (uint32_t)v.size();
    h.bytes = sizeof(SnapHeader);
    for (const Section &x : v) h.bytes += x.bytes;
    return GLINT_OK;
}

}  // namespace

extern "C" {

glint_status glint_snapshot_size(glint_ctx *c, uint64_t *bytes)
{
    if (!c || !bytes) return GLINT_ERR_INVALID_ARGUMENT;
    HIP_TRY(c, hipSetDevice(c->device));
    SnapHeader h;
    glint_status st = snapshot_header(c, h);
    if (st) return st;
    *bytes = h.bytes;
    return GLINT_OK;
}

glint_status glint_snapshot(glint_ctx *c, void *buf, uint64_t bytes)
{
    if (!c || !buf) return GLINT_ERR_INVALID_ARGUMENT;
    HIP_TRY(c, hipSetDevice(c->device));
    SnapHeader h;
    glint_status st = snapshot_header(c, h);
    if (st) return st;
    if (bytes < h.bytes) {
        c->err = "snapshot buffer too small (glint_snapshot_size)";
        return GLINT_ERR_OUTPUT_CAPACITY;
    }
    uint8_t *p = static_cast<uint8_t *>(buf);
    std::memcpy(p, &h, sizeof h);
    p += sizeof h;
    for (const Section &x : state_sections(c, h.gen_count, h.mlog_count, h.kvlog_count)) {
        if (x.bytes) HIP_TRY(c, hipMemcpy(p, x.dev, x.bytes, hipMemcpyDeviceToHost));
        p += x.bytes;
    }
    return GLINT_OK;
}

glint_status glint_restore(glint_ctx *c, const void *buf, uint64_t bytes)
{
    if (!c || !buf || bytes < sizeof(SnapHeader)) return GLINT_ERR_INVALID_ARGUMENT;
    SnapHeader h;
    std::memcpy(&h, buf, sizeof h);
    DevState &s = c->s;
    if (h.magic != kSnapMagic || h.bytes != bytes || h.n != s.n || h.node_cap != s.node_cap ||
        h.rs_total != s.rs_total || h.precision != s.precision || h.store_all != s.store_all ||
        h.kv != (c->kv != nullptr) || h.dynamic_schema != c->cfg.dynamic_schema || h.clist_cap != s.clist_cap ||
        h.mlog_cap != s.mlog_cap || h.kvlog_cap != s.kvlog_cap || h.gen_count > s.clist_cap || h.cur_gen > 1) {
        c->err = "the snapshot does not belong to a context of this configuration";
        return GLINT_ERR_INVALID_ARGUMENT;
    }
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->cur_gen = (int)h.cur_gen;   // the sections follow the generation parity of the snapshot
    const uint8_t *p = static_cast<const uint8_t *>(buf) + sizeof h;
    const std::vector<Section> v = state_sections(c, h.gen_count, h.mlog_count, h.kvlog_count);
    if (v.size() != h.nsections) {
        c->err = "the snapshot's sections do not match this context";
        return GLINT_ERR_INVALID_ARGUMENT;
    }
    for (const Section &x : v) {
        if (x.bytes) HIP_TRY(c, hipMemcpy(x.dev, p, x.bytes, hipMemcpyHostToDevice));
        p += x.bytes;
    }
    c->built = h.built != 0;
    c->has_prev = h.has_prev != 0;
    c->tick = h.tick;
    c->prev_tick = h.prev_tick;
    c->has_positions = c->partitioned = c->refined = c->evaluated = true;
    // the last tick's positions (the history the next tick reads) become the
    // resident ones, from which update ingestion continues
    HIP_TRY(c, hipMemcpy(c->own_xy, c->hist[1 - c->cur_gen], (size_t)s.n * 16, hipMemcpyDeviceToDevice));
    c->last_xy = c->own_xy;
    c->resident = true;
    if (c->kv) kv_set_end_time_min(c->kv, h.kv_end_time_min);
    glint_status st = read_counters(c);
    if (st) return st;
    fill_stats(c);
    return status_from_flags(c, c->h_ctr->error);
}

glint_status glint_kv_pending(glint_ctx *c, uint64_t *keys)
{
    if (!c || !keys) return GLINT_ERR_INVALID_ARGUMENT;
    if (!c->kv) {
        c->err = "key-value output is off (kv_capacity = 0)";
        return GLINT_ERR_STATE;
    }
    HIP_TRY(c, hipSetDevice(c->device));
    glint_status st = read_counters(c);
    if (st) return st;
    *keys = 2 * std::min<uint64_t>(c->h_ctr->n_kvlog, kv_log_capacity(c->kv));
    return status_from_flags(c, c->h_ctr->error);
}

glint_status glint_kv_flush(glint_ctx *c, glint_kv_summary *out)
{
    if (!c) return GLINT_ERR_INVALID_ARGUMENT;
    if (!c->kv) {
        c->err = "key-value output is off (kv_capacity = 0)";
        return GLINT_ERR_STATE;
    }
    if (!c->has_prev || (c->has_positions && !c->evaluated)) {
        c->err = "flush needs an evaluated tick";
        return GLINT_ERR_STATE;
    }
    HIP_TRY(c, hipSetDevice(c->device));
    glint_status st = read_counters(c);
    if (st) return st;
    if (c->h_ctr->error) return status_from_flags(c, c->h_ctr->error);
    st = kv_flush(c->kv, c->h_ctr->n_kvlog, c->prev_tick, out, c->err);
    if (st) return st;
    // the log starts over (the flush runs outside any captured graph)
    HIP_TRY(c, hipMemsetAsync(&c->s.ctr->n_kvlog, 0, sizeof(unsigned long long), c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return GLINT_OK;
}

glint_status glint_kv_part_size(glint_ctx *c, uint32_t part, uint64_t *bytes)
{
    if (!c || !bytes || part > GLINT_KV_PACKED) return GLINT_ERR_INVALID_ARGUMENT;
    if (!c->kv) {
        c->err = "key-value output is off (kv_capacity = 0)";
        return GLINT_ERR_STATE;
    }
    *bytes = kv_part_bytes(c->kv, part);
    return GLINT_OK;
}

glint_status glint_kv_copy(glint_ctx *c, uint32_t part, void *dst, uint64_t bytes)
{
    if (!c) return GLINT_ERR_INVALID_ARGUMENT;
    if (!c->kv) {
        c->err = "key-value output is off (kv_capacity = 0)";
        return GLINT_ERR_STATE;
    }
    HIP_TRY(c, hipSetDevice(c->device));
    return kv_copy(c->kv, part, dst, bytes, c->err);
}

glint_status glint_debug_inject_failure(glint_ctx *c, const char *stage)
{
    if (!c || !stage) return GLINT_ERR_INVALID_ARGUMENT;
    std::string s(stage);
    if (s == "partition")
        c->inject |= kInjPartition;
    else if (s == "refine")
        c->inject |= kInjRefine;
    else if (s == "evaluate")
        c->inject |= kInjEvaluate;
    else if (s == "update_schema")
        c->inject |= kInjSchema;
    else
        return GLINT_ERR_INVALID_ARGUMENT;
    return GLINT_OK;
}

glint_status glint_device_malloc(void **ptr, uint64_t bytes, uint32_t device)
{
    if (!ptr) return GLINT_ERR_INVALID_ARGUMENT;
    if (hipSetDevice((int)device) != hipSuccess) return GLINT_ERR_HIP;
    if (hipMalloc(ptr, bytes ? bytes : 1) != hipSuccess) {
        (void)hipGetLastError();
        return GLINT_ERR_OUT_OF_MEMORY;
    }
    return GLINT_OK;
}

namespace {
__global__ void k_copy(const uint4 *src, uint4 *dst, uint64_t n16)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n16;
         i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}
}  // namespace

glint_status glint_device_copy(void *dst, const void *src, uint64_t bytes)
{
    if (!dst || !src || (bytes % 16) || ((uintptr_t)dst % 16) || ((uintptr_t)src % 16))
        return GLINT_ERR_INVALID_ARGUMENT;
    uint64_t n16 = bytes / 16;
    k_copy<<<grid_for(n16, 16384), kBlock>>>((const uint4 *)src, (uint4 *)dst, n16);
    if (hipGetLastError() != hipSuccess) return GLINT_ERR_HIP;
    return hipDeviceSynchronize() == hipSuccess ? GLINT_OK : GLINT_ERR_HIP;
}

glint_status glint_device_free(void *ptr)
{
    return hipFree(ptr) == hipSuccess ? GLINT_OK : GLINT_ERR_HIP;
}

glint_status glint_memcpy_to_device(void *dst, const void *src, uint64_t bytes)
{
    return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess ? GLINT_OK : GLINT_ERR_HIP;
}

glint_status glint_memcpy_to_host(void *dst, const void *src, uint64_t bytes)
{
    return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess ? GLINT_OK : GLINT_ERR_HIP;
}

glint_status glint_device_count(int *count)
{
    if (!count) return GLINT_ERR_INVALID_ARGUMENT;
    hipError_t e = hipGetDeviceCount(count);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *count = 0;
    }
    return GLINT_OK;
}

}  // extern "C"
