// kv.hip - sorted key-value meeting output on gfx950.
//
// The flush branch of the reference's GPU path (src/cuda/geometry.cu:
// 1752-2138 with the kernels at 665-1107, the CTF encoding of
// src/tracing/step_merge.cpp:140-310), re-planned for one device:
//  - the table update only logs ended meetings (32 B each); the per-object
//    aggregates the reference updates with atomics during evaluation (key
//    counts, uint boxes, duration-oversize flags) are rebuilt from the log at
//    flush time, which keeps them out of the per-tick path;
//  - every sort is a stable LSD radix sort over only the bits that differ
//    across the keys (rocPRIM's onesweep sort of 64-bit keys with 32-bit
//    values; a 128-bit key is sorted by its low word, then stably by its high
//    word, the rows following through index permutations); the STR slab sort
//    is one sort of (slab + 1, y) keys;
//  - per-file limits, bitmaps (in LDS) and packed keys run one workgroup per
//    file.
// Results are the reference's: the final key order is the total order of the
// unique keys; the STR sorts are stable, as the reference's radix sorts are.
// Where the reference is undefined (a zero-width grid, cells past the grid, a
// key range past the kept keys, an empty file) the engine does what the
// oracle states.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>

#include "kv_engine.h"

using glint::KvLogEntry;

namespace {

constexpr uint32_t kB = 256;
constexpr uint32_t kBitmapEdge = 256;
constexpr uint32_t kBitmapBytes = kBitmapEdge * kBitmapEdge / 8;   // bench->bit_count / 8 per file
constexpr uint64_t kLow52 = (1ull << 52) - 1;     // key bits below the sid (bit 116 = hi bit 52)
constexpr uint32_t kOidMask = (1u << 26) - 1;

struct KvCtr {
    unsigned long long nk, no;
    unsigned long long and_hi, and_lo, or_hi, or_lo;
    uint32_t sid_count, ovs_meet, ovs_area, osmin, osmax, pad;
};

struct SortView {
    uint64_t *hi;   // null for 64-bit keys
    uint64_t *lo;
    uint32_t *val;
};

__device__ __forceinline__ uint32_t lane() { return __lane_id(); }

__device__ __forceinline__ uint32_t f2u(float v)
{
    // float_to_uint (cuda_util.cuh:115-119)
    v += 180.0f;
    return (uint32_t)(v * 100000.0f);
}

__device__ __forceinline__ float u2f(uint32_t f)
{
    // uint_to_float (cuda_util.cuh:121-125)
    return (float)f / 100000.0f - 180.0f;
}

__device__ __forceinline__ float ratio(float num, float den) { return den == 0.0f ? 0.0f : num / den; }

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *total, uint32_t *wsum)
{
    uint32_t x = v;
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(x, d);
        if (lane() >= (uint32_t)d) x += y;
    }
    if (lane() == 63) wsum[threadIdx.x / 64] = x;
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (uint32_t w = 0; w < kB / 64; w++) {
        if (w < threadIdx.x / 64) before += wsum[w];
        all += wsum[w];
    }
    __syncthreads();
    *total = all;
    return before + x - v;
}

// --------------------------------------------------------------- radix sort

__global__ __launch_bounds__(kB) void k_iota(uint32_t *p, uint64_t n)
{
    for (uint64_t i = blockIdx.x * (uint64_t)kB + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kB)
        p[i] = (uint32_t)i;
}

// dst64[i] = src64[perm[i]], dst32[i] = src32[perm[i]].
__global__ __launch_bounds__(kB) void k_permute(const uint32_t *perm, const uint64_t *src64, uint64_t *dst64,
                                                const uint32_t *src32, uint32_t *dst32, uint64_t n)
{
    for (uint64_t i = blockIdx.x * (uint64_t)kB + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kB) {
        const uint32_t j = perm[i];
        dst64[i] = src64[j];
        dst32[i] = src32[j];
    }
}

// AND and OR of all keys: the bits that differ decide which digits to sort.
__global__ __launch_bounds__(kB) void k_andor(const uint64_t *hi, const uint64_t *lo, uint64_t n, KvCtr *ctr)
{
    uint64_t ah = ~0ull, al = ~0ull, oh = 0, ol = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)kB + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kB) {
        const uint64_t h = hi ? hi[i] : 0, l = lo[i];
        ah &= h;
        al &= l;
        oh |= h;
        ol |= l;
    }
    for (int d = 32; d >= 1; d >>= 1) {
        ah &= __shfl_xor(ah, d);
        al &= __shfl_xor(al, d);
        oh |= __shfl_xor(oh, d);
        ol |= __shfl_xor(ol, d);
    }
    if (lane() == 0) {
        atomicAnd(&ctr->and_hi, (unsigned long long)ah);
        atomicAnd(&ctr->and_lo, (unsigned long long)al);
        atomicOr(&ctr->or_hi, (unsigned long long)oh);
        atomicOr(&ctr->or_lo, (unsigned long long)ol);
    }
}

__global__ void k_ctr_reset_masks(KvCtr *c)
{
    c->and_hi = ~0ull;
    c->and_lo = ~0ull;
    c->or_hi = 0;
    c->or_lo = 0;
}

__global__ void k_ctr_reset(KvCtr *c)
{
    c->nk = 0;
    c->no = 0;
    c->sid_count = 0;
    c->ovs_meet = 0;
    c->ovs_area = 0;
    c->osmin = 1u << 30;
    c->osmax = 0;
}

// ------------------------------------------------------------- flush kernels

__global__ __launch_bounds__(kB) void k_obj_init(uint32_t n, uint32_t *cnt, uint4 *obox, uint32_t *sids)
{
    for (uint32_t i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB) {
        cnt[i] = 0;
        obox[i] = make_uint4(100000000u, 100000000u, 0u, 0u);   // cuda_init_o_boxs (geometry.cu:1254-1263)
        sids[i] = 0;
    }
}

// Two directed keys per logged meeting, and the per-object aggregates the
// reference keeps during evaluation (geometry.cu:704-735).
__global__ __launch_bounds__(kB) void k_expand(const KvLogEntry *log, uint64_t nlog, SortView out, float4 *boxes,
                                               uint32_t *cnt, uint4 *obox, uint32_t *sids, KvCtr *ctr,
                                               uint32_t end_time_min, uint32_t max_meet)
{
    uint32_t ovs = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)kB + threadIdx.x; i < nlog; i += (uint64_t)gridDim.x * kB) {
        const KvLogEntry e = log[i];
        const uint32_t dur = e.end - e.start;
        if (dur > max_meet) {
            ovs++;
            sids[e.pid_a] = 1;
            sids[e.pid_b] = 1;
        }
        const uint32_t bx0 = f2u(e.box.x), by0 = f2u(e.box.y), bx1 = f2u(e.box.z), by1 = f2u(e.box.w);
        for (uint32_t dir = 0; dir < 2; dir++) {
            const uint32_t pid = dir ? e.pid_b : e.pid_a, target = dir ? e.pid_a : e.pid_b;
            const uint64_t j = 2 * i + dir;
            out.hi[j] = ((uint64_t)pid << 26) + target;                      // bits 90 and 64
            out.lo[j] = ((uint64_t)dur << 32) + (uint32_t)(e.end - end_time_min);
            out.val[j] = (uint32_t)j;
            boxes[j] = e.box;
            atomicAdd(&cnt[pid], 1u);
            // the box only shrinks towards its final value: an update that
            // the current value already covers needs no atomic
            const uint4 cur = obox[pid];   // an older value only costs an atomic more
            if (bx0 < cur.x) atomicMin(&obox[pid].x, bx0);
            if (by0 < cur.y) atomicMin(&obox[pid].y, by0);
            if (bx1 > cur.z) atomicMax(&obox[pid].z, bx1);
            if (by1 > cur.w) atomicMax(&obox[pid].w, by1);
        }
    }
    for (int d = 32; d >= 1; d >>= 1) ovs += __shfl_xor(ovs, d);
    if (lane() == 0 && ovs) atomicAdd(&ctr->ovs_meet, ovs);
}

// set_oid (geometry.cu:740-766): area test and STR mid of every object.
__global__ __launch_bounds__(kB) void k_objects(uint32_t n, const uint32_t *cnt, const uint4 *obox, uint32_t *sids,
                                                uint64_t *mid, uint32_t *oid, double area_limit, KvCtr *ctr)
{
    uint32_t present = 0, cut = 0;
    for (uint32_t i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB) {
        oid[i] = i;
        uint64_t m = 0;
        if (cnt[i] >= 1) {
            present++;
            const uint4 o = obox[i];
            const float l0 = u2f(o.x), h0 = u2f(o.z), l1 = u2f(o.y), h1 = u2f(o.w);
            const float area = (h1 - l1) * (h0 - l0);
            if ((double)area > area_limit) {
                cut++;
                sids[i] = 1;
            } else {
                const uint32_t m0 = o.x / 2 + o.z / 2;
                const uint32_t m1 = o.w / 2 + o.w / 2;   // high1 twice (geometry.cu:764)
                m = ((uint64_t)m0 << 32) + m1;
            }
        }
        mid[i] = m;
    }
    for (int d = 32; d >= 1; d >>= 1) {
        present += __shfl_xor(present, d);
        cut += __shfl_xor(cut, d);
    }
    if (lane() == 0) {
        if (present) atomicAdd(&ctr->sid_count, present);
        if (cut) atomicAdd(&ctr->ovs_area, cut);
    }
}

// Keys of sid-1 objects to the tail, the rest to the head (geometry.cu:1776-1797).
// Every workgroup takes 4096 keys at a time and reserves its places with one
// atomic per side (the order within each side is free: both are sorted
// afterwards).
constexpr uint32_t kPartItems = 16;
__global__ __launch_bounds__(kB) void k_partition(SortView in, SortView out, uint64_t n, const uint32_t *sids,
                                                  KvCtr *ctr)
{
    __shared__ uint32_t wk[kB / 64], wo[kB / 64];
    __shared__ unsigned long long base_k, base_o;
    const uint32_t w = threadIdx.x / 64;
    const unsigned long long below = (1ull << lane()) - 1ull;
    for (uint64_t t0 = blockIdx.x * (uint64_t)kB * kPartItems; t0 < n; t0 += (uint64_t)gridDim.x * kB * kPartItems) {
        uint64_t h[kPartItems], l[kPartItems];
        uint32_t v[kPartItems];
        uint32_t keepbits = 0, validbits = 0;
#pragma unroll
        for (uint32_t q = 0; q < kPartItems; q++) {
            const uint64_t i = t0 + (uint64_t)q * kB + threadIdx.x;   // coalesced rows
            h[q] = 0;
            l[q] = 0;
            v[q] = 0;
            if (i < n) {
                h[q] = in.hi[i];
                l[q] = in.lo[i];
                v[q] = in.val[i];
                validbits |= 1u << q;
                if (sids[(uint32_t)(h[q] >> 26) & kOidMask] != 1) keepbits |= 1u << q;
            }
        }
        // this wave's counts, then the workgroup's, then one reservation per side
        uint32_t ck = 0, co = 0;
#pragma unroll
        for (uint32_t q = 0; q < kPartItems; q++) {
            ck += (uint32_t)__popcll(__ballot((keepbits >> q) & 1u));
            co += (uint32_t)__popcll(__ballot(((validbits & ~keepbits) >> q) & 1u));
        }
        if (lane() == 0) {
            wk[w] = ck;
            wo[w] = co;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t tk = 0, to = 0;
            for (uint32_t k = 0; k < kB / 64; k++) {
                const uint32_t a = wk[k], b = wo[k];
                wk[k] = tk;
                wo[k] = to;
                tk += a;
                to += b;
            }
            base_k = tk ? atomicAdd(&ctr->nk, (unsigned long long)tk) : 0;
            base_o = to ? atomicAdd(&ctr->no, (unsigned long long)to) : 0;
        }
        __syncthreads();
        unsigned long long pk = base_k + wk[w], po = base_o + wo[w];
#pragma unroll
        for (uint32_t q = 0; q < kPartItems; q++) {
            const bool keep = (keepbits >> q) & 1u, other = ((validbits & ~keepbits) >> q) & 1u;
            const unsigned long long mk = __ballot(keep), mo = __ballot(other);
            if (keep || other) {
                const uint64_t pos = keep ? pk + __popcll(mk & below) : n - 1 - (po + __popcll(mo & below));
                out.hi[pos] = h[q];
                out.lo[pos] = l[q];
                out.val[pos] = v[q];
            }
            pk += (unsigned long long)__popcll(mk);
            po += (unsigned long long)__popcll(mo);
        }
        __syncthreads();
    }
}

// Slab keys for the second STR sort: positions before z keep key 0; the
// rest sort by (slab + 1, low word of mid) (geometry.cu:1841-1880).
__global__ __launch_bounds__(kB) void k_narrow(uint64_t *mid, uint32_t n, uint32_t z, uint32_t xcap, uint32_t split)
{
    for (uint32_t p = blockIdx.x * kB + threadIdx.x; p < n; p += gridDim.x * kB) {
        if (p < z) {
            mid[p] = 0;
        } else {
            const uint32_t slab = min((p - z) / xcap, split - 1);
            mid[p] = ((uint64_t)(slab + 1) << 32) | (mid[p] & 0xFFFFFFFFull);
        }
    }
}

// make_sid (geometry.cu:780-796)
__global__ __launch_bounds__(kB) void k_make_sid(const uint32_t *oid, uint32_t n, uint32_t z, uint32_t xcap,
                                                 uint32_t part, uint32_t split, uint32_t *sids)
{
    for (uint32_t p = blockIdx.x * kB + threadIdx.x; p < n; p += gridDim.x * kB) {
        if (p < z) continue;
        const uint32_t xi = (p - z) / xcap;
        const uint32_t yi = ((p - z) - xi * xcap) / part;
        sids[oid[p]] = xi * split + yi + 2;
    }
}

// write_sid_in_key (geometry.cu:798-806)
__global__ __launch_bounds__(kB) void k_write_sid(uint64_t *hi, uint64_t n, const uint32_t *sids)
{
    for (uint64_t i = blockIdx.x * (uint64_t)kB + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kB) {
        const uint64_t h = hi[i];
        hi[i] = (h & kLow52) + ((uint64_t)(sids[(uint32_t)(h >> 26) & kOidMask] & 0xFFFFu) << 52);
    }
}

__global__ __launch_bounds__(kB) void k_gather_boxes(const uint32_t *val, const float4 *in, float4 *out, uint64_t n)
{
    for (uint64_t i = blockIdx.x * (uint64_t)kB + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kB)
        out[i] = in[val[i]];
}

// get_CTF_capacity (geometry.cu:808-818): per-block histogram in LDS, then
// one global add per file and block.
__global__ __launch_bounds__(kB) void k_capacity(const uint32_t *oid, uint32_t n, uint32_t z, const uint32_t *sids,
                                                 const uint32_t *cnt, uint32_t nctf, uint32_t *cap)
{
    __shared__ uint32_t h[4096];
    for (uint32_t k = threadIdx.x; k < nctf; k += kB) h[k] = 0;
    __syncthreads();
    for (uint32_t p = z + blockIdx.x * kB + threadIdx.x; p < n; p += gridDim.x * kB) {
        const uint32_t o = oid[p];
        atomicAdd(&h[sids[o] - 2], cnt[o]);
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < nctf; k += kB)
        if (h[k]) atomicAdd(&cap[k], h[k]);
}

__global__ void k_fill_u32(uint32_t *p, uint64_t n, uint32_t v)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// cuda_get_limit (geometry.cu:820-890): one workgroup per file.
__global__ __launch_bounds__(kB) void k_limits(const uint32_t *cap, uint32_t nctf, const uint64_t *hi,
                                               const uint64_t *lo, const float4 *boxes, uint64_t nk,
                                               uint32_t end_time_min, uint32_t cur_time, glint_ctf *ctfs)
{
    __shared__ float r4[4][kB / 64];
    __shared__ uint32_t rt[2][kB / 64];
    __shared__ unsigned long long rk[kB / 64];
    const uint32_t t = blockIdx.x;
    unsigned long long before = 0;
    for (uint32_t i = threadIdx.x; i < t; i += kB) before += cap[i];
    for (int d = 32; d >= 1; d >>= 1) before += __shfl_xor(before, d);
    if (lane() == 0) rk[threadIdx.x / 64] = before;
    __syncthreads();
    uint64_t kidx = 0;
    for (uint32_t w = 0; w < kB / 64; w++) kidx += rk[w];
    const uint64_t b0 = min(kidx, nk), b1 = min(kidx + cap[t], nk);
    float lo0 = 100000.0f, lo1 = 100000.0f, hi0 = -100000.0f, hi1 = -100000.0f;
    uint32_t smin = 1u << 30, smax = 0;
    for (uint64_t i = b0 + threadIdx.x; i < b1; i += kB) {
        const float4 b = boxes[i];
        lo0 = fminf(lo0, b.x);
        lo1 = fminf(lo1, b.y);
        hi0 = fmaxf(hi0, b.z);
        hi1 = fmaxf(hi1, b.w);
        const uint32_t dur = (uint32_t)(lo[i] >> 32);
        const uint32_t end = (uint32_t)lo[i] + end_time_min;
        const uint32_t st = end - dur;
        smin = min(smin, st);
        smax = max(smax, st);
    }
    for (int d = 32; d >= 1; d >>= 1) {
        lo0 = fminf(lo0, __shfl_xor(lo0, d));
        lo1 = fminf(lo1, __shfl_xor(lo1, d));
        hi0 = fmaxf(hi0, __shfl_xor(hi0, d));
        hi1 = fmaxf(hi1, __shfl_xor(hi1, d));
        smin = min(smin, (uint32_t)__shfl_xor(smin, d));
        smax = max(smax, (uint32_t)__shfl_xor(smax, d));
    }
    const uint32_t w = threadIdx.x / 64;
    if (lane() == 0) {
        r4[0][w] = lo0;
        r4[1][w] = lo1;
        r4[2][w] = hi0;
        r4[3][w] = hi1;
        rt[0][w] = smin;
        rt[1][w] = smax;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (uint32_t k = 1; k < kB / 64; k++) {
            lo0 = fminf(lo0, r4[0][k]);
            lo1 = fminf(lo1, r4[1][k]);
            hi0 = fmaxf(hi0, r4[2][k]);
            hi1 = fmaxf(hi1, r4[3][k]);
            smin = min(smin, rt[0][k]);
            smax = max(smax, rt[1][k]);
        }
        glint_ctf f{};
        f.mbr[0] = lo0;
        f.mbr[1] = lo1;
        f.mbr[2] = hi0;
        f.mbr[3] = hi1;
        f.key_offset = kidx;
        f.kv_count = cap[t];
        f.start_time_min = smin;
        f.start_time_max = smax;
        f.end_time_min = end_time_min;
        f.end_time_max = cur_time;
        ctfs[t] = f;
    }
}

__device__ __forceinline__ void key_range(const glint_ctf &f, uint64_t nk, uint64_t &b0, uint64_t &b1)
{
    b0 = min((uint64_t)f.key_offset, nk);
    b1 = min((uint64_t)f.key_offset + f.kv_count, nk);
}

// cuda_write_bitmap (geometry.cu:992-1019) with cells clipped to the grid;
// gridDim.y slices of every file's keys, each OR-ed in from LDS.
__global__ __launch_bounds__(kB) void k_bitmaps(const glint_ctf *ctfs, const float4 *boxes, uint64_t nk,
                                                uint32_t *bitmaps)
{
    __shared__ uint32_t bm[kBitmapBytes / 4];
    const glint_ctf f = ctfs[blockIdx.x];
    if (!f.key_bit) return;
    for (uint32_t k = threadIdx.x; k < kBitmapBytes / 4; k += kB) bm[k] = 0;
    __syncthreads();
    {
        uint64_t r0, r1;
        key_range(f, nk, r0, r1);
        const uint64_t per = (r1 - r0 + gridDim.y - 1) / gridDim.y;
        const uint64_t b0 = min(r1, r0 + per * blockIdx.y), b1 = min(r1, b0 + per);
        const float w = f.mbr[2] - f.mbr[0], h = f.mbr[3] - f.mbr[1];
        for (uint64_t i = b0 + threadIdx.x; i < b1; i += kB) {
            const float4 b = boxes[i];
            const uint32_t l0 = (uint32_t)(ratio(b.x - f.mbr[0], w) * f.x_grid);
            const uint32_t l1 = (uint32_t)(ratio(b.y - f.mbr[1], h) * f.y_grid);
            const uint32_t h0 = (uint32_t)(ratio(b.z - f.mbr[0], w) * f.x_grid);
            const uint32_t h1 = (uint32_t)(ratio(b.w - f.mbr[1], h) * f.y_grid);
            for (uint32_t x = l0; x <= h0 && x < f.x_grid; x++)
                for (uint32_t y = l1; y <= h1 && y < f.y_grid; y++) {
                    const uint32_t pos = x + y * f.x_grid;
                    atomicOr(&bm[pos >> 5], 1u << (pos & 31));
                }
        }
    }
    __syncthreads();
    uint32_t *out = bitmaps + (size_t)blockIdx.x * (kBitmapBytes / 4);
    for (uint32_t k = threadIdx.x; k < kBitmapBytes / 4; k += kB)
        if (bm[k]) atomicOr(&out[k], bm[k]);
}

__device__ __forceinline__ void pack_one(const glint_ctf &f, uint64_t i, const uint64_t *hi, const uint64_t *lo,
                                         const float4 *boxes, uint8_t *packed)
{
    const uint32_t kb = f.key_bit / 8u;
    const float w = f.mbr[2] - f.mbr[0], h = f.mbr[3] - f.mbr[1];
    const float longer = fmaxf(w, h);
    const float qx = (float)((1ull << f.low_x_bit) - 1), qy = (float)((1ull << f.low_y_bit) - 1);
    const float qe = (float)((1ull << f.edge_bit) - 1);
    const float4 b = boxes[i];
    const uint64_t low0 = (uint64_t)(ratio(b.x - f.mbr[0], w) * qx);
    const uint64_t low1 = (uint64_t)(ratio(b.y - f.mbr[1], h) * qy);
    const uint64_t ex = (uint64_t)(ratio(b.z - b.x, longer) * qe);
    const uint64_t ey = (uint64_t)(ratio(b.w - b.y, longer) * qe);
    const uint64_t vm = (low0 << (f.low_y_bit + 2 * f.edge_bit)) + (low1 << (2 * f.edge_bit)) + (ex << f.edge_bit) + ey;
    const uint32_t s_pid = f.id_bit + f.duration_bit + f.end_bit + f.mbr_bit;
    const uint32_t s_tgt = f.duration_bit + f.end_bit + f.mbr_bit;
    const uint32_t s_dur = f.end_bit + f.mbr_bit;
    const uint64_t h_ = hi[i], l_ = lo[i];
    const unsigned __int128 pid = (unsigned __int128)(h_ >> 26);   // sid bits stay above the oid
    const unsigned __int128 tgt = h_ & kOidMask;
    const unsigned __int128 dur = l_ >> 32;
    const unsigned __int128 end = l_ & 0xFFFFFFFFull;
    const unsigned __int128 v =
        (pid << s_pid) + (tgt << s_tgt) + (dur << s_dur) + (end << f.mbr_bit) + (unsigned __int128)vm;
    uint8_t *o = packed + f.packed_offset + (i - f.key_offset) * kb;
    for (uint32_t j = 0; j < kb; j++) o[j] = (uint8_t)(v >> (8 * j));
}

// write_key_and_mbr (geometry.cu:1076-1107), serialize_mbr (geometry.cu:162-172):
// one thread per kept key; its file is the last one starting at or before it.
__global__ __launch_bounds__(kB) void k_pack(const glint_ctf *ctfs, uint32_t nctf, const uint64_t *hi,
                                             const uint64_t *lo, const float4 *boxes, uint64_t nk, uint8_t *packed)
{
    for (uint64_t i = blockIdx.x * (uint64_t)kB + threadIdx.x; i < nk; i += (uint64_t)gridDim.x * kB) {
        uint32_t a = 0, b = nctf;   // first file with key_offset > i
        while (a < b) {
            const uint32_t m = (a + b) / 2;
            if (ctfs[m].key_offset <= i)
                a = m + 1;
            else
                b = m;
        }
        if (a == 0) continue;
        const glint_ctf &f = ctfs[a - 1];
        if (!f.key_bit || i >= f.key_offset + f.kv_count) continue;
        pack_one(f, i, hi, lo, boxes, packed);
    }
}

// get_o_limit start times (geometry.cu:929-979) and cuda_write_o_bitmap
// (geometry.cu:895-913) over the root box, cells clipped as in
// step_merge.cpp:57-58.
__global__ __launch_bounds__(kB) void k_oversize(const uint64_t *lo, const float4 *boxes, uint64_t b0, uint64_t b1,
                                                 double mlx, double mly, double mhx, double mhy,
                                                 uint32_t end_time_min, uint32_t *obitmap, KvCtr *ctr)
{
    uint32_t smin = 1u << 30, smax = 0;
    for (uint64_t i = b0 + blockIdx.x * (uint64_t)kB + threadIdx.x; i < b1; i += (uint64_t)gridDim.x * kB) {
        const uint32_t dur = (uint32_t)(lo[i] >> 32);
        const uint32_t end = (uint32_t)lo[i] + end_time_min;
        smin = min(smin, end - dur);
        smax = max(smax, end - dur);
        const float4 b = boxes[i];
        const uint32_t l0 = (uint32_t)(((double)b.x - mlx) / (mhx - mlx) * kBitmapEdge);
        const uint32_t l1 = (uint32_t)(((double)b.y - mly) / (mhy - mly) * kBitmapEdge);
        const uint32_t h0 = (uint32_t)(((double)b.z - mlx) / (mhx - mlx) * kBitmapEdge);
        const uint32_t h1 = (uint32_t)(((double)b.w - mly) / (mhy - mly) * kBitmapEdge);
        for (uint32_t x = l0; x <= h0 && x < kBitmapEdge; x++)
            for (uint32_t y = l1; y <= h1 && y < kBitmapEdge; y++) {
                const uint32_t pos = x + y * kBitmapEdge;
                atomicOr(&obitmap[pos >> 5], 1u << (pos & 31));
            }
    }
    for (int d = 32; d >= 1; d >>= 1) {
        smin = min(smin, (uint32_t)__shfl_xor(smin, d));
        smax = max(smax, (uint32_t)__shfl_xor(smax, d));
    }
    if (lane() == 0) {
        atomicMin(&ctr->osmin, smin);
        atomicMax(&ctr->osmax, smax);
    }
}

inline uint32_t grid_of(uint64_t n, uint32_t cap = 4096)
{
    uint64_t g = (n + kB - 1) / kB;
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(g, cap));
}

// ---------------------------------------------------------- file encoding

uint16_t min_bits(uint64_t v)
{
    // min_bits_to_store (util.h:528-533)
    if (v == 0) return 1;
    return (uint16_t)((uint32_t)std::log2((double)v) + 1);
}

// CTF::get_ctf_bits (step_merge.cpp:145-186) on the host, as the reference
// does; false if the key does not fit 128 bits.
bool ctf_bits(glint_ctf &t, const double map[4], uint32_t num_objects, uint32_t max_meet_time)
{
    t.id_bit = min_bits(num_objects);
    t.duration_bit = min_bits((uint32_t)(max_meet_time - 1));
    t.end_bit = t.duration_bit;
    const double gx = (map[2] - map[0]) / 3000;
    const double gy = (map[3] - map[1]) / 4200;
    const float wx = t.mbr[2] - t.mbr[0], wy = t.mbr[3] - t.mbr[1];
    const uint32_t lxg = (uint32_t)((double)wx / gx);
    const uint32_t lyg = (uint32_t)((double)wy / gy);
    t.low_x_bit = min_bits(std::max<uint32_t>(lxg, 1) - 1);
    t.low_y_bit = min_bits(std::max<uint32_t>(lyg, 1) - 1);
    t.edge_bit = std::max(t.low_x_bit, t.low_y_bit);
    t.mbr_bit = (uint16_t)(t.low_x_bit + t.low_y_bit + 2 * t.edge_bit);
    if (t.mbr_bit > 64) return false;
    const uint32_t temp = 2u * t.id_bit + t.duration_bit + t.end_bit + t.mbr_bit;   // eight_parallel
    if (temp > 128) return false;
    t.key_bit = (uint16_t)(temp / 8 * 8);
    if (t.key_bit < temp) t.key_bit += 8;
    const int ow = (int)((double)wx / (map[2] - map[0]) * 300);
    const int oh = (int)((double)wy / (map[3] - map[1]) * 420);
    int xg = std::max(ow + 1, 32), yg = std::max(oh + 1, 32);
    if (xg >= 256 || yg >= 256) {
        xg = std::min(ow + 1, 255);
        yg = std::min(oh + 1, 255);
    }
    t.x_grid = (uint16_t)xg;
    t.y_grid = (uint16_t)yg;
    t.bitmap_bytes = (uint32_t)(xg * yg + 7) / 8;
    return true;
}

}  // namespace

struct KvEngine {
    KvParams p{};
    hipStream_t st = nullptr;
    uint32_t end_time_min = 0;
    uint32_t nctf = 0;
    uint64_t key_cap = 0;
    std::vector<void *> allocs;
    KvLogEntry *log = nullptr;
    SortView a{}, b{};
    float4 *box_in = nullptr, *box_out = nullptr;
    uint32_t *cnt = nullptr, *sids = nullptr, *cap = nullptr;
    uint32_t *idx0 = nullptr, *idx1 = nullptr;   // row indices of the two-word sorts
    void *sort_tmp = nullptr;                    // the sort's temporary storage
    size_t sort_tmp_bytes = 0;
    uint4 *obox = nullptr;
    SortView oa{}, ob{};   // (mid, oid) pairs; hi unused
    glint_ctf *ctfs = nullptr;
    uint32_t *bitmaps = nullptr, *obitmap = nullptr;
    uint8_t *packed = nullptr;
    KvCtr *ctr = nullptr;
    KvCtr *h_ctr = nullptr;
    std::vector<glint_ctf> h_ctfs;
    glint_kv_summary sum{};
    bool flushed = false;
};

namespace {

template <typename T>
bool kalloc(KvEngine *e, T **p, uint64_t count, std::string &err)
{
    void *v = nullptr;
    if (count == 0) count = 1;
    if (hipMalloc(&v, count * sizeof(T)) != hipSuccess) {
        (void)hipGetLastError();
        err = "key-value output: hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed";
        return false;
    }
    e->allocs.push_back(v);
    *p = (T *)v;
    return true;
}

#define KV_TRY(call)                                                     \
    do {                                                                 \
        hipError_t e_ = (call);                                          \
        if (e_ != hipSuccess) {                                          \
            err = std::string(#call) + ": " + hipGetErrorString(e_);     \
            return GLINT_ERR_HIP;                                        \
        }                                                                \
    } while (0)

glint_status read_ctr(KvEngine *e, std::string &err)
{
    KV_TRY(hipMemcpyAsync(e->h_ctr, e->ctr, sizeof(KvCtr), hipMemcpyDeviceToHost, e->st));
    KV_TRY(hipStreamSynchronize(e->st));
    return GLINT_OK;
}

// The bits that differ across the n keys of v: [*_lo, *_hi) of the low and
// the high word (empty ranges: lo >= hi).
struct VaryingBits {
    uint32_t lo_begin, lo_end, hi_begin, hi_end;
};

glint_status varying_bits(KvEngine *e, const SortView &v, uint64_t n, bool wide, VaryingBits &vb,
                          std::string &err)
{
    vb = VaryingBits{0, 0, 0, 0};
    if (n < 2) return GLINT_OK;
    k_ctr_reset_masks<<<1, 1, 0, e->st>>>(e->ctr);
    k_andor<<<grid_of(n, 1024), kB, 0, e->st>>>(wide ? v.hi : nullptr, v.lo, n, e->ctr);
    KV_TRY(hipGetLastError());
    glint_status s = read_ctr(e, err);
    if (s) return s;
    const uint64_t dh = wide ? e->h_ctr->or_hi ^ e->h_ctr->and_hi : 0, dl = e->h_ctr->or_lo ^ e->h_ctr->and_lo;
    if (dl) vb.lo_begin = (uint32_t)__builtin_ctzll(dl), vb.lo_end = 64 - (uint32_t)__builtin_clzll(dl);
    if (dh) vb.hi_begin = (uint32_t)__builtin_ctzll(dh), vb.hi_end = 64 - (uint32_t)__builtin_clzll(dh);
    return GLINT_OK;
}

SortView offset(const SortView &v, uint64_t k)
{
    return SortView{v.hi ? v.hi + k : nullptr, v.lo + k, v.val + k};
}

// Stable sort of n (key, value) pairs by key bits [b0, b1): the pairs start
// in (k0, v0), with (k1, v1) as scratch; returns whether they ended in the
// scratch pair.
glint_status sort_pairs(KvEngine *e, uint64_t *k0, uint64_t *k1, uint32_t *v0, uint32_t *v1, uint64_t n, uint32_t b0,
                        uint32_t b1, bool &in_scratch, std::string &err)
{
    rocprim::double_buffer<uint64_t> keys(k0, k1);
    rocprim::double_buffer<uint32_t> vals(v0, v1);
    size_t bytes = e->sort_tmp_bytes;
    KV_TRY(rocprim::radix_sort_pairs(e->sort_tmp, bytes, keys, vals, (size_t)n, b0, b1, e->st));
    in_scratch = keys.current() != k0;
    return GLINT_OK;
}

// Stable sort of n items of `a` by their keys (the high word, if wide, then
// the low word), `b` as scratch; the result is left in `a`.  A wide key is
// sorted by its low word carrying the row index, the rows permuted, then by
// its high word carrying the row index again, and the rows permuted; only
// the bits that differ across the keys are sorted.
glint_status radix_sort(KvEngine *e, SortView a, SortView b, uint64_t n, bool wide, std::string &err)
{
    VaryingBits vb;
    glint_status s = varying_bits(e, a, n, wide, vb, err);
    if (s) return s;
    const uint32_t g = grid_of(n);
    bool scratch = false;
    if (!wide) {
        if (vb.lo_begin >= vb.lo_end) return GLINT_OK;
        s = sort_pairs(e, a.lo, b.lo, a.val, b.val, n, vb.lo_begin, vb.lo_end, scratch, err);
        if (s) return s;
        if (scratch) {
            KV_TRY(hipMemcpyAsync(a.lo, b.lo, n * 8, hipMemcpyDeviceToDevice, e->st));
            KV_TRY(hipMemcpyAsync(a.val, b.val, n * 4, hipMemcpyDeviceToDevice, e->st));
        }
        return GLINT_OK;
    }
    // rows (hi, lo, val) live in a; b and the two index buffers are scratch
    if (vb.lo_begin < vb.lo_end) {
        k_iota<<<g, kB, 0, e->st>>>(e->idx0, n);
        s = sort_pairs(e, a.lo, b.lo, e->idx0, e->idx1, n, vb.lo_begin, vb.lo_end, scratch, err);
        if (s) return s;
        const uint32_t *perm = scratch ? e->idx1 : e->idx0;
        if (scratch) KV_TRY(hipMemcpyAsync(a.lo, b.lo, n * 8, hipMemcpyDeviceToDevice, e->st));
        k_permute<<<g, kB, 0, e->st>>>(perm, a.hi, b.hi, a.val, b.val, n);
        KV_TRY(hipMemcpyAsync(a.hi, b.hi, n * 8, hipMemcpyDeviceToDevice, e->st));
        KV_TRY(hipMemcpyAsync(a.val, b.val, n * 4, hipMemcpyDeviceToDevice, e->st));
    }
    if (vb.hi_begin < vb.hi_end) {
        k_iota<<<g, kB, 0, e->st>>>(e->idx0, n);
        s = sort_pairs(e, a.hi, b.hi, e->idx0, e->idx1, n, vb.hi_begin, vb.hi_end, scratch, err);
        if (s) return s;
        const uint32_t *perm = scratch ? e->idx1 : e->idx0;
        if (scratch) KV_TRY(hipMemcpyAsync(a.hi, b.hi, n * 8, hipMemcpyDeviceToDevice, e->st));
        k_permute<<<g, kB, 0, e->st>>>(perm, a.lo, b.lo, a.val, b.val, n);
        KV_TRY(hipMemcpyAsync(a.lo, b.lo, n * 8, hipMemcpyDeviceToDevice, e->st));
        KV_TRY(hipMemcpyAsync(a.val, b.val, n * 4, hipMemcpyDeviceToDevice, e->st));
    }
    KV_TRY(hipGetLastError());
    return GLINT_OK;
}

}  // namespace

glint_status kv_create(const KvParams &p, hipStream_t stream, KvEngine **out, std::string &err)
{
    *out = nullptr;
    if (p.split_num == 0 || p.split_num * p.split_num + 2 > 4096 || p.max_meet_time == 0 ||
        p.num_objects > (1u << 26) || p.key_capacity < 2 || p.key_capacity >= (1ull << 32) ||
        !(p.area_limit == p.area_limit)) {
        err = "invalid key-value output configuration (split_num^2 + 2 <= 4096, objects <= 2^26, "
              "2 <= kv_capacity < 2^32, max_meet_time >= 1)";
        return GLINT_ERR_INVALID_ARGUMENT;
    }
    if (!(p.map[0] >= -180.0) || !(p.map[2] < 42000.0) || !(p.map[1] >= -180.0) || !(p.map[3] < 42000.0)) {
        err = "key-value output: coordinates must lie in [-180, 42000) for float_to_uint";
        return GLINT_ERR_INVALID_ARGUMENT;
    }
    KvEngine *e = new (std::nothrow) KvEngine();
    if (!e) return GLINT_ERR_OUT_OF_MEMORY;
    e->p = p;
    e->st = stream;
    e->end_time_min = p.first_end_time_min;
    e->nctf = p.split_num * p.split_num;
    e->key_cap = p.key_capacity & ~1ull;
    const uint64_t K = e->key_cap, N = p.num_objects;
    bool ok = kalloc(e, &e->log, K / 2, err) && kalloc(e, &e->a.hi, K, err) && kalloc(e, &e->a.lo, K, err) &&
              kalloc(e, &e->a.val, K, err) && kalloc(e, &e->b.hi, K, err) && kalloc(e, &e->b.lo, K, err) &&
              kalloc(e, &e->b.val, K, err) && kalloc(e, &e->box_in, K, err) && kalloc(e, &e->box_out, K, err) &&
              kalloc(e, &e->cnt, N, err) && kalloc(e, &e->sids, N, err) && kalloc(e, &e->obox, N, err) &&
              kalloc(e, &e->oa.lo, N, err) && kalloc(e, &e->oa.val, N, err) && kalloc(e, &e->ob.lo, N, err) &&
              kalloc(e, &e->ob.val, N, err) && kalloc(e, &e->cap, e->nctf, err) &&
              kalloc(e, &e->idx0, K, err) && kalloc(e, &e->idx1, K, err) && kalloc(e, &e->ctfs, e->nctf, err) &&
              kalloc(e, &e->bitmaps, (uint64_t)e->nctf * (kBitmapBytes / 4), err) &&
              kalloc(e, &e->obitmap, kBitmapBytes / 4, err) && kalloc(e, &e->packed, 16 * K + 16, err) &&
              kalloc(e, &e->ctr, 1, err);
    if (ok && hipHostMalloc((void **)&e->h_ctr, sizeof(KvCtr), hipHostMallocDefault) != hipSuccess) {
        err = "key-value output: hipHostMalloc failed";
        ok = false;
    }
    if (!ok) {
        kv_destroy(e);
        return GLINT_ERR_OUT_OF_MEMORY;
    }
    {
        // temporary storage of the largest sort (every bit of a 64-bit key)
        rocprim::double_buffer<uint64_t> keys(e->a.lo, e->b.lo);
        rocprim::double_buffer<uint32_t> vals(e->idx0, e->idx1);
        size_t bytes = 0;
        const uint64_t most = std::max<uint64_t>(K, N);
        if (rocprim::radix_sort_pairs(nullptr, bytes, keys, vals, (size_t)most, 0, 64, stream) != hipSuccess ||
            !kalloc(e, reinterpret_cast<uint8_t **>(&e->sort_tmp), bytes, err)) {
            if (err.empty()) err = "key-value output: sort storage query failed";
            kv_destroy(e);
            return GLINT_ERR_OUT_OF_MEMORY;
        }
        e->sort_tmp_bytes = bytes;
    }
    std::memset(e->h_ctr, 0, sizeof(KvCtr));
    *out = e;
    return GLINT_OK;
}

void kv_destroy(KvEngine *e)
{
    if (!e) return;
    if (e->st) (void)hipStreamSynchronize(e->st);
    for (void *v : e->allocs) (void)hipFree(v);
    if (e->h_ctr) (void)hipHostFree(e->h_ctr);
    delete e;
}

KvLogEntry *kv_log(KvEngine *e) { return e->log; }
uint32_t kv_end_time_min(const KvEngine *e) { return e->end_time_min; }
void kv_set_end_time_min(KvEngine *e, uint32_t t) { e->end_time_min = t; }
uint64_t kv_log_capacity(const KvEngine *e) { return e->key_cap / 2; }

glint_status kv_flush(KvEngine *e, uint64_t nlog, uint32_t cur_time, glint_kv_summary *out, std::string &err)
{
    const uint32_t N = e->p.num_objects, split = e->p.split_num, C = e->nctf;
    nlog = std::min(nlog, kv_log_capacity(e));
    const uint64_t nkeys = 2 * nlog;
    hipStream_t st = e->st;
    const uint32_t gN = grid_of(N), gK = grid_of(nkeys);

    k_ctr_reset<<<1, 1, 0, st>>>(e->ctr);
    k_obj_init<<<gN, kB, 0, st>>>(N, e->cnt, e->obox, e->sids);
    if (nlog)
        k_expand<<<grid_of(nlog), kB, 0, st>>>(e->log, nlog, e->a, e->box_in, e->cnt, e->obox, e->sids, e->ctr,
                                               e->end_time_min, e->p.max_meet_time);
    k_objects<<<gN, kB, 0, st>>>(N, e->cnt, e->obox, e->sids, e->oa.lo, e->oa.val, e->p.area_limit, e->ctr);
    if (nkeys) k_partition<<<gK, kB, 0, st>>>(e->a, e->b, nkeys, e->sids, e->ctr);
    KV_TRY(hipGetLastError());
    glint_status s = read_ctr(e, err);
    if (s) return s;
    const uint64_t nk = e->h_ctr->nk, no = e->h_ctr->no;
    const uint32_t sid_count = e->h_ctr->sid_count;
    const uint32_t ovs = e->h_ctr->ovs_meet + e->h_ctr->ovs_area;

    // oversize buffer: sorted keys
    s = radix_sort(e, offset(e->b, nk), offset(e->a, nk), no, true, err);
    if (s) return s;

    // STR (geometry.cu:1832-1895)
    s = radix_sort(e, e->oa, e->ob, N, false, err);
    if (s) return s;
    const uint32_t z = (N - sid_count) + ovs;
    const uint32_t xcap = (sid_count - ovs) / split + 1;
    const uint32_t part = xcap / split + 1;
    if (z < N) {
        k_narrow<<<gN, kB, 0, st>>>(e->oa.lo, N, z, xcap, split);
        s = radix_sort(e, e->oa, e->ob, N, false, err);
        if (s) return s;
        k_make_sid<<<gN, kB, 0, st>>>(e->oa.val, N, z, xcap, part, split, e->sids);
    }

    // kept keys: sid written, sorted; boxes follow their keys
    if (nk) k_write_sid<<<grid_of(nk), kB, 0, st>>>(e->b.hi, nk, e->sids);
    s = radix_sort(e, e->b, e->a, nk, true, err);
    if (s) return s;
    if (nkeys) k_gather_boxes<<<gK, kB, 0, st>>>(e->b.val, e->box_in, e->box_out, nkeys);

    // files: capacities, limits; widths on the host
    k_fill_u32<<<grid_of(C), kB, 0, st>>>(e->cap, C, 0);
    if (z < N) k_capacity<<<grid_of(N - z, 1024), kB, 0, st>>>(e->oa.val, N, z, e->sids, e->cnt, C, e->cap);
    k_limits<<<C, kB, 0, st>>>(e->cap, C, e->b.hi, e->b.lo, e->box_out, nk, e->end_time_min, cur_time, e->ctfs);
    KV_TRY(hipGetLastError());
    e->h_ctfs.assign(C, glint_ctf{});
    KV_TRY(hipMemcpyAsync(e->h_ctfs.data(), e->ctfs, C * sizeof(glint_ctf), hipMemcpyDeviceToHost, st));
    KV_TRY(hipStreamSynchronize(st));
    uint64_t bytes = 0;
    bool fits = true;
    for (uint32_t t = 0; t < C; t++) {
        glint_ctf &f = e->h_ctfs[t];
        f.packed_offset = bytes;
        const bool has_keys = f.key_offset < nk && f.kv_count > 0;
        if (has_keys && !ctf_bits(f, e->p.map, N, e->p.max_meet_time)) fits = false;
        bytes += (uint64_t)f.kv_count * (f.key_bit / 8u);
    }
    if (!fits) {
        err = "key-value output: a file's key does not fit 128 bits";
        return GLINT_ERR_INVALID_ARGUMENT;
    }
    if (bytes > 16 * e->key_cap) {
        err = "key-value output: packed keys exceed their buffer";
        return GLINT_ERR_OUTPUT_CAPACITY;
    }
    KV_TRY(hipMemcpyAsync(e->ctfs, e->h_ctfs.data(), C * sizeof(glint_ctf), hipMemcpyHostToDevice, st));
    // bytes of keys past the kept keys (clipped ranges) stay zero
    k_fill_u32<<<grid_of((bytes + 3) / 4), kB, 0, st>>>((uint32_t *)e->packed, (bytes + 3) / 4, 0);
    k_fill_u32<<<grid_of((uint64_t)C * (kBitmapBytes / 4)), kB, 0, st>>>(e->bitmaps, (uint64_t)C * (kBitmapBytes / 4), 0);
    {
        const uint32_t slices = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(64, nk / ((uint64_t)C * 4096) + 1));
        k_bitmaps<<<dim3(C, slices), kB, 0, st>>>(e->ctfs, e->box_out, nk, e->bitmaps);
    }
    if (nk) k_pack<<<grid_of(nk), kB, 0, st>>>(e->ctfs, C, e->b.hi, e->b.lo, e->box_out, nk, e->packed);

    // oversize buffer: start times and map bitmap
    k_fill_u32<<<8, kB, 0, st>>>(e->obitmap, kBitmapBytes / 4, 0);
    if (no)
        k_oversize<<<grid_of(no, 1024), kB, 0, st>>>(e->b.lo, e->box_out, nk, nk + no, e->p.map[0], e->p.map[1],
                                                     e->p.map[2], e->p.map[3], e->end_time_min, e->obitmap, e->ctr);
    KV_TRY(hipGetLastError());
    s = read_ctr(e, err);
    if (s) return s;

    glint_kv_summary &m = e->sum;
    m = glint_kv_summary{};
    m.kept_keys = nk;
    m.oversize_keys = no;
    m.packed_bytes = bytes;
    m.logged_meetings = nlog;
    m.ctf_count = C;
    m.sid_count = sid_count;
    m.oversize_oid_count = ovs;
    m.zero_one_count = z;
    m.end_time_min = e->end_time_min;
    m.end_time_max = cur_time;
    m.o_start_time_min = e->h_ctr->osmin;
    m.o_start_time_max = e->h_ctr->osmax;
    e->flushed = true;
    // the next run's end offsets count from this tick (process.cpp:549,271)
    e->end_time_min = cur_time;
    if (out) *out = m;
    return GLINT_OK;
}

uint64_t kv_part_bytes(const KvEngine *e, uint32_t part)
{
    if (!e->flushed) return 0;
    const glint_kv_summary &m = e->sum;
    switch (part) {
    case GLINT_KV_KEYS: return 16 * m.kept_keys;
    case GLINT_KV_BOXES: return 16 * m.kept_keys;
    case GLINT_KV_OVERSIZE_KEYS: return 16 * m.oversize_keys;
    case GLINT_KV_OVERSIZE_BOXES: return 16 * m.oversize_keys;
    case GLINT_KV_SIDS: return 2ull * e->p.num_objects;
    case GLINT_KV_CTFS: return (uint64_t)e->nctf * sizeof(glint_ctf);
    case GLINT_KV_BITMAPS: return (uint64_t)e->nctf * kBitmapBytes;
    case GLINT_KV_OVERSIZE_BITMAP: return kBitmapBytes;
    case GLINT_KV_PACKED: return m.packed_bytes;
    default: return 0;
    }
}

glint_status kv_copy(KvEngine *e, uint32_t part, void *dst, uint64_t bytes, std::string &err)
{
    if (!e->flushed) {
        err = "no key-value flush yet";
        return GLINT_ERR_STATE;
    }
    if (part > GLINT_KV_PACKED || bytes != kv_part_bytes(e, part) || (bytes && !dst)) {
        err = "key-value copy: unknown part or wrong size";
        return GLINT_ERR_INVALID_ARGUMENT;
    }
    if (!bytes) return GLINT_OK;
    const uint64_t nk = e->sum.kept_keys, no = e->sum.oversize_keys;
    KV_TRY(hipStreamSynchronize(e->st));
    auto keys = [&](uint64_t first, uint64_t n) -> glint_status {
        std::vector<uint64_t> h(n), l(n);
        KV_TRY(hipMemcpy(h.data(), e->b.hi + first, n * 8, hipMemcpyDeviceToHost));
        KV_TRY(hipMemcpy(l.data(), e->b.lo + first, n * 8, hipMemcpyDeviceToHost));
        uint64_t *o = (uint64_t *)dst;
        for (uint64_t i = 0; i < n; i++) {
            o[2 * i] = l[i];
            o[2 * i + 1] = h[i];
        }
        return GLINT_OK;
    };
    switch (part) {
    case GLINT_KV_KEYS: return keys(0, nk);
    case GLINT_KV_OVERSIZE_KEYS: return keys(nk, no);
    case GLINT_KV_BOXES: KV_TRY(hipMemcpy(dst, e->box_out, bytes, hipMemcpyDeviceToHost)); return GLINT_OK;
    case GLINT_KV_OVERSIZE_BOXES:
        KV_TRY(hipMemcpy(dst, e->box_out + nk, bytes, hipMemcpyDeviceToHost));
        return GLINT_OK;
    case GLINT_KV_SIDS: {
        std::vector<uint32_t> v(e->p.num_objects);
        KV_TRY(hipMemcpy(v.data(), e->sids, v.size() * 4, hipMemcpyDeviceToHost));
        uint16_t *o = (uint16_t *)dst;
        for (size_t i = 0; i < v.size(); i++) o[i] = (uint16_t)v[i];
        return GLINT_OK;
    }
    case GLINT_KV_CTFS: std::memcpy(dst, e->h_ctfs.data(), bytes); return GLINT_OK;
    case GLINT_KV_BITMAPS: KV_TRY(hipMemcpy(dst, e->bitmaps, bytes, hipMemcpyDeviceToHost)); return GLINT_OK;
    case GLINT_KV_OVERSIZE_BITMAP: KV_TRY(hipMemcpy(dst, e->obitmap, bytes, hipMemcpyDeviceToHost)); return GLINT_OK;
    case GLINT_KV_PACKED: KV_TRY(hipMemcpy(dst, e->packed, bytes, hipMemcpyDeviceToHost)); return GLINT_OK;
    default: return GLINT_ERR_INVALID_ARGUMENT;
    }
}
