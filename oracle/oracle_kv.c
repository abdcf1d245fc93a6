/*
 * oracle_kv.c - CPU restatement of the sorted key-value meeting output (the
 * flush branch of the reference's GPU path).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Compiled with -ffp-contract=off.
 *
 * Reference lines each part follows (paths relative to the reference repo):
 *  - meeting box: f_box grown by both points at every contact tick, starting
 *    from (1e5, 1e5, -1e5, -1e5)            src/cuda/geometry.cu:469-494,82-97
 *  - emission of an ended meeting: dropped unless its box lies in the map
 *    box; duration > max_meet_time marks both objects oversize (sid 1) and
 *    counts one; two directed keys
 *      [sid:12 | oid:26 | target:26 | duration:32 | end - end_time_min:32]
 *    with the meeting box; per-object key count and a uint box through
 *    float_to_uint(v) = (uint)((v + 180) * 100000)
 *                                           src/cuda/geometry.cu:665-735,115-121
 *                                           src/cuda/cuda_util.cuh:115-125
 *                                           src/util/util.h:47-52
 *  - object pass: area of the object's box (back through uint_to_float)
 *    above the area limit -> sid 1; else mid = (low0/2 + high0/2) << 32 |
 *    (high1/2 + high1/2)                    src/cuda/geometry.cu:740-766
 *  - keys of sid-1 objects go to the oversize buffer, sorted; its start
 *    times and a 256x256 map bitmap         src/cuda/geometry.cu:1766-1830,895-979
 *                                           src/tracing/step_merge.cpp:41-64
 *  - STR: objects sorted by mid; from position z = (N - sid_count) +
 *    oversize_count on, split into split_num slabs of x_cap positions, each
 *    sorted by the low word; sid = slab * split_num + part + 2
 *                                           src/cuda/geometry.cu:1832-1895,768-796
 *  - sid written into the kept keys, keys sorted, per-CTF capacities,
 *    limits (box and start times)           src/cuda/geometry.cu:798-890,1897-1914
 *  - per-CTF widths and grids               src/tracing/step_merge.cpp:140-186
 *  - per-CTF bitmaps and variable-width keys src/cuda/geometry.cu:992-1019,1076-1107
 *                                           src/tracing/step_merge.cpp:284-310
 * Sorting: the reference sorts with Thrust (radix sorts, stable in practice)
 * and partitions unstably; here every sort is stable and the partition keeps
 * the emission order; the final key order is the total order of unique keys.
 * Where the reference's result is undefined the restatement defines it:
 *  - a zero denominator in a grid or box quantisation gives 0;
 *  - bitmap cells are clipped to the grid (the CPU loops'
 *    i < x_grid, j < y_grid, step_merge.cpp:57-58,470-471);
 *  - an STR slab or CTF key range past the end is clipped (capacities can
 *    exceed the kept keys when duration-oversize objects receive a sid:
 *    the reference counts them once per meeting, geometry.cu:708-712);
 *  - a CTF with no keys in its range gets zero widths and grids; offsets
 *    still advance by its capacity, as the reference's do.
 */
#include "oracle.h"
#include "oracle_kv_internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

enum { kBitmapEdge = 256, kBitmapBytes = kBitmapEdge * kBitmapEdge / 8 };

typedef struct kvrec {
    u128 key;
    float box[4];
} kvrec;

struct okv {
    oracle_kv_config cfg;
    uint32_t n;
    double map[4];
    uint32_t end_time_min;
    uint64_t logged;
    /* accumulated since the last flush */
    kvrec *keys;
    uint64_t nkeys, capkeys;
    uint32_t *cnt;          /* same_pid_count */
    uint32_t *obox;         /* 4 per object: low0, low1, high0, high1 */
    uint16_t *sids;
    uint32_t sid_count, oversize;
    /* last flush */
    oracle_kv_summary sum;
    uint64_t *out_keys[2];  /* (lo, hi) per key */
    float *out_boxes[2];
    uint16_t *out_sids;
    oracle_ctf *ctfs;
    uint8_t *bitmaps, *obitmap, *packed;
};

static uint32_t f2u(float v)
{
    v += 180;
    return (uint32_t)(v * 100000);
}

static float u2f(uint32_t f) { return (float)f / 100000 - 180; }

static void reset_objects(okv *k)
{
    for (uint32_t i = 0; i < k->n; i++) {
        k->cnt[i] = 0;
        k->sids[i] = 0;
        k->obox[4 * i] = 100000000u;      /* geometry.cu:1254-1263 */
        k->obox[4 * i + 1] = 100000000u;
        k->obox[4 * i + 2] = 0;
        k->obox[4 * i + 3] = 0;
    }
    k->nkeys = 0;
    k->sid_count = 0;
    k->oversize = 0;
    k->logged = 0;
}

okv *okv_create(uint32_t n, const double map[4], uint32_t end_time_min, const oracle_kv_config *kc)
{
    okv *k = (okv *)calloc(1, sizeof(okv));
    k->cfg = *kc;
    k->n = n;
    memcpy(k->map, map, sizeof(k->map));
    k->end_time_min = end_time_min;
    k->cnt = (uint32_t *)malloc((size_t)n * sizeof(uint32_t));
    k->obox = (uint32_t *)malloc((size_t)n * 4 * sizeof(uint32_t));
    k->sids = (uint16_t *)malloc((size_t)n * sizeof(uint16_t));
    reset_objects(k);
    return k;
}

static void free_outputs(okv *k)
{
    for (int p = 0; p < 2; p++) {
        free(k->out_keys[p]);
        free(k->out_boxes[p]);
        k->out_keys[p] = NULL;
        k->out_boxes[p] = NULL;
    }
    free(k->out_sids);
    free(k->ctfs);
    free(k->bitmaps);
    free(k->obitmap);
    free(k->packed);
    k->out_sids = NULL;
    k->ctfs = NULL;
    k->bitmaps = k->obitmap = k->packed = NULL;
}

void okv_destroy(okv *k)
{
    if (!k) return;
    free_outputs(k);
    free(k->keys);
    free(k->cnt);
    free(k->obox);
    free(k->sids);
    free(k);
}

uint64_t okv_pending(const okv *k) { return k->nkeys; }

void okv_emit(okv *k, uint32_t a, uint32_t b, uint32_t start, uint32_t end, const float box[4])
{
    /* geometry.cu:667-670, box_contain (geometry.cu:53-58) */
    if (!(box[0] >= k->map[0] && box[2] <= k->map[2] && box[1] >= k->map[1] && box[3] <= k->map[3])) return;
    k->logged++;
    const uint32_t duration = end - start;
    if (duration > k->cfg.max_meet_time) {     /* geometry.cu:708-712 */
        k->oversize++;
        k->sids[a] = 1;
        k->sids[b] = 1;
    }
    for (int dir = 0; dir < 2; dir++) {
        const uint32_t pid = dir ? b : a, target = dir ? a : b;
        if (k->nkeys == k->capkeys) {
            k->capkeys = k->capkeys ? 2 * k->capkeys : 4096;
            k->keys = (kvrec *)realloc(k->keys, k->capkeys * sizeof(kvrec));
        }
        kvrec *r = &k->keys[k->nkeys++];
        r->key = ((u128)pid << 90) + ((u128)target << 64) + ((u128)duration << 32) + (u128)(uint32_t)(end - k->end_time_min);
        memcpy(r->box, box, sizeof(r->box));
        if (k->cnt[pid]++ == 0) k->sid_count++;
        uint32_t *o = &k->obox[4 * (size_t)pid];
        uint32_t v;
        v = f2u(box[0]); if (v < o[0]) o[0] = v;
        v = f2u(box[2]); if (v > o[2]) o[2] = v;
        v = f2u(box[1]); if (v < o[1]) o[1] = v;
        v = f2u(box[3]); if (v > o[3]) o[3] = v;
    }
}

/* ------------------------------------------------------------ sorting */

static int cmp_rec(const void *x, const void *y)
{
    const u128 a = ((const kvrec *)x)->key, b = ((const kvrec *)y)->key;
    return a < b ? -1 : (a > b ? 1 : 0);
}

typedef struct objrec {
    uint64_t mid;
    uint32_t oid;
    uint32_t pos;   /* position before the sort: ties keep it (stable) */
} objrec;

static int cmp_obj(const void *x, const void *y)
{
    const objrec *a = (const objrec *)x, *b = (const objrec *)y;
    if (a->mid != b->mid) return a->mid < b->mid ? -1 : 1;
    return a->pos < b->pos ? -1 : (a->pos > b->pos ? 1 : 0);
}

static void stable_sort_objs(objrec *o, uint64_t n)
{
    for (uint64_t i = 0; i < n; i++) o[i].pos = (uint32_t)i;
    qsort(o, n, sizeof(objrec), cmp_obj);
}

/* ------------------------------------------------------- CTF encoding */

static uint16_t min_bits(uint64_t v)
{
    /* min_bits_to_store (util.h:528-533) */
    if (v == 0) return 1;
    return (uint16_t)((uint32_t)log2((double)v) + 1);
}

static int ctf_bits(oracle_ctf *t, const double map[4], uint32_t num_objects, uint32_t max_meet_time)
{
    /* CTF::get_ctf_bits (step_merge.cpp:145-186) */
    t->id_bit = min_bits(num_objects);
    t->duration_bit = min_bits((uint32_t)(max_meet_time - 1));
    t->end_bit = t->duration_bit;
    const double gx = (map[2] - map[0]) / 3000;
    const double gy = (map[3] - map[1]) / 4200;
    const uint32_t lxg = (uint32_t)((double)(t->mbr[2] - t->mbr[0]) / gx);
    const uint32_t lyg = (uint32_t)((double)(t->mbr[3] - t->mbr[1]) / gy);
    t->low_x_bit = min_bits((lxg > 1 ? lxg : 1) - 1);
    t->low_y_bit = min_bits((lyg > 1 ? lyg : 1) - 1);
    t->edge_bit = t->low_x_bit > t->low_y_bit ? t->low_x_bit : t->low_y_bit;
    t->mbr_bit = (uint16_t)(t->low_x_bit + t->low_y_bit + 2 * t->edge_bit);
    if (t->mbr_bit > 64) return -1;
    const uint32_t temp = 2u * t->id_bit + t->duration_bit + t->end_bit + t->mbr_bit;   /* eight_parallel */
    if (temp > 128) return -1;
    t->key_bit = (uint16_t)(temp / 8 * 8);
    if (t->key_bit < temp) t->key_bit += 8;
    const int ow = (int)((double)(t->mbr[2] - t->mbr[0]) / (map[2] - map[0]) * 300);
    const int oh = (int)((double)(t->mbr[3] - t->mbr[1]) / (map[3] - map[1]) * 420);
    int xg = ow + 1 > 32 ? ow + 1 : 32, yg = oh + 1 > 32 ? oh + 1 : 32;
    if (xg >= 256 || yg >= 256) {
        xg = ow + 1 < 255 ? ow + 1 : 255;
        yg = oh + 1 < 255 ? oh + 1 : 255;
    }
    t->x_grid = (uint16_t)xg;
    t->y_grid = (uint16_t)yg;
    t->bitmap_bytes = (uint32_t)(xg * yg + 7) / 8;
    return 0;
}

static float ratio(float num, float den) { return den == 0 ? 0.0f : num / den; }

static uint64_t serial_mbr(const float *b, const oracle_ctf *t)
{
    /* serialize_mbr (geometry.cu:162-172) */
    const float w = t->mbr[2] - t->mbr[0], h = t->mbr[3] - t->mbr[1];
    const float longer = w > h ? w : h;
    const uint64_t low0 = (uint64_t)(ratio(b[0] - t->mbr[0], w) * (float)((1ull << t->low_x_bit) - 1));
    const uint64_t low1 = (uint64_t)(ratio(b[1] - t->mbr[1], h) * (float)((1ull << t->low_y_bit) - 1));
    const uint64_t x = (uint64_t)(ratio(b[2] - b[0], longer) * (float)((1ull << t->edge_bit) - 1));
    const uint64_t y = (uint64_t)(ratio(b[3] - b[1], longer) * (float)((1ull << t->edge_bit) - 1));
    return (low0 << (t->low_y_bit + 2 * t->edge_bit)) + (low1 << (2 * t->edge_bit)) + (x << t->edge_bit) + y;
}

static void set_bit(uint8_t *bm, uint32_t pos) { bm[pos / 8] |= (uint8_t)(1u << (pos % 8)); }

/* ---------------------------------------------------------------- flush */

int okv_flush(okv *k, uint32_t cur_time)
{
    const uint32_t N = k->n, split = k->cfg.split_num, C = split * split;
    free_outputs(k);
    memset(&k->sum, 0, sizeof(k->sum));

    /* set_oid (geometry.cu:740-766) */
    objrec *obj = (objrec *)malloc((size_t)(N ? N : 1) * sizeof(objrec));
    uint32_t ovs = k->oversize;
    for (uint32_t id = 0; id < N; id++) {
        obj[id].oid = id;
        obj[id].mid = 0;
        if (k->cnt[id] >= 1) {
            const uint32_t *o = &k->obox[4 * (size_t)id];
            const float l0 = u2f(o[0]), h0 = u2f(o[2]), l1 = u2f(o[1]), h1 = u2f(o[3]);
            const float area = (h1 - l1) * (h0 - l0);
            if ((double)area > k->cfg.area_limit) {
                ovs++;
                k->sids[id] = 1;
                continue;
            }
            const uint32_t m0 = o[0] / 2 + o[2] / 2;
            const uint32_t m1 = o[3] / 2 + o[3] / 2;   /* high1 twice, geometry.cu:764 */
            obj[id].mid = ((uint64_t)m0 << 32) + m1;
        }
    }

    /* partition: keys of sid-1 objects to the oversize buffer (geometry.cu:1776-1797) */
    kvrec *kept = (kvrec *)malloc((k->nkeys ? k->nkeys : 1) * sizeof(kvrec));
    kvrec *over = (kvrec *)malloc((k->nkeys ? k->nkeys : 1) * sizeof(kvrec));
    uint64_t nk = 0, no = 0;
    for (uint64_t i = 0; i < k->nkeys; i++) {
        const uint32_t oid = (uint32_t)(k->keys[i].key >> 90) & ((1u << 26) - 1);
        if (k->sids[oid] != 1)
            kept[nk++] = k->keys[i];
        else
            over[no++] = k->keys[i];
    }

    /* oversize buffer: sorted keys, start times, map bitmap */
    qsort(over, no, sizeof(kvrec), cmp_rec);
    uint32_t osmin = 1u << 30, osmax = 0;
    k->obitmap = (uint8_t *)calloc(kBitmapBytes, 1);
    for (uint64_t i = 0; i < no; i++) {
        const uint32_t dur = (uint32_t)(over[i].key >> 32);
        const uint32_t end = (uint32_t)over[i].key + k->end_time_min;
        const uint32_t st = end - dur;
        if (st < osmin) osmin = st;
        if (st > osmax) osmax = st;
        const float *b = over[i].box;
        const uint32_t l0 = (uint32_t)((b[0] - k->map[0]) / (k->map[2] - k->map[0]) * kBitmapEdge);
        const uint32_t l1 = (uint32_t)((b[1] - k->map[1]) / (k->map[3] - k->map[1]) * kBitmapEdge);
        const uint32_t h0 = (uint32_t)((b[2] - k->map[0]) / (k->map[2] - k->map[0]) * kBitmapEdge);
        const uint32_t h1 = (uint32_t)((b[3] - k->map[1]) / (k->map[3] - k->map[1]) * kBitmapEdge);
        for (uint32_t x = l0; x <= h0 && x < kBitmapEdge; x++)
            for (uint32_t y = l1; y <= h1 && y < kBitmapEdge; y++) set_bit(k->obitmap, x + y * kBitmapEdge);
    }

    /* STR over the object mids (geometry.cu:1832-1880) */
    stable_sort_objs(obj, N);
    const uint32_t z = (N - k->sid_count) + ovs;
    for (uint64_t p = z; p < N; p++) obj[p].mid &= 0xFFFFFFFFull;
    const uint32_t xcap = (k->sid_count - ovs) / split + 1;
    for (uint32_t s = 0; s < split; s++) {
        uint64_t lo = (uint64_t)z + (uint64_t)xcap * s;
        uint64_t hi = s + 1 < split ? lo + xcap : N;
        if (lo > N) lo = N;
        if (hi > N) hi = N;
        if (hi > lo) stable_sort_objs(obj + lo, hi - lo);
    }
    /* make_sid (geometry.cu:780-796) */
    const uint32_t part = xcap / split + 1;
    for (uint64_t p = z; p < N; p++) {
        const uint32_t xi = (uint32_t)(p - z) / xcap;
        const uint32_t yi = ((uint32_t)(p - z) - xi * xcap) / part;
        k->sids[obj[p].oid] = (uint16_t)(xi * split + yi + 2);
    }

    /* write_sid_in_key, sort (geometry.cu:798-806,1897) */
    const u128 low116 = ((u128)1 << 116) - 1;
    for (uint64_t i = 0; i < nk; i++) {
        const uint32_t oid = (uint32_t)(kept[i].key >> 90) & ((1u << 26) - 1);
        kept[i].key = (kept[i].key & low116) + ((u128)k->sids[oid] << 116);
    }
    qsort(kept, nk, sizeof(kvrec), cmp_rec);

    /* capacities (geometry.cu:808-818) */
    uint32_t *cap = (uint32_t *)calloc(C ? C : 1, sizeof(uint32_t));
    for (uint64_t p = z; p < N; p++) cap[k->sids[obj[p].oid] - 2] += k->cnt[obj[p].oid];

    /* limits and widths (cuda_get_limit, geometry.cu:820-890; get_ctf_bits) */
    k->ctfs = (oracle_ctf *)calloc(C ? C : 1, sizeof(oracle_ctf));
    k->bitmaps = (uint8_t *)calloc((size_t)(C ? C : 1) * kBitmapBytes, 1);
    uint64_t kidx = 0, bidx = 0;
    int bad = 0;
    for (uint32_t t = 0; t < C; t++) {
        oracle_ctf *f = &k->ctfs[t];
        f->kv_count = cap[t];
        f->key_offset = kidx;
        f->packed_offset = bidx;
        const uint64_t b0 = kidx < nk ? kidx : nk, b1 = kidx + cap[t] < nk ? kidx + cap[t] : nk;
        float lo0 = 100000.0f, lo1 = 100000.0f, hi0 = -100000.0f, hi1 = -100000.0f;
        uint32_t smin = 1u << 30, smax = 0;
        for (uint64_t i = b0; i < b1; i++) {
            const float *b = kept[i].box;
            if (b[0] < lo0) lo0 = b[0];
            if (b[1] < lo1) lo1 = b[1];
            if (b[2] > hi0) hi0 = b[2];
            if (b[3] > hi1) hi1 = b[3];
            const uint32_t dur = (uint32_t)(kept[i].key >> 32);
            const uint32_t end = (uint32_t)kept[i].key + k->end_time_min;
            const uint32_t st = end - dur;
            if (st < smin) smin = st;
            if (st > smax) smax = st;
        }
        f->mbr[0] = lo0;
        f->mbr[1] = lo1;
        f->mbr[2] = hi0;
        f->mbr[3] = hi1;
        f->start_time_min = smin;
        f->start_time_max = smax;
        f->end_time_min = k->end_time_min;
        f->end_time_max = cur_time;
        if (b1 > b0 && ctf_bits(f, k->map, N, k->cfg.max_meet_time) != 0) bad = 1;
        kidx += cap[t];
        bidx += (uint64_t)cap[t] * (f->key_bit / 8u);
    }
    /* bitmaps and variable-width keys of the keys present in each range */
    k->packed = (uint8_t *)calloc(bidx + 1, 1);
    for (uint32_t t = 0; t < C && !bad; t++) {
        const oracle_ctf *f = &k->ctfs[t];
        const uint64_t b0 = f->key_offset < nk ? f->key_offset : nk;
        const uint64_t b1 = f->key_offset + f->kv_count < nk ? f->key_offset + f->kv_count : nk;
        uint8_t *bm = k->bitmaps + (size_t)t * kBitmapBytes;
        const uint32_t kb = f->key_bit / 8u;
        for (uint64_t i = b0; i < b1; i++) {
            const float *b = kept[i].box;
            /* cuda_write_bitmap (geometry.cu:1002-1016) */
            const float w = f->mbr[2] - f->mbr[0], h = f->mbr[3] - f->mbr[1];
            const uint32_t l0 = (uint32_t)(ratio(b[0] - f->mbr[0], w) * f->x_grid);
            const uint32_t l1 = (uint32_t)(ratio(b[1] - f->mbr[1], h) * f->y_grid);
            const uint32_t h0 = (uint32_t)(ratio(b[2] - f->mbr[0], w) * f->x_grid);
            const uint32_t h1 = (uint32_t)(ratio(b[3] - f->mbr[1], h) * f->y_grid);
            for (uint32_t x = l0; x <= h0 && x < f->x_grid; x++)
                for (uint32_t y = l1; y <= h1 && y < f->y_grid; y++) set_bit(bm, x + y * f->x_grid);
            /* write_key_and_mbr (geometry.cu:1089-1101), extract_fields (cuda_util.cuh:164-169) */
            const u128 key = kept[i].key;
            const u128 pid = key >> 90;   /* the sid bits stay above the oid */
            const u128 target = (key >> 64) & ((1u << 26) - 1);
            const u128 dur = (key >> 32) & 0xFFFFFFFFu;
            const u128 end = key & 0xFFFFFFFFu;
            const u128 vm = serial_mbr(b, f);
            const u128 packed = (pid << (f->id_bit + f->duration_bit + f->end_bit + f->mbr_bit)) +
                                (target << (f->duration_bit + f->end_bit + f->mbr_bit)) +
                                (dur << (f->end_bit + f->mbr_bit)) + (end << f->mbr_bit) + vm;
            memcpy(k->packed + f->packed_offset + (i - b0) * kb, &packed, kb);
        }
    }

    /* outputs */
    k->out_keys[0] = (uint64_t *)malloc((nk ? nk : 1) * 16);
    k->out_boxes[0] = (float *)malloc((nk ? nk : 1) * 16);
    for (uint64_t i = 0; i < nk; i++) {
        k->out_keys[0][2 * i] = (uint64_t)kept[i].key;
        k->out_keys[0][2 * i + 1] = (uint64_t)(kept[i].key >> 64);
        memcpy(k->out_boxes[0] + 4 * i, kept[i].box, 16);
    }
    k->out_keys[1] = (uint64_t *)malloc((no ? no : 1) * 16);
    k->out_boxes[1] = (float *)malloc((no ? no : 1) * 16);
    for (uint64_t i = 0; i < no; i++) {
        k->out_keys[1][2 * i] = (uint64_t)over[i].key;
        k->out_keys[1][2 * i + 1] = (uint64_t)(over[i].key >> 64);
        memcpy(k->out_boxes[1] + 4 * i, over[i].box, 16);
    }
    k->out_sids = (uint16_t *)malloc((size_t)(N ? N : 1) * sizeof(uint16_t));
    memcpy(k->out_sids, k->sids, (size_t)N * sizeof(uint16_t));

    k->sum.kept_keys = nk;
    k->sum.oversize_keys = no;
    k->sum.packed_bytes = bidx;
    k->sum.logged_meetings = k->logged;
    k->sum.ctf_count = C;
    k->sum.sid_count = k->sid_count;
    k->sum.oversize_oid_count = ovs;
    k->sum.zero_one_count = z;
    k->sum.end_time_min = k->end_time_min;
    k->sum.end_time_max = cur_time;
    k->sum.o_start_time_min = osmin;
    k->sum.o_start_time_max = osmax;

    free(obj);
    free(kept);
    free(over);
    free(cap);
    /* the next run starts empty; its end offsets count from this tick
     * (process.cpp:549,271) */
    reset_objects(k);
    k->end_time_min = cur_time;
    return bad ? -1 : 0;
}

const oracle_kv_summary *okv_summary(const okv *k) { return &k->sum; }
const uint64_t *okv_keys(const okv *k, int part) { return k->out_keys[part ? 1 : 0]; }
const float *okv_boxes(const okv *k, int part) { return k->out_boxes[part ? 1 : 0]; }
const uint16_t *okv_sids(const okv *k) { return k->out_sids; }
const oracle_ctf *okv_ctfs(const okv *k) { return k->ctfs; }
const uint8_t *okv_bitmaps(const okv *k) { return k->bitmaps; }
const uint8_t *okv_obitmap(const okv *k) { return k->obitmap; }
const uint8_t *okv_packed(const okv *k) { return k->packed; }
