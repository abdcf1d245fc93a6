/*
 * oracle.c - CPU restatement of GLINT's per-tick contact-tracing path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Compiled with -ffp-contract=off so
 * every floating-point expression below rounds exactly as written.
 *
 * Semantics, with the reference lines each part follows:
 *  - Tree build from the first tick's positions (QTree.h:157-214): a node
 *    splits iff the number of points routed to it is >= grid_capacity and
 *    width/2 > r; children BL, BR, TL, TR with boxes (lo,mid), (mid,hi)
 *    (QTree.h:163-166); mid = (hi+lo)/2 (QTree.h:103-104); a point goes to
 *    child (y>mid_y)*2 + (x>mid_x) (QTree.h:174, filter.cpp:231).
 *  - Filter (filter.cpp:229-261): descend, keeping last_valid = deepest node
 *    on the path with x+r < node.high.x && y+r < node.high.y; the object is
 *    inserted into its leaf (one "inside" check unit); if last_valid is not
 *    the leaf, lookup_rec from last_valid adds border units.
 *  - lookup_rec (filter.cpp:130-152): visit children whose box distance to
 *    the point is <= r (geometry.h:174-185); a leaf is accepted iff
 *    p.y < leaf.low.y || (p.y < leaf.high.y && p.x < leaf.low.x).
 *  - Refine (refinement.cpp:19-58): for a unit (pid, leaf, inside) and every
 *    q in the leaf: if pid < q || !inside, test sqrt(dx*dx+dy*dy) <= r
 *    (geometry.h:40-48) and record the unordered pair.
 *  - Evaluate (meeting.cpp:134-155): a pair seen this tick stays active; a
 *    pair not seen this tick ends; it is emitted iff cur - start >= min + 1;
 *    either way it leaves the table.
 *  - Schema update (geometry.cu:1193-1234 collect, 1114-1158 split,
 *    1160-1191 merge): a leaf with height > 2r and count > cap splits one
 *    level (child boxes lo + (i%2)*half, mid + (i%2)*half); a branch whose
 *    four children are leaves with total count < cap merges; decisions use
 *    this tick's counts before any change.
 *  - Grid rebuild (geometry.cu:1270-1393,1629-1645), the reference's GPU
 *    path with dynamic_schema off: every tick, the tree of the 16384^2 cell
 *    histogram in which a block is a branch iff it holds more than
 *    grid_capacity points (see grid_build below).
 */
#include "oracle.h"
#include "oracle_kv_internal.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

enum { T_INVALID = 0, T_BRANCH = 1, T_LEAF = 2 };

typedef struct onode {
    double lo[2], hi[2], mid[2];
    uint32_t child[4];
    int32_t type;
    int32_t level;
    uint32_t overflow, underflow;
} onode;

typedef struct ubuf {   /* growable u64 buffer */
    uint64_t *v;
    uint64_t n, cap;
} ubuf;

typedef struct octx {
    oracle_config cfg;
    onode *nodes;
    uint32_t nnodes, cap_nodes;
    uint32_t *freelist;
    uint32_t nfree;
    int built;
    uint32_t *leaf_of, *lv_of;
    uint32_t *count;     /* per node, leaf population of the last tick */
    uint32_t *mstart;    /* per node, first member index */
    uint32_t *members;
    ubuf contacts;
    uint64_t *akeys;     /* active pairs, sorted */
    uint32_t *astart, *alast;
    float *abox;         /* 4 per active pair when key-value output is on */
    uint64_t nactive, cap_active;
    okv *kv;
    uint32_t last_tick;
    uint32_t *meet;      /* 4 u32 per meeting */
    uint64_t nmeet, cap_meet;
    oracle_stats st;
} octx;

/* ---------------------------------------------------------------- helpers */

static void ubuf_push(ubuf *b, uint64_t x)
{
    if (b->n == b->cap) {
        b->cap = b->cap ? b->cap * 2 : 1024;
        b->v = (uint64_t *)realloc(b->v, b->cap * sizeof(uint64_t));
    }
    b->v[b->n++] = x;
}

int oracle_point_within(double x1, double y1, double x2, double y2, double r)
{
    /* Point::distance with geography=false (geometry.h:40-48). */
    double dx = x1 - x2;
    double dy = y1 - y2;
    return sqrt(dx * dx + dy * dy) <= r;
}

/* util.h:38,57: the reference's PI and degrees of latitude per metre */
#define ORACLE_REF_PI 3.14159265
static const double kDegPerMeterLat = 360.0 / (40076.0 * 1000.0);

double oracle_degree_per_meter_longitude(double latitude)
{
    /* degree_per_meter_longitude (util.h:59-66); 90 is taken as 89 */
    double absla = fabs(latitude);
    if (absla == 90) absla = 89;
    return 360.0 / (sin((90 - absla) * ORACLE_REF_PI / 180) * 40076.0 * 1000.0);
}

int oracle_point_within_geo(double x1, double y1, double x2, double y2, double r)
{
    /* Point::distance with geography=true (geometry.h:40-48): p1 = (x1, y1) */
    double dx = x1 - x2;
    double dy = y1 - y2;
    dy = dy / kDegPerMeterLat;
    dx = dx / oracle_degree_per_meter_longitude(y1);
    return sqrt(dx * dx + dy * dy) <= r;
}

static int box_contains(const onode *n, double px, double py)
{
    /* box::contain(Point&) (geometry.h:140-145) */
    return px >= n->lo[0] && px <= n->hi[0] && py >= n->lo[1] && py <= n->hi[1];
}

static double maxd(double a, double b) { return a < b ? b : a; }   /* std::max */

double oracle_box_distance(double lx, double ly, double hx, double hy, double px, double py)
{
    /* box::distance(Point&, false) (geometry.h:174-185) */
    if (px >= lx && px <= hx && py >= ly && py <= hy) return 0;
    double dx = maxd(fabs(px - (lx + hx) / 2) - (hx - lx) / 2, 0.0);
    double dy = maxd(fabs(py - (ly + hy) / 2) - (hy - ly) / 2, 0.0);
    return sqrt(dx * dx + dy * dy);
}

double oracle_box_distance_geo(double lx, double ly, double hx, double hy, double px, double py)
{
    /* box::distance(Point&, true) (geometry.h:174-185) */
    if (px >= lx && px <= hx && py >= ly && py <= hy) return 0;
    double dx = maxd(fabs(px - (lx + hx) / 2) - (hx - lx) / 2, 0.0);
    double dy = maxd(fabs(py - (ly + hy) / 2) - (hy - ly) / 2, 0.0);
    dy = dy / kDegPerMeterLat;
    dx = dx / oracle_degree_per_meter_longitude(py);
    return sqrt(dx * dx + dy * dy);
}

static double node_distance(const octx *c, const onode *n, double px, double py)
{
    if (box_contains(n, px, py)) return 0;
    if (c->cfg.geography) return oracle_box_distance_geo(n->lo[0], n->lo[1], n->hi[0], n->hi[1], px, py);
    return oracle_box_distance(n->lo[0], n->lo[1], n->hi[0], n->hi[1], px, py);
}

/* The pair test of reachability_unit (refinement.cpp:33): p1 is the unit's object. */
static int pair_within(const octx *c, double x1, double y1, double x2, double y2, double r)
{
    return c->cfg.geography ? oracle_point_within_geo(x1, y1, x2, y2, r) : oracle_point_within(x1, y1, x2, y2, r);
}

static uint32_t alloc_node(octx *c)
{
    if (c->nfree) return c->freelist[--c->nfree];
    if (c->nnodes == c->cap_nodes) {
        uint32_t nc = c->cap_nodes ? c->cap_nodes * 2 : 1024;
        c->nodes = (onode *)realloc(c->nodes, (size_t)nc * sizeof(onode));
        c->count = (uint32_t *)realloc(c->count, (size_t)nc * sizeof(uint32_t));
        c->mstart = (uint32_t *)realloc(c->mstart, (size_t)(nc + 1) * sizeof(uint32_t));
        c->freelist = (uint32_t *)realloc(c->freelist, (size_t)nc * sizeof(uint32_t));
        memset(c->count + c->cap_nodes, 0, (size_t)(nc - c->cap_nodes) * sizeof(uint32_t));
        c->cap_nodes = nc;
    }
    return c->nnodes++;
}

static void init_node(onode *n, double lx, double ly, double hx, double hy, int level)
{
    memset(n, 0, sizeof(*n));
    n->lo[0] = lx;
    n->lo[1] = ly;
    n->hi[0] = hx;
    n->hi[1] = hy;
    n->mid[0] = (hx + lx) / 2;   /* QTree.h:103 */
    n->mid[1] = (hy + ly) / 2;
    n->type = T_LEAF;
    n->level = level;
}

/* ------------------------------------------------------------------ build */

static void build_rec(octx *c, uint32_t node, uint32_t *idx, uint32_t n, uint32_t *tmp,
                      const double *xy)
{
    onode nd = c->nodes[node];
    double width = nd.hi[0] - nd.lo[0];
    /* box::width(geography) (geometry.h:166-172): metres at the low latitude */
    if (c->cfg.geography) width = width / oracle_degree_per_meter_longitude(nd.lo[1]);
    /* QTNode::split (QTree.h:158-159): object_index >= cap && width/2 > r */
    if (!(n >= c->cfg.grid_capacity && width / 2 > c->cfg.reach_distance)) {
        c->nodes[node].type = T_LEAF;
        return;
    }
    uint32_t ch[4];
    for (int i = 0; i < 4; i++) ch[i] = alloc_node(c);
    /* QTree.h:163-166 */
    init_node(&c->nodes[ch[0]], nd.lo[0], nd.lo[1], nd.mid[0], nd.mid[1], nd.level + 1);
    init_node(&c->nodes[ch[1]], nd.mid[0], nd.lo[1], nd.hi[0], nd.mid[1], nd.level + 1);
    init_node(&c->nodes[ch[2]], nd.lo[0], nd.mid[1], nd.mid[0], nd.hi[1], nd.level + 1);
    init_node(&c->nodes[ch[3]], nd.mid[0], nd.mid[1], nd.hi[0], nd.hi[1], nd.level + 1);
    onode *p = &c->nodes[node];
    p->type = T_BRANCH;
    for (int i = 0; i < 4; i++) p->child[i] = ch[i];

    uint32_t cnt[4] = {0, 0, 0, 0};
    for (uint32_t k = 0; k < n; k++) {
        double x = xy[2 * (size_t)idx[k]], y = xy[2 * (size_t)idx[k] + 1];
        cnt[(y > nd.mid[1]) * 2 + (x > nd.mid[0])]++;
    }
    uint32_t off[4] = {0, cnt[0], cnt[0] + cnt[1], cnt[0] + cnt[1] + cnt[2]};
    uint32_t pos[4] = {off[0], off[1], off[2], off[3]};
    for (uint32_t k = 0; k < n; k++) {
        double x = xy[2 * (size_t)idx[k]], y = xy[2 * (size_t)idx[k] + 1];
        tmp[pos[(y > nd.mid[1]) * 2 + (x > nd.mid[0])]++] = idx[k];
    }
    memcpy(idx, tmp, (size_t)n * sizeof(uint32_t));
    for (int i = 0; i < 4; i++) build_rec(c, ch[i], idx + off[i], cnt[i], tmp + off[i], xy);
}

int oracle_build(void *h, const double *xy)
{
    octx *c = (octx *)h;
    c->nnodes = 0;
    c->nfree = 0;
    uint32_t root = alloc_node(c);
    init_node(&c->nodes[root], c->cfg.low_x, c->cfg.low_y, c->cfg.high_x, c->cfg.high_y, 0);
    uint32_t n = c->cfg.num_objects;
    uint32_t *idx = (uint32_t *)malloc((size_t)n * sizeof(uint32_t));
    uint32_t *tmp = (uint32_t *)malloc((size_t)n * sizeof(uint32_t));
    for (uint32_t i = 0; i < n; i++) idx[i] = i;
    build_rec(c, root, idx, n, tmp, xy);
    free(idx);
    free(tmp);
    memset(c->count, 0, (size_t)c->cap_nodes * sizeof(uint32_t));
    c->built = 1;
    return 0;
}

/* ------------------------------------------------------ grid rebuild (GPU, static) */

/* The reference's GPU path without a dynamic schema rebuilds the tree from
 * its points every tick (geometry.cu:1629-1645): a 16384 x 16384 histogram of
 * cells (cuda_build_qtree, geometry.cu:1270-1279), merged bottom-up one level
 * at a time (cuda_merge_qtree, geometry.cu:1295-1393).  A block of cells is a
 * branch iff it holds more than grid_capacity points; a branch cell gets four
 * leaf children; node 0 is the root.  Restated here top-down: a branch's
 * children are branches iff their own cell count exceeds the capacity (the
 * branch set is closed upwards, so both orders give one tree).  Boxes:
 * a branch block (bx, by) at depth d (xdim = 2^d blocks per side) spans
 * low + bx*(high-low)/xdim .. low + (bx+1)*(high-low)/xdim (geometry.cu:
 * 1304-1307,1335-1338); a leaf child i spans lo + (i%2)*xhalf .. mid +
 * (i%2)*xhalf with xhalf = mid - lo (geometry.cu:1311-1326).  When no block
 * exceeds the capacity the reference writes no node at all and keeps the
 * node 0 of the tick before; here the root is then one leaf with the
 * depth-0 block box. */
#define GRID_BITS 14
#define GRID_DIM (1u << GRID_BITS)

static void grid_block_box(const octx *c, int d, uint32_t bx, uint32_t by, onode *n)
{
    const double xdim = (double)(1u << d);
    const double lx = c->cfg.low_x, ly = c->cfg.low_y, hx = c->cfg.high_x, hy = c->cfg.high_y;
    n->lo[0] = lx + (double)bx * (hx - lx) / xdim;
    n->lo[1] = ly + (double)by * (hy - ly) / xdim;
    n->hi[0] = lx + (double)(bx + 1) * (hx - lx) / xdim;
    n->hi[1] = ly + (double)(by + 1) * (hy - ly) / xdim;
    n->mid[0] = (n->lo[0] + n->hi[0]) / 2;
    n->mid[1] = (n->lo[1] + n->hi[1]) / 2;
}

static void grid_rec(octx *c, uint32_t node, int d, uint32_t bx, uint32_t by, uint32_t *idx, uint32_t n,
                     uint32_t *tmp, const uint32_t *cell)
{
    onode b;
    memset(&b, 0, sizeof(b));
    grid_block_box(c, d, bx, by, &b);
    b.type = T_BRANCH;
    b.level = d;
    const double xhalf = b.mid[0] - b.lo[0], yhalf = b.mid[1] - b.lo[1];
    /* the points of each child block: cell bit (13 - d) of x and y */
    uint32_t cnt[4] = {0, 0, 0, 0}, off[4], pos[4];
    if (d < GRID_BITS) {
        const int sh = GRID_BITS - 1 - d;
        for (uint32_t k = 0; k < n; k++) {
            const uint32_t lin = cell[idx[k]];
            cnt[(((lin >> GRID_BITS) >> sh) & 1) * 2 + (((lin & (GRID_DIM - 1)) >> sh) & 1)]++;
        }
        off[0] = 0;
        for (int i = 1; i < 4; i++) off[i] = off[i - 1] + cnt[i - 1];
        for (int i = 0; i < 4; i++) pos[i] = off[i];
        for (uint32_t k = 0; k < n; k++) {
            const uint32_t lin = cell[idx[k]];
            tmp[pos[(((lin >> GRID_BITS) >> sh) & 1) * 2 + (((lin & (GRID_DIM - 1)) >> sh) & 1)]++] = idx[k];
        }
        memcpy(idx, tmp, (size_t)n * sizeof(uint32_t));
    }
    uint32_t ch[4];
    for (int i = 0; i < 4; i++) ch[i] = alloc_node(c);
    for (int i = 0; i < 4; i++) b.child[i] = ch[i];
    c->nodes[node] = b;
    for (int i = 0; i < 4; i++) {
        const uint32_t cbx = 2 * bx + (uint32_t)(i & 1), cby = 2 * by + (uint32_t)(i >> 1);
        if (d < GRID_BITS && cnt[i] > c->cfg.grid_capacity) {
            grid_rec(c, ch[i], d + 1, cbx, cby, idx + off[i], cnt[i], tmp + off[i], cell);
            continue;
        }
        onode l;
        memset(&l, 0, sizeof(l));
        l.lo[0] = b.lo[0] + (double)(i % 2 == 1) * xhalf;
        l.lo[1] = b.lo[1] + (double)(i / 2 == 1) * yhalf;
        l.hi[0] = b.mid[0] + (double)(i % 2 == 1) * xhalf;
        l.hi[1] = b.mid[1] + (double)(i / 2 == 1) * yhalf;
        l.mid[0] = (l.lo[0] + l.hi[0]) / 2;
        l.mid[1] = (l.lo[1] + l.hi[1]) / 2;
        l.type = T_LEAF;
        l.level = d + 1;
        c->nodes[ch[i]] = l;
    }
}

/* The cell of every point as the reference computes it (geometry.cu:1276-
 * 1278), x + y*16384 in 32-bit arithmetic: a point on the high x edge lands
 * in the first cell of the next row, as there; one on the high y edge has
 * no cell (the reference writes past its histogram) and is refused. */
int oracle_grid_cells(const oracle_config *cfg, const double *xy, uint32_t *cell)
{
    int bad = 0;
    const double lx = cfg->low_x, ly = cfg->low_y, hx = cfg->high_x, hy = cfg->high_y;
#pragma omp parallel for reduction(| : bad)
    for (int64_t i = 0; i < (int64_t)cfg->num_objects; i++) {
        const uint32_t x = (uint32_t)((xy[2 * i] - lx) / (hx - lx) * GRID_DIM);
        const uint32_t y = (uint32_t)((xy[2 * i + 1] - ly) / (hy - ly) * GRID_DIM);
        const uint32_t lin = x + y * GRID_DIM;
        cell[i] = lin;
        if (lin >= GRID_DIM * GRID_DIM) bad = 1;
    }
    return bad ? -2 : 0;
}

static int grid_build(octx *c, const double *xy)
{
    const uint32_t n = c->cfg.num_objects;
    uint32_t *cell = (uint32_t *)malloc((size_t)n * sizeof(uint32_t));
    if (oracle_grid_cells(&c->cfg, xy, cell)) {
        free(cell);
        return -2;
    }
    c->nnodes = 0;
    c->nfree = 0;
    const uint32_t root = alloc_node(c);
    if (n > c->cfg.grid_capacity) {
        uint32_t *idx = (uint32_t *)malloc((size_t)n * sizeof(uint32_t));
        uint32_t *tmp = (uint32_t *)malloc((size_t)n * sizeof(uint32_t));
        for (uint32_t i = 0; i < n; i++) idx[i] = i;
        grid_rec(c, root, 0, 0, 0, idx, n, tmp, cell);
        free(idx);
        free(tmp);
    } else {
        onode *r = &c->nodes[root];
        memset(r, 0, sizeof(*r));
        grid_block_box(c, 0, 0, 0, r);
        r->type = T_LEAF;
        r->level = 0;
    }
    free(cell);
    memset(c->count, 0, (size_t)c->cap_nodes * sizeof(uint32_t));
    c->built = 1;
    return 0;
}

/* ------------------------------------------------------------ filter/lookup */

static uint32_t descend(const octx *c, double x, double y, uint32_t *last_valid)
{
    uint32_t cur = 0, lv = 0;
    /* x_buffer, y_buffer (filter.cpp:30-31): r, or r in degrees at the root's low latitude */
    const double r = c->cfg.reach_distance;
    const double xb = c->cfg.geography ? r * oracle_degree_per_meter_longitude(c->cfg.low_y) : r;
    const double yb = c->cfg.geography ? r * kDegPerMeterLat : r;
    if (c->nodes[0].type == T_LEAF) {   /* single-leaf tree: nothing to look up */
        *last_valid = 0;
        return 0;
    }
    for (;;) {
        const onode *n = &c->nodes[cur];
        int ch = (y > n->mid[1]) * 2 + (x > n->mid[0]);
        cur = n->child[ch];
        const onode *m = &c->nodes[cur];
        if (x + xb < m->hi[0] && y + yb < m->hi[1]) lv = cur;   /* filter.cpp:235-238 */
        if (m->type == T_LEAF) break;
    }
    *last_valid = lv;
    return cur;
}

typedef struct cbuf {
    uint32_t *v;
    uint32_t n, cap;
} cbuf;

static void cbuf_push(cbuf *b, uint32_t x)
{
    if (b->n == b->cap) {
        b->cap = b->cap ? b->cap * 2 : 16;
        b->v = (uint32_t *)realloc(b->v, b->cap * sizeof(uint32_t));
    }
    b->v[b->n++] = x;
}

/* lookup_rec (filter.cpp:130-152); in complete mode every leaf within r
 * except the point's own is accepted. */
static void lookup(const octx *c, double x, double y, uint32_t node, uint32_t own, cbuf *out)
{
    const double r = c->cfg.reach_distance;
    for (int i = 0; i < 4; i++) {
        uint32_t ch = c->nodes[node].child[i];
        const onode *m = &c->nodes[ch];
        double dist = node_distance(c, m, x, y);
        if (dist <= r) {
            if (m->type == T_LEAF) {
                if (c->cfg.complete_mode) {
                    if (ch != own) cbuf_push(out, ch);
                } else if (y < m->lo[1] || (y < m->hi[1] && x < m->lo[0])) {
                    cbuf_push(out, ch);
                }
            } else {
                lookup(c, x, y, ch, own, out);
            }
        }
    }
}

/* ------------------------------------------------------------- radix sort */

static void radix_sort_u64(uint64_t *a, uint64_t n)
{
    if (n < 2) return;
    uint64_t *b = (uint64_t *)malloc(n * sizeof(uint64_t));
    uint64_t or_all = 0, and_all = ~0ull;
    for (uint64_t i = 0; i < n; i++) {
        or_all |= a[i];
        and_all &= a[i];
    }
    int nt = omp_get_max_threads();
    uint64_t *hist = (uint64_t *)malloc((size_t)nt * 256 * sizeof(uint64_t));
    for (int shift = 0; shift < 64; shift += 8) {
        if ((((or_all ^ and_all) >> shift) & 0xff) == 0) continue;   /* digit constant */
        memset(hist, 0, (size_t)nt * 256 * sizeof(uint64_t));
#pragma omp parallel num_threads(nt)
        {
            int t = omp_get_thread_num(), T = omp_get_num_threads();
            uint64_t lo = n * t / T, hi = n * (t + 1) / T;
            uint64_t *hh = hist + (size_t)t * 256;
            for (uint64_t i = lo; i < hi; i++) hh[(a[i] >> shift) & 0xff]++;
#pragma omp barrier
#pragma omp single
            {
                uint64_t sum = 0;
                for (int d = 0; d < 256; d++)
                    for (int tt = 0; tt < T; tt++) {
                        uint64_t v = hist[(size_t)tt * 256 + d];
                        hist[(size_t)tt * 256 + d] = sum;
                        sum += v;
                    }
            }
            for (uint64_t i = lo; i < hi; i++) b[hh[(a[i] >> shift) & 0xff]++] = a[i];
        }
        uint64_t *t = a;
        memcpy(t, b, n * sizeof(uint64_t));
    }
    free(hist);
    free(b);
}

static uint64_t unique_u64(uint64_t *a, uint64_t n)
{
    if (!n) return 0;
    uint64_t k = 1;
    for (uint64_t i = 1; i < n; i++)
        if (a[i] != a[k - 1]) a[k++] = a[i];
    return k;
}

/* --------------------------------------------------------------- evaluate */

static void meet_push(octx *c, uint64_t key, uint32_t start, uint32_t end)
{
    if (c->nmeet == c->cap_meet) {
        c->cap_meet = c->cap_meet ? c->cap_meet * 2 : 1024;
        c->meet = (uint32_t *)realloc(c->meet, c->cap_meet * 4 * sizeof(uint32_t));
    }
    uint32_t *m = c->meet + 4 * c->nmeet++;
    m[0] = (uint32_t)(key >> 32);
    m[1] = (uint32_t)key;
    m[2] = start;
    m[3] = end;
}

/* mbr_update_by_point (geometry.cu:82-97): float box, double point. */
static void box_point(float *m, double px, double py)
{
    if (m[0] > px) m[0] = (float)px;
    if (m[2] < px) m[2] = (float)px;
    if (m[1] > py) m[1] = (float)py;
    if (m[3] < py) m[3] = (float)py;
}

/* The two points of a contact grow the pair's box (geometry.cu:474-489). */
static void box_contact(float *m, const double *xy, uint64_t key)
{
    const size_t a = (size_t)(key >> 32), b = (size_t)(uint32_t)key;
    box_point(m, xy[2 * a], xy[2 * a + 1]);
    box_point(m, xy[2 * b], xy[2 * b + 1]);
}

/* keys: sorted unique contact pairs of this tick; xy: this tick's
 * positions (needed only for key-value output). */
static void evaluate(octx *c, const uint64_t *keys, uint64_t nk, uint32_t tick, const double *xy)
{
    uint64_t cap = c->nactive + nk;
    uint64_t *nkeys = (uint64_t *)malloc((cap ? cap : 1) * sizeof(uint64_t));
    uint32_t *nstart = (uint32_t *)malloc((cap ? cap : 1) * sizeof(uint32_t));
    uint32_t *nlast = (uint32_t *)malloc((cap ? cap : 1) * sizeof(uint32_t));
    const int boxes = c->kv != NULL && xy != NULL;
    float *nbox = boxes ? (float *)malloc((cap ? cap : 1) * 4 * sizeof(float)) : NULL;
    uint64_t i = 0, j = 0, k = 0;
    const uint32_t minm = c->cfg.min_meet_time;
    c->nmeet = 0;
    while (i < c->nactive || j < nk) {
        if (j >= nk || (i < c->nactive && c->akeys[i] < keys[j])) {
            /* not seen this tick: ended (meeting.cpp:147-154) */
            if (tick - c->astart[i] >= minm + 1) {
                meet_push(c, c->akeys[i], c->astart[i], c->alast[i]);
                if (boxes)
                    okv_emit(c->kv, (uint32_t)(c->akeys[i] >> 32), (uint32_t)c->akeys[i], c->astart[i], c->alast[i],
                             c->abox + 4 * i);
            }
            i++;
        } else if (i >= c->nactive || keys[j] < c->akeys[i]) {
            nkeys[k] = keys[j];   /* new bucket: start = end = cur (refinement.cpp:47-50) */
            nstart[k] = tick;
            nlast[k] = tick;
            if (boxes) {
                float *m = nbox + 4 * k;
                m[0] = m[1] = 100000.0f;
                m[2] = m[3] = -100000.0f;
                box_contact(m, xy, keys[j]);
            }
            k++;
            j++;
        } else {
            nkeys[k] = keys[j];   /* still active: end = cur (refinement.cpp:40-41) */
            nstart[k] = c->astart[i];
            nlast[k] = tick;
            if (boxes) {
                memcpy(nbox + 4 * k, c->abox + 4 * i, 4 * sizeof(float));
                box_contact(nbox + 4 * k, xy, keys[j]);
            }
            k++;
            i++;
            j++;
        }
    }
    free(c->akeys);
    free(c->astart);
    free(c->alast);
    free(c->abox);
    c->akeys = nkeys;
    c->astart = nstart;
    c->alast = nlast;
    c->abox = nbox;
    c->nactive = k;
    c->cap_active = cap;
}

int oracle_evaluate_contacts(void *h, const uint64_t *keys, uint64_t n, uint32_t tick)
{
    octx *c = (octx *)h;
    uint64_t *tmp = (uint64_t *)malloc((n ? n : 1) * sizeof(uint64_t));
    memcpy(tmp, keys, n * sizeof(uint64_t));
    radix_sort_u64(tmp, n);
    n = unique_u64(tmp, n);
    evaluate(c, tmp, n, tick, NULL);
    free(tmp);
    return 0;
}

/* ---------------------------------------------------------- schema update */

static void split_node(octx *c, uint32_t node)
{
    /* geometry.cu:1123-1157 */
    onode nd = c->nodes[node];
    double xhalf = nd.mid[0] - nd.lo[0];
    double yhalf = nd.mid[1] - nd.lo[1];
    uint32_t ch[4];
    for (int i = 0; i < 4; i++) {
        ch[i] = alloc_node(c);
        onode *m = &c->nodes[ch[i]];
        memset(m, 0, sizeof(*m));
        m->lo[0] = nd.lo[0] + (i % 2 == 1) * xhalf;
        m->lo[1] = nd.lo[1] + (i / 2 == 1) * yhalf;
        m->hi[0] = nd.mid[0] + (i % 2 == 1) * xhalf;
        m->hi[1] = nd.mid[1] + (i / 2 == 1) * yhalf;
        m->mid[0] = (m->lo[0] + m->hi[0]) / 2;
        m->mid[1] = (m->lo[1] + m->hi[1]) / 2;
        m->type = T_LEAF;
        m->level = nd.level + 1;
        c->count[ch[i]] = 0;
    }
    onode *p = &c->nodes[node];
    p->type = T_BRANCH;
    for (int i = 0; i < 4; i++) p->child[i] = ch[i];
}

static void merge_node(octx *c, uint32_t node)
{
    /* geometry.cu:1169-1190 */
    for (int i = 0; i < 4; i++) {
        uint32_t ch = c->nodes[node].child[i];
        c->nodes[ch].type = T_INVALID;
        c->count[ch] = 0;
        c->freelist[c->nfree++] = ch;
    }
    c->nodes[node].type = T_LEAF;
    c->count[node] = 0;
}

static void update_schema(octx *c)
{
    const double r = c->cfg.reach_distance;
    const uint32_t cap = c->cfg.grid_capacity, delay = c->cfg.schema_update_delay;
    uint32_t *splits = (uint32_t *)malloc((size_t)c->nnodes * sizeof(uint32_t) + 4);
    uint32_t *merges = (uint32_t *)malloc((size_t)c->nnodes * sizeof(uint32_t) + 4);
    uint32_t ns = 0, nm = 0;
    for (uint32_t n = 0; n < c->nnodes; n++) {
        onode *nd = &c->nodes[n];
        if (nd->type == T_LEAF) {
            /* geometry.cu:1200-1210 (height(), which equals width() on square boxes;
             * in metres with geography, geometry.cu:30) */
            const double ht = c->cfg.geography ? (nd->hi[1] - nd->lo[1]) / kDegPerMeterLat : nd->hi[1] - nd->lo[1];
            if (ht > 2 * r && c->count[n] > cap) {
                if (++nd->overflow >= delay) {
                    splits[ns++] = n;
                    nd->overflow = 0;
                }
            } else {
                nd->overflow = 0;
            }
        } else if (nd->type == T_BRANCH) {
            int leafchild = 0;
            int ncounter = 0;
            for (int i = 0; i < 4; i++) {
                uint32_t ch = nd->child[i];
                if (c->nodes[ch].type == T_LEAF) {
                    leafchild++;
                    ncounter += (int)c->count[ch];
                }
            }
            if (leafchild == 4 && ncounter < (int)cap) {   /* geometry.cu:1222 */
                if (++nd->underflow >= delay) {
                    merges[nm++] = n;
                    nd->underflow = 0;
                }
            } else {
                nd->underflow = 0;
            }
        }
    }
    for (uint32_t k = 0; k < ns; k++) split_node(c, splits[k]);
    for (uint32_t k = 0; k < nm; k++) merge_node(c, merges[k]);
    c->st.splits = ns;
    c->st.merges = nm;
    free(splits);
    free(merges);
}

/* ------------------------------------------------------------------- tick */

void *oracle_create(const oracle_config *cfg)
{
    if (!cfg || cfg->num_objects == 0 || cfg->grid_capacity == 0) return NULL;
    if (!(cfg->reach_distance > 0)) return NULL;
    if (cfg->geography && cfg->dynamic_schema == ORACLE_SCHEMA_GRID_REBUILD) return NULL;
    if (!(cfg->high_x > cfg->low_x) || !(cfg->high_y > cfg->low_y)) return NULL;
    octx *c = (octx *)calloc(1, sizeof(octx));
    c->cfg = *cfg;
    if (c->cfg.schema_update_delay == 0) c->cfg.schema_update_delay = 1;
    size_t n = cfg->num_objects;
    c->leaf_of = (uint32_t *)malloc(n * sizeof(uint32_t));
    c->lv_of = (uint32_t *)malloc(n * sizeof(uint32_t));
    c->members = (uint32_t *)malloc(n * sizeof(uint32_t));
    return c;
}

void oracle_destroy(void *h)
{
    octx *c = (octx *)h;
    if (!c) return;
    free(c->nodes);
    free(c->count);
    free(c->mstart);
    free(c->freelist);
    free(c->leaf_of);
    free(c->lv_of);
    free(c->members);
    free(c->contacts.v);
    free(c->akeys);
    free(c->astart);
    free(c->alast);
    free(c->abox);
    okv_destroy(c->kv);
    free(c->meet);
    free(c);
}

static int check_inputs(const octx *c, const double *xy)
{
    const uint64_t n = c->cfg.num_objects;
    int bad = 0;
#pragma omp parallel for reduction(| : bad)
    for (int64_t i = 0; i < (int64_t)n; i++) {
        double x = xy[2 * i], y = xy[2 * i + 1];
        if (!(x >= c->cfg.low_x && x <= c->cfg.high_x && y >= c->cfg.low_y && y <= c->cfg.high_y))
            bad = 1;
    }
    return bad ? -2 : 0;
}

int oracle_tick(void *h, const double *xy, uint32_t tick)
{
    octx *c = (octx *)h;
    const uint32_t N = c->cfg.num_objects;
    const double r = c->cfg.reach_distance;
    if (c->cfg.num_threads) omp_set_num_threads((int)c->cfg.num_threads);
    if (check_inputs(c, xy)) return -2;
    memset(&c->st, 0, sizeof(c->st));
    double t0 = omp_get_wtime();
    if (c->cfg.dynamic_schema == ORACLE_SCHEMA_GRID_REBUILD) {
        if (grid_build(c, xy)) return -2;
    } else if (!c->built) {
        oracle_build(c, xy);
    }

    /* ---- filter: descent, last_valid, leaf membership (filter.cpp:222-261) */
    memset(c->count, 0, (size_t)c->cap_nodes * sizeof(uint32_t));
    uint64_t nborder = 0;
#pragma omp parallel for schedule(static) reduction(+ : nborder)
    for (int64_t i = 0; i < (int64_t)N; i++) {
        uint32_t lv;
        uint32_t leaf = descend(c, xy[2 * i], xy[2 * i + 1], &lv);
        c->leaf_of[i] = leaf;
        c->lv_of[i] = lv;
        nborder += (lv != leaf) || c->cfg.complete_mode;
#pragma omp atomic
        c->count[leaf]++;
    }
    /* members of each leaf, ascending pid */
    uint32_t acc = 0;
    for (uint32_t n = 0; n < c->nnodes; n++) {
        c->mstart[n] = acc;
        acc += c->count[n];
    }
    c->mstart[c->nnodes] = acc;
    {
        uint32_t *cur = (uint32_t *)malloc((size_t)c->nnodes * sizeof(uint32_t) + 4);
        memcpy(cur, c->mstart, (size_t)c->nnodes * sizeof(uint32_t));
        for (uint32_t i = 0; i < N; i++) c->members[cur[c->leaf_of[i]]++] = i;
        free(cur);
    }
    double t1 = omp_get_wtime();

    /* ---- refine: inside units and border units (refinement.cpp:19-58) */
    int nt = omp_get_max_threads();
    ubuf *tb = (ubuf *)calloc((size_t)nt, sizeof(ubuf));
    uint64_t ntests = 0, nunits = N;
#pragma omp parallel reduction(+ : ntests, nunits)
    {
        ubuf *mine = &tb[omp_get_thread_num()];
        cbuf cands = {0, 0, 0};
#pragma omp for schedule(dynamic, 1024)
        for (int64_t i = 0; i < (int64_t)N; i++) {
            const uint32_t pid = (uint32_t)i;
            const double px = xy[2 * i], py = xy[2 * i + 1];
            const uint32_t leaf = c->leaf_of[i];
            /* inside unit (pid, leaf, inside=true) */
            for (uint32_t k = c->mstart[leaf]; k < c->mstart[leaf] + c->count[leaf]; k++) {
                uint32_t q = c->members[k];
                if (pid < q) {
                    ntests++;
                    if (pair_within(c, px, py, xy[2 * (size_t)q], xy[2 * (size_t)q + 1], r))
                        ubuf_push(mine, ((uint64_t)pid << 32) | q);
                }
            }
            /* border units (pid, leaf', inside=false) */
            cands.n = 0;
            if (c->cfg.complete_mode) {
                if (c->nodes[0].type == T_BRANCH) lookup(c, px, py, 0, leaf, &cands);
            } else if (c->lv_of[i] != leaf) {
                lookup(c, px, py, c->lv_of[i], leaf, &cands);
            }
            nunits += cands.n;
            for (uint32_t u = 0; u < cands.n; u++) {
                uint32_t l2 = cands.v[u];
                for (uint32_t k = c->mstart[l2]; k < c->mstart[l2] + c->count[l2]; k++) {
                    uint32_t q = c->members[k];
                    ntests++;
                    if (pair_within(c, px, py, xy[2 * (size_t)q], xy[2 * (size_t)q + 1], r)) {
                        uint32_t a = pid < q ? pid : q, b = pid < q ? q : pid;
                        ubuf_push(mine, ((uint64_t)a << 32) | b);
                    }
                }
            }
        }
        free(cands.v);
    }
    uint64_t total = 0;
    for (int t = 0; t < nt; t++) total += tb[t].n;
    c->contacts.n = 0;
    if (c->contacts.cap < total) {
        c->contacts.cap = total;
        c->contacts.v = (uint64_t *)realloc(c->contacts.v, (total ? total : 1) * sizeof(uint64_t));
    }
    for (int t = 0; t < nt; t++) {
        memcpy(c->contacts.v + c->contacts.n, tb[t].v, tb[t].n * sizeof(uint64_t));
        c->contacts.n += tb[t].n;
        free(tb[t].v);
    }
    free(tb);
    radix_sort_u64(c->contacts.v, c->contacts.n);
    c->contacts.n = unique_u64(c->contacts.v, c->contacts.n);
    double t2 = omp_get_wtime();

    /* ---- evaluate (meeting.cpp:134-155) */
    evaluate(c, c->contacts.v, c->contacts.n, tick, xy);
    c->last_tick = tick;
    double t3 = omp_get_wtime();

    /* ---- stats, then schema update (process.cpp:608-612) */
    uint64_t nleaves = 0, maxc = 0, drops = 0, drop_objs = 0;
    for (uint32_t n = 0; n < c->nnodes; n++) {
        if (c->nodes[n].type != T_LEAF) continue;
        nleaves++;
        if (c->count[n] > maxc) maxc = c->count[n];
        if (c->count[n] > 2ull * c->cfg.grid_capacity) {
            drops++;
            drop_objs += c->count[n] - 2ull * c->cfg.grid_capacity;
        }
    }
    c->st.num_leaves = nleaves;
    c->st.num_nodes = c->nnodes - c->nfree;
    c->st.num_border = nborder;
    c->st.num_units = nunits;
    c->st.num_tests = ntests;
    c->st.num_contacts = c->contacts.n;
    c->st.num_active = c->nactive;
    c->st.num_meetings = c->nmeet;
    c->st.max_leaf_count = maxc;
    c->st.ref_drop_leaves = drops;
    c->st.ref_drop_objects = drop_objs;
    if (c->cfg.dynamic_schema == ORACLE_SCHEMA_DYNAMIC) update_schema(c);
    double t4 = omp_get_wtime();
    c->st.t_filter = t1 - t0;
    c->st.t_refine = t2 - t1;
    c->st.t_evaluate = t3 - t2;
    c->st.t_schema = t4 - t3;
    return 0;
}

uint64_t oracle_num_contacts(void *h) { return ((octx *)h)->contacts.n; }
const uint64_t *oracle_contacts(void *h) { return ((octx *)h)->contacts.v; }
uint64_t oracle_num_meetings(void *h) { return ((octx *)h)->nmeet; }
const uint32_t *oracle_meetings(void *h) { return ((octx *)h)->meet; }
void oracle_get_stats(void *h, oracle_stats *out) { *out = ((octx *)h)->st; }
uint64_t oracle_num_active(void *h) { return ((octx *)h)->nactive; }

void oracle_get_active(void *h, uint64_t *keys, uint32_t *starts)
{
    octx *c = (octx *)h;
    memcpy(keys, c->akeys, c->nactive * sizeof(uint64_t));
    memcpy(starts, c->astart, c->nactive * sizeof(uint32_t));
}

uint64_t oracle_flush(void *h, uint32_t *out4, uint64_t cap)
{
    octx *c = (octx *)h;
    uint64_t k = 0;
    for (uint64_t i = 0; i < c->nactive && k < cap; i++, k++) {
        out4[4 * k] = (uint32_t)(c->akeys[i] >> 32);
        out4[4 * k + 1] = (uint32_t)c->akeys[i];
        out4[4 * k + 2] = c->astart[i];
        out4[4 * k + 3] = c->alast[i];
    }
    return k;
}

/* ------------------------------------------------------- test accessors */

static void leaf_order(const octx *c, uint32_t node, uint32_t *out, uint64_t *n)
{
    if (c->nodes[node].type == T_LEAF) {
        out[(*n)++] = node;
        return;
    }
    for (int i = 0; i < 4; i++) leaf_order(c, c->nodes[node].child[i], out, n);
}

uint64_t oracle_num_leaves(void *h)
{
    octx *c = (octx *)h;
    uint64_t k = 0;
    for (uint32_t n = 0; n < c->nnodes; n++) k += c->nodes[n].type == T_LEAF;
    return k;
}

uint64_t oracle_get_leaves(void *h, double *boxes, uint32_t *levels, uint32_t *counts, uint64_t cap)
{
    octx *c = (octx *)h;
    if (!c->nnodes) return 0;
    uint32_t *ord = (uint32_t *)malloc((size_t)c->nnodes * sizeof(uint32_t));
    uint64_t n = 0;
    leaf_order(c, 0, ord, &n);
    for (uint64_t k = 0; k < n && k < cap; k++) {
        const onode *m = &c->nodes[ord[k]];
        if (boxes) {
            boxes[4 * k] = m->lo[0];
            boxes[4 * k + 1] = m->lo[1];
            boxes[4 * k + 2] = m->hi[0];
            boxes[4 * k + 3] = m->hi[1];
        }
        if (levels) levels[k] = (uint32_t)m->level;
        if (counts) counts[k] = c->count[ord[k]];
    }
    free(ord);
    return n;
}

static int64_t leaf_rank(const octx *c, uint32_t node)
{
    uint32_t *ord = (uint32_t *)malloc((size_t)c->nnodes * sizeof(uint32_t));
    uint64_t n = 0;
    leaf_order(c, 0, ord, &n);
    int64_t r = -1;
    for (uint64_t k = 0; k < n; k++)
        if (ord[k] == node) r = (int64_t)k;
    free(ord);
    return r;
}

int64_t oracle_locate(void *h, double x, double y, int64_t *lv_span_lo, int64_t *lv_span_hi,
                      uint32_t *cands, uint32_t cap, uint32_t *ncand)
{
    octx *c = (octx *)h;
    uint32_t lv;
    uint32_t leaf = descend(c, x, y, &lv);
    cbuf cb = {0, 0, 0};
    if (c->cfg.complete_mode) {
        if (c->nodes[0].type == T_BRANCH) lookup(c, x, y, 0, leaf, &cb);
    } else if (lv != leaf) {
        lookup(c, x, y, lv, leaf, &cb);
    }
    *ncand = cb.n;
    for (uint32_t k = 0; k < cb.n && k < cap; k++) cands[k] = (uint32_t)leaf_rank(c, cb.v[k]);
    free(cb.v);
    uint32_t *ord = (uint32_t *)malloc((size_t)c->nnodes * sizeof(uint32_t));
    uint64_t n = 0;
    leaf_order(c, lv, ord, &n);
    *lv_span_lo = leaf_rank(c, ord[0]);
    *lv_span_hi = *lv_span_lo + (int64_t)n;
    free(ord);
    return leaf_rank(c, leaf);
}

int oracle_update_schema_counts(void *h, const uint32_t *counts, uint64_t nleaves)
{
    octx *c = (octx *)h;
    uint32_t *ord = (uint32_t *)malloc((size_t)c->nnodes * sizeof(uint32_t));
    uint64_t n = 0;
    leaf_order(c, 0, ord, &n);
    if (n != nleaves) {
        free(ord);
        return -1;
    }
    memset(c->count, 0, (size_t)c->cap_nodes * sizeof(uint32_t));
    for (uint64_t k = 0; k < n; k++) c->count[ord[k]] = counts[k];
    free(ord);
    update_schema(c);
    return 0;
}

/* --------------------------------------------------- key-value output */

int oracle_kv_enable(void *h, const oracle_kv_config *kc)
{
    octx *c = (octx *)h;
    if (!kc || kc->split_num == 0 || kc->split_num * kc->split_num + 2 > 4096 || kc->max_meet_time == 0) return -1;
    if (c->built || c->kv) return -1;
    const double map[4] = {c->cfg.low_x, c->cfg.low_y, c->cfg.high_x, c->cfg.high_y};
    /* first end_time_min = start_time + min_meet_time (process.cpp:404) */
    c->kv = okv_create(c->cfg.num_objects, map, c->cfg.start_time + c->cfg.min_meet_time, kc);
    return 0;
}

uint64_t oracle_kv_pending(void *h) { return ((octx *)h)->kv ? okv_pending(((octx *)h)->kv) : 0; }

int oracle_kv_flush(void *h, oracle_kv_summary *out)
{
    octx *c = (octx *)h;
    if (!c->kv) return -1;
    int rc = okv_flush(c->kv, c->last_tick);
    if (out) *out = *okv_summary(c->kv);
    return rc;
}

const uint64_t *oracle_kv_keys(void *h, int part) { return okv_keys(((octx *)h)->kv, part); }
const float *oracle_kv_boxes(void *h, int part) { return okv_boxes(((octx *)h)->kv, part); }
const uint16_t *oracle_kv_sids(void *h) { return okv_sids(((octx *)h)->kv); }
const oracle_ctf *oracle_kv_ctfs(void *h) { return okv_ctfs(((octx *)h)->kv); }
const uint8_t *oracle_kv_bitmaps(void *h) { return okv_bitmaps(((octx *)h)->kv); }
const uint8_t *oracle_kv_obitmap(void *h) { return okv_obitmap(((octx *)h)->kv); }
const uint8_t *oracle_kv_packed(void *h) { return okv_packed(((octx *)h)->kv); }
