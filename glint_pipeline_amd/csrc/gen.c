/*
 * Seeded Manhattan-grid trace generator (host side).
 *
 * The per-tick step (walk_one) is the same sequence of IEEE operations as the
 * device stepper in walk.hip; both are compiled without floating-point
 * contraction, so positions agree bit for bit.  Initial placement uses libm
 * (log, cos, sin, pow) and runs on the host only.
 */
#include "glint_gen.h"

#include <math.h>
#include <stddef.h>

enum {
    STREAM_MOVE = 1,
    STREAM_TURN = 2,
    STREAM_KIND = 10,
    STREAM_POS_X = 11,
    STREAM_POS_Y = 12,
    STREAM_SPOT = 13,
    STREAM_GAUSS_U = 14,
    STREAM_GAUSS_V = 15,
    STREAM_SPEED = 16,
    STREAM_DIR = 17,
    STREAM_CENTRE_X = 18,
    STREAM_CENTRE_Y = 19,
};

/* SplitMix64 output function (Steele, Lea, Flood, OOPSLA 2014). */
static inline uint64_t mix64(uint64_t z)
{
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

uint64_t glint_gen_hash(uint64_t seed, uint64_t a, uint64_t b, uint64_t c)
{
    return mix64(seed ^ mix64(a ^ mix64(b ^ mix64(c))));
}

/* Uniform in [0, 1) with 53 random bits. */
static inline double unit_double(uint64_t h) { return (double)(h >> 11) * 0x1.0p-53; }

int glint_gen_validate(const glint_gen_config *c)
{
    if (!c || c->num_objects == 0) return -1;
    if (!(c->street_gap > 0) || !(c->side >= c->street_gap)) return -2;
    double cells = c->side / c->street_gap;
    if (cells != floor(cells) || cells > 1.0e6) return -3;      /* box edges must be streets */
    if (!(c->speed_min > 0) || !(c->speed_max >= c->speed_min)) return -4;
    if (!(c->speed_max < c->street_gap)) return -5;             /* at most one turn per tick */
    if (c->num_hotspots > 0 && !(c->hotspot_sigma > 0)) return -6;
    if (c->num_hotspots > 4096) return -7;
    return 0;
}

static double clampd(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }

int glint_gen_init(const glint_gen_config *c, glint_walkers *w, uint32_t first, uint32_t count)
{
    int rc = glint_gen_validate(c);
    if (rc) return rc;
    if ((uint64_t)first + count > c->num_objects) return -8;
    const double S = c->side, G = c->street_gap;
    const int32_t ncell = (int32_t)(S / G);

    /* Hotspot centres and the cumulative Zipf weights. */
    double cx[4096], cy[4096], cdf[4096];
    double total = 0.0;
    for (uint32_t k = 0; k < c->num_hotspots; k++) {
        cx[k] = unit_double(glint_gen_hash(c->seed, k, 0, STREAM_CENTRE_X)) * S;
        cy[k] = unit_double(glint_gen_hash(c->seed, k, 0, STREAM_CENTRE_Y)) * S;
        total += pow((double)(k + 1), -c->zipf_s);
        cdf[k] = total;
    }

    for (uint64_t i = first; i < (uint64_t)first + count; i++) {
        const uint64_t s = c->seed;
        double x, y;
        int in_spot = c->num_hotspots > 0 &&
                      unit_double(glint_gen_hash(s, i, 0, STREAM_KIND)) < c->hotspot_fraction;
        if (in_spot) {
            double u = unit_double(glint_gen_hash(s, i, 0, STREAM_SPOT)) * total;
            uint32_t k = 0;
            while (k + 1 < c->num_hotspots && cdf[k] <= u) k++;
            /* Box-Muller; u1 in (0, 1] keeps the log finite. */
            double u1 = ((double)(glint_gen_hash(s, i, 0, STREAM_GAUSS_U) >> 11) + 1.0) * 0x1.0p-53;
            double u2 = unit_double(glint_gen_hash(s, i, 0, STREAM_GAUSS_V));
            double rad = sqrt(-2.0 * log(u1));
            double ang = 6.283185307179586 * u2;
            x = cx[k] + c->hotspot_sigma * rad * cos(ang);
            y = cy[k] + c->hotspot_sigma * rad * sin(ang);
        } else {
            x = unit_double(glint_gen_hash(s, i, 0, STREAM_POS_X)) * S;
            y = unit_double(glint_gen_hash(s, i, 0, STREAM_POS_Y)) * S;
        }
        x = clampd(x, 0.0, S);
        y = clampd(y, 0.0, S);

        /* Snap to the nearest street point. */
        int32_t ix = (int32_t)rint(x / G), iy = (int32_t)rint(y / G);
        if (ix > ncell) ix = ncell;
        if (iy > ncell) iy = ncell;
        double ddx = fabs(x - (double)ix * G), ddy = fabs(y - (double)iy * G);
        int up = (glint_gen_hash(s, i, 0, STREAM_DIR) >> 63) != 0;
        uint8_t dir;
        int32_t tx, ty;
        if (ddx <= ddy) {
            /* vertical street x = ix*G, moving along y */
            x = (double)ix * G;
            int32_t lo = (int32_t)floor(y / G);
            if (lo > ncell) lo = ncell;
            if (up && lo >= ncell) up = 0;
            if (!up && y <= 0.0) up = 1;
            if (up) {
                dir = 2;
                ty = lo + 1;
            } else {
                dir = 3;
                ty = ((double)lo * G < y) ? lo : lo - 1;
            }
            tx = ix;
        } else {
            y = (double)iy * G;
            int32_t lo = (int32_t)floor(x / G);
            if (lo > ncell) lo = ncell;
            if (up && lo >= ncell) up = 0;
            if (!up && x <= 0.0) up = 1;
            if (up) {
                dir = 0;
                tx = lo + 1;
            } else {
                dir = 1;
                tx = ((double)lo * G < x) ? lo : lo - 1;
            }
            ty = iy;
        }
        double v = c->speed_min +
                   unit_double(glint_gen_hash(s, i, 0, STREAM_SPEED)) * (c->speed_max - c->speed_min);
        w->xy[2 * i] = x;
        w->xy[2 * i + 1] = y;
        w->speed[i] = v;
        w->tx[i] = tx;
        w->ty[i] = ty;
        w->dir[i] = dir;
    }
    return 0;
}

/* One tick of one walker.  Mirrored operation for operation by walk.hip. */
static void walk_one(const glint_gen_config *c, glint_walkers *w, uint32_t tick, uint64_t i,
                     int32_t ncell)
{
    if (c->move_fraction < 1.0 &&
        !(unit_double(glint_gen_hash(c->seed, i, tick, STREAM_MOVE)) < c->move_fraction))
        return;
    const double G = c->street_gap;
    double x = w->xy[2 * i], y = w->xy[2 * i + 1];
    double rem = w->speed[i];
    int32_t tx = w->tx[i], ty = w->ty[i];
    uint8_t dir = w->dir[i];

    double pos = (dir < 2) ? x : y;
    double target = (double)((dir < 2) ? tx : ty) * G;
    double d = (dir & 1) ? pos - target : target - pos;
    if (rem < d) {
        pos = (dir & 1) ? pos - rem : pos + rem;
        if (dir < 2) x = pos; else y = pos;
    } else {
        rem = rem - d;
        int32_t ix = tx, iy = ty;
        x = (double)ix * G;
        y = (double)iy * G;
        uint8_t valid[4];
        int nv = 0;
        if (ix < ncell) valid[nv++] = 0;
        if (ix > 0) valid[nv++] = 1;
        if (iy < ncell) valid[nv++] = 2;
        if (iy > 0) valid[nv++] = 3;
        uint64_t h = glint_gen_hash(c->seed, i, tick, STREAM_TURN);
        dir = valid[h % (uint64_t)nv];
        switch (dir) {
        case 0: tx = ix + 1; ty = iy; x = x + rem; break;
        case 1: tx = ix - 1; ty = iy; x = x - rem; break;
        case 2: ty = iy + 1; tx = ix; y = y + rem; break;
        default: ty = iy - 1; tx = ix; y = y - rem; break;
        }
    }
    w->xy[2 * i] = x;
    w->xy[2 * i + 1] = y;
    w->tx[i] = tx;
    w->ty[i] = ty;
    w->dir[i] = dir;
}

int glint_gen_step(const glint_gen_config *c, glint_walkers *w, uint32_t tick, uint32_t first,
                   uint32_t count)
{
    int rc = glint_gen_validate(c);
    if (rc) return rc;
    if ((uint64_t)first + count > c->num_objects || tick == 0) return -8;
    const int32_t ncell = (int32_t)(c->side / c->street_gap);
#pragma omp parallel for schedule(static)
    for (int64_t i = first; i < (int64_t)first + count; i++) walk_one(c, w, tick, (uint64_t)i, ncell);
    return 0;
}
