/*
 * glint_gen.h - seeded synthetic trace generator for the contact-tracing path.
 *
 * No counterpart in the reference's hot path: the reference's generator
 * (src/tracing/generator.cpp:387-406, trace_generator::generate_trace) walks a
 * Chicago street map with an unseeded multi-threaded rand() (srand(time(NULL))
 * at generator.cpp:388), so its traces cannot be reproduced.  This generator
 * replaces it with a counter-based one: every random draw is a pure function of
 * (seed, object id, tick, stream), so the host stepper below and the device
 * stepper (glint_gen_walk_device in glint.h) produce bit-identical positions.
 *
 * Model (Manhattan street grid):
 *   - box [0, side]^2 metres, streets on x = k*gap and y = k*gap (side is a
 *     multiple of gap, so the box edges are streets);
 *   - an object starts uniform in the box or in one of num_hotspots Gaussian
 *     hotspots (centres uniform, weights proportional to k^-zipf_s), is clamped
 *     into the box and snapped to the nearest street point;
 *   - speed U[speed_min, speed_max) m/s, constant; speed_max < gap;
 *   - each tick (1 s) the object moves `speed` metres along its street; on
 *     reaching an intersection it picks a new direction uniformly among the
 *     directions that stay inside the box (this is the box-edge reflection) and
 *     spends the rest of its step on the new street;
 *   - with move_fraction < 1 only a seeded random subset moves each tick.
 * The step uses only IEEE +, -, compares and integer*gap products, which are
 * exact or correctly rounded on every IEEE machine.
 */
#ifndef GLINT_GEN_H
#define GLINT_GEN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct glint_gen_config {
    uint64_t seed;
    uint32_t num_objects;
    uint32_t num_hotspots;      /* 0: every object starts uniform */
    double side;                /* box side, metres; multiple of street_gap */
    double street_gap;          /* distance between parallel streets, metres */
    double speed_min;           /* m/s */
    double speed_max;           /* m/s, must be < street_gap */
    double hotspot_fraction;    /* share of objects that start in hotspots */
    double hotspot_sigma;       /* Gaussian sigma of a hotspot, metres */
    double zipf_s;              /* hotspot weight exponent */
    double move_fraction;       /* share of objects that move per tick; >= 1: all */
} glint_gen_config;

/* Walker state, structure of arrays, N entries each (xy has 2N, AoS x,y). */
typedef struct glint_walkers {
    double *xy;       /* positions, AoS: xy[2*i] = x, xy[2*i+1] = y */
    double *speed;    /* metres per tick */
    int32_t *tx;      /* x index of the street/target intersection */
    int32_t *ty;      /* y index of the street/target intersection */
    uint8_t *dir;     /* 0:+x 1:-x 2:+y 3:-y */
} glint_walkers;

/* Returns 0 on success, negative on an invalid configuration. */
int glint_gen_validate(const glint_gen_config *cfg);

/* Positions at tick 0 (and the walk state) for objects [first, first+count). */
int glint_gen_init(const glint_gen_config *cfg, glint_walkers *w,
                   uint32_t first, uint32_t count);

/* Advance objects [first, first+count) from tick-1 to tick (tick >= 1). */
int glint_gen_step(const glint_gen_config *cfg, glint_walkers *w, uint32_t tick,
                   uint32_t first, uint32_t count);

/* Counter-based draws shared with the device stepper. */
uint64_t glint_gen_hash(uint64_t seed, uint64_t a, uint64_t b, uint64_t c);

#ifdef __cplusplus
}
#endif

#endif /* GLINT_GEN_H */
