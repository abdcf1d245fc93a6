// Device stepper of the seeded Manhattan-grid generator (include/glint_gen.h).
// Operation for operation the same as walk_one() in gen.c; with contraction
// off both produce the same doubles, so a trace generated on the device
// equals the host trace bit for bit.
#include <hip/hip_runtime.h>

#include "glint.h"

namespace {

constexpr uint64_t kStreamMove = 1;
constexpr uint64_t kStreamTurn = 2;

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ uint64_t hash4(uint64_t seed, uint64_t a, uint64_t b, uint64_t c)
{
    return mix64(seed ^ mix64(a ^ mix64(b ^ mix64(c))));
}

__device__ __forceinline__ double unit_double(uint64_t h) { return (double)(h >> 11) * 0x1.0p-53; }

struct WalkParams {
    uint64_t seed;
    uint32_t n;
    int32_t ncell;
    double gap;
    double move_fraction;
};

__global__ void k_walk(WalkParams p, double *xy, const double *speed, int32_t *txa, int32_t *tya, uint8_t *dira,
                       uint32_t tick, double *snapshot, uint32_t *upd_pid, double *upd_xy, uint32_t *upd_count)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < p.n; i += gridDim.x * blockDim.x) {
        double x = xy[2 * (size_t)i], y = xy[2 * (size_t)i + 1];
        bool moves = !(p.move_fraction < 1.0) ||
                     unit_double(hash4(p.seed, i, tick, kStreamMove)) < p.move_fraction;
        if (moves) {
            const double G = p.gap;
            double rem = speed[i];
            int32_t tx = txa[i], ty = tya[i];
            uint8_t dir = dira[i];
            double pos = (dir < 2) ? x : y;
            double target = (double)((dir < 2) ? tx : ty) * G;
            double d = (dir & 1) ? pos - target : target - pos;
            if (rem < d) {
                pos = (dir & 1) ? pos - rem : pos + rem;
                if (dir < 2)
                    x = pos;
                else
                    y = pos;
            } else {
                rem = rem - d;
                int32_t ix = tx, iy = ty;
                x = (double)ix * G;
                y = (double)iy * G;
                uint8_t valid[4];
                int nv = 0;
                if (ix < p.ncell) valid[nv++] = 0;
                if (ix > 0) valid[nv++] = 1;
                if (iy < p.ncell) valid[nv++] = 2;
                if (iy > 0) valid[nv++] = 3;
                uint64_t h = hash4(p.seed, i, tick, kStreamTurn);
                dir = valid[h % (uint64_t)nv];
                switch (dir) {
                case 0: tx = ix + 1; ty = iy; x = x + rem; break;
                case 1: tx = ix - 1; ty = iy; x = x - rem; break;
                case 2: ty = iy + 1; tx = ix; y = y + rem; break;
                default: ty = iy - 1; tx = ix; y = y - rem; break;
                }
            }
            xy[2 * (size_t)i] = x;
            xy[2 * (size_t)i + 1] = y;
            txa[i] = tx;
            tya[i] = ty;
            dira[i] = dir;
        }
        if (upd_count) {
            // the objects that moved, as a location-update list (any order)
            const unsigned long long m = __ballot(moves);
            uint32_t base = 0;
            if (__lane_id() == 0 && m) base = atomicAdd(upd_count, (uint32_t)__popcll(m));
            base = __shfl(base, 0);
            if (moves) {
                const uint32_t j = base + (uint32_t)__popcll(m & ((1ull << __lane_id()) - 1ull));
                upd_pid[j] = i;
                upd_xy[2 * (size_t)j] = x;
                upd_xy[2 * (size_t)j + 1] = y;
            }
        }
        if (snapshot) {
            snapshot[2 * (size_t)i] = x;
            snapshot[2 * (size_t)i + 1] = y;
        }
    }
}

}  // namespace

extern "C" glint_status glint_gen_walk_device(const glint_gen_config *cfg, glint_walkers *w, uint32_t first_tick,
                                              uint32_t nticks, double *d_trace, uint32_t device)
{
    if (!cfg || !w || first_tick == 0) return GLINT_ERR_INVALID_ARGUMENT;
    if (glint_gen_validate(cfg) != 0) return GLINT_ERR_INVALID_ARGUMENT;
    if (hipSetDevice((int)device) != hipSuccess) return GLINT_ERR_HIP;
    WalkParams p;
    p.seed = cfg->seed;
    p.n = cfg->num_objects;
    p.ncell = (int32_t)(cfg->side / cfg->street_gap);
    p.gap = cfg->street_gap;
    p.move_fraction = cfg->move_fraction;
    uint32_t grid = (p.n + 255) / 256;
    if (grid > 16384) grid = 16384;
    for (uint32_t k = 0; k < nticks; k++) {
        double *snap = d_trace ? d_trace + 2ull * p.n * k : nullptr;
        k_walk<<<grid, 256>>>(p, w->xy, w->speed, w->tx, w->ty, w->dir, first_tick + k, snap, nullptr, nullptr,
                              nullptr);
    }
    if (hipGetLastError() != hipSuccess) return GLINT_ERR_HIP;
    if (hipDeviceSynchronize() != hipSuccess) return GLINT_ERR_HIP;
    return GLINT_OK;
}

// One tick of the walk that also writes the objects that moved as an update
// list: d_pids[j], d_xy[2j..2j+1] for j < *d_count (device memory; capacity
// num_objects).
extern "C" glint_status glint_gen_walk_updates_device(const glint_gen_config *cfg, glint_walkers *w, uint32_t tick,
                                                      uint32_t *d_pids, double *d_xy, uint32_t *d_count,
                                                      uint32_t device)
{
    if (!cfg || !w || tick == 0 || !d_pids || !d_xy || !d_count) return GLINT_ERR_INVALID_ARGUMENT;
    if (glint_gen_validate(cfg) != 0) return GLINT_ERR_INVALID_ARGUMENT;
    if (hipSetDevice((int)device) != hipSuccess) return GLINT_ERR_HIP;
    WalkParams p;
    p.seed = cfg->seed;
    p.n = cfg->num_objects;
    p.ncell = (int32_t)(cfg->side / cfg->street_gap);
    p.gap = cfg->street_gap;
    p.move_fraction = cfg->move_fraction;
    uint32_t grid = (p.n + 255) / 256;
    if (grid > 16384) grid = 16384;
    if (hipMemsetAsync(d_count, 0, sizeof(uint32_t), nullptr) != hipSuccess) return GLINT_ERR_HIP;
    k_walk<<<grid, 256>>>(p, w->xy, w->speed, w->tx, w->ty, w->dir, tick, nullptr, d_pids, d_xy, d_count);
    if (hipGetLastError() != hipSuccess) return GLINT_ERR_HIP;
    if (hipDeviceSynchronize() != hipSuccess) return GLINT_ERR_HIP;
    return GLINT_OK;
}
