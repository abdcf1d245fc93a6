// trace_io.hip - the reference's trace files read on the host, and the
// lon/lat -> planar metres conversion on the GPU.
//
// File layouts (paths relative to the reference repo):
//  - chunk files "<prefix><st>_<N>.tr": 100 ticks of N Points, tick-major,
//    no header (src/tools/datagen.cpp:18-26, src/tracing/process.cpp:128-144);
//  - trace files: u32 objects, u32 ticks, box (low x, low y, high x,
//    high y as doubles), then tick-major Points (process.cpp:146-187).
// The conversion restates util.h:38-39,57-66 (degree_per_meter_latitude,
// degree_per_meter_longitude with PI = 3.14159265) at a fixed latitude, the
// box's low one, as the filter's x_buffer is (filter.cpp:30-31).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <string>

#include "glint.h"

namespace {

constexpr double kPi = 3.14159265;                  // util.h:38
constexpr double kEarthMetres = 40076.0 * 1000.0;   // util.h:57,65
constexpr uint64_t kPoint = 16;                     // sizeof(Point): two doubles

struct File {
    FILE *f = nullptr;
    explicit File(const char *path) : f(path ? std::fopen(path, "rb") : nullptr) {}
    ~File()
    {
        if (f) std::fclose(f);
    }
    bool read(void *dst, uint64_t bytes) { return bytes == 0 || std::fread(dst, 1, bytes, f) == bytes; }
    bool skip(uint64_t bytes) { return bytes == 0 || fseeko(f, (off_t)bytes, SEEK_CUR) == 0; }
};

bool read_header(File &in, glint_tr_header &h)
{
    double box[4];
    if (!in.read(&h.num_objects, 4) || !in.read(&h.duration, 4) || !in.read(box, sizeof(box))) return false;
    h.low_x = box[0];
    h.low_y = box[1];
    h.high_x = box[2];
    h.high_y = box[3];
    return true;
}

__global__ void k_project(const double *in, double *out, uint64_t n, double lon0, double lat0, double dlon,
                          double dlat)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const double lon = in[2 * i], lat = in[2 * i + 1];
        out[2 * i] = (lon - lon0) / dlon;
        out[2 * i + 1] = (lat - lat0) / dlat;
    }
}

}  // namespace

extern "C" {

glint_status glint_tr_header_read(const char *path, glint_tr_header *out)
{
    if (!path || !out) return GLINT_ERR_INVALID_ARGUMENT;
    File in(path);
    if (!in.f || !read_header(in, *out)) return GLINT_ERR_BAD_INPUT;
    return GLINT_OK;
}

glint_status glint_tr_read(const char *path, uint32_t num_objects, uint32_t st, uint32_t duration, double *out)
{
    if (!path || (!out && num_objects && duration)) return GLINT_ERR_INVALID_ARGUMENT;
    File in(path);
    glint_tr_header h;
    if (!in.f || !read_header(in, h)) return GLINT_ERR_BAD_INPUT;
    // the reference asserts both (process.cpp:163-165); its st * objects is a
    // 32-bit product, exact here
    if ((uint64_t)num_objects * ((uint64_t)st + duration) > (uint64_t)h.num_objects * h.duration ||
        (uint64_t)st + duration > h.duration || (h.num_objects == 0 && num_objects > 0))
        return GLINT_ERR_INVALID_ARGUMENT;
    if (!in.skip((uint64_t)st * h.num_objects * kPoint)) return GLINT_ERR_BAD_INPUT;
    // more objects than the file holds: the next ones come from its later ticks
    uint64_t loaded = 0;
    while (loaded < num_objects) {
        const uint64_t cur = std::min<uint64_t>(h.num_objects, num_objects - loaded);
        for (uint32_t i = 0; i < duration; i++) {
            if (!in.read(out + 2 * ((uint64_t)i * num_objects + loaded), cur * kPoint)) return GLINT_ERR_BAD_INPUT;
            if (h.num_objects > cur && !in.skip((h.num_objects - cur) * kPoint)) return GLINT_ERR_BAD_INPUT;
        }
        loaded += cur;
    }
    return GLINT_OK;
}

glint_status glint_tr_read_chunk(const char *prefix, uint32_t st, uint32_t num_objects, double *out)
{
    if (!prefix || (!out && num_objects)) return GLINT_ERR_INVALID_ARGUMENT;
    const std::string name = std::string(prefix) + std::to_string(st) + "_" + std::to_string(num_objects) + ".tr";
    File in(name.c_str());
    if (!in.f || !in.read(out, 100ull * num_objects * kPoint)) return GLINT_ERR_BAD_INPUT;
    return GLINT_OK;
}

glint_status glint_projection_for_box(double lon_lo, double lat_lo, glint_projection *out)
{
    if (!out || !std::isfinite(lon_lo) || !std::isfinite(lat_lo)) return GLINT_ERR_INVALID_ARGUMENT;
    double absla = std::fabs(lat_lo);
    if (absla == 90) absla = 89;   // util.h:61-63
    if (!(absla < 90)) return GLINT_ERR_INVALID_ARGUMENT;
    out->lon0 = lon_lo;
    out->lat0 = lat_lo;
    out->deg_per_m_lon = 360.0 / (std::sin((90 - absla) * kPi / 180) * 40076.0 * 1000.0);
    out->deg_per_m_lat = 360.0 / kEarthMetres;
    return GLINT_OK;
}

glint_status glint_project_device(const glint_projection *proj, const double *d_lonlat, double *d_xy, uint64_t n,
                                  uint32_t device)
{
    if (!proj || (n && (!d_lonlat || !d_xy))) return GLINT_ERR_INVALID_ARGUMENT;
    if (n == 0) return GLINT_OK;
    if (hipSetDevice((int)device) != hipSuccess) return GLINT_ERR_HIP;
    const uint64_t blocks = std::min<uint64_t>((n + 255) / 256, 8192);
    k_project<<<(uint32_t)blocks, 256>>>(d_lonlat, d_xy, n, proj->lon0, proj->lat0, proj->deg_per_m_lon,
                                         proj->deg_per_m_lat);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) return GLINT_ERR_HIP;
    return GLINT_OK;
}

}  // extern "C"
