// runtime.cpp -- status strings, device memory helpers, geometry and bin
// edges of libsupertrack_amd (host code).
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <mutex>
#include <vector>

#include "st_common.hpp"
#include "st_internal.hpp"

namespace st {

namespace {
thread_local std::string g_last_error;
}

int set_error(int code, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

void clear_error() { g_last_error.clear(); }

namespace {
// st_debug_fail_allocation: allocations left before the one that fails (0: off)
std::atomic<int64_t> g_fail_countdown{0};

bool injected_failure()
{
    int64_t c = g_fail_countdown.load();
    while (c > 0) {
        if (g_fail_countdown.compare_exchange_weak(c, c - 1)) return c == 1;
    }
    return false;
}
}  // namespace

hipError_t dev_malloc(void **p, size_t bytes)
{
    if (injected_failure()) {
        *p = nullptr;
        return hipErrorOutOfMemory;
    }
    return hipMalloc(p, bytes);
}

hipError_t pinned_malloc(void **p, size_t bytes, unsigned int flags)
{
    if (injected_failure()) {
        *p = nullptr;
        return hipErrorOutOfMemory;
    }
    return hipHostMalloc(p, bytes, flags);
}

namespace {
// per-device default streams, created on first use, destroyed by st_shutdown
std::mutex g_stream_mu;
std::vector<hipStream_t> g_streams;
}  // namespace

hipStream_t default_stream()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lock(g_stream_mu);
    if ((int)g_streams.size() <= dev) g_streams.resize(dev + 1, nullptr);
    if (!g_streams[dev]) {
        hipStream_t s = nullptr;
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return nullptr;
        g_streams[dev] = s;
    }
    return g_streams[dev];
}

int device_cu_count()
{
    static int cached = 0;
    if (cached) return cached;
    int dev = 0;
    hipDeviceProp_t p;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&p, dev) != hipSuccess) return 256;
    cached = p.multiProcessorCount;
    return cached;
}

// SphericalGeometry::SphericalGeometry (SphericalGeometry.cu:34-58).
void make_geometry(double half_length_ini, double diameter_ini, st_geometry *g)
{
    const double D = diameter_ini;
    const double r = D / 2;
    int N = (int)std::floor((half_length_ini * 2) / D);
    double L = N * D;
    double H = L / 2;
    double gso = -H + r;
    if (r > H) {  // sphere larger than the voxel
        H = r;
        N = 1;
        L = H * 2;
        gso = -H + r;
    }
    g->half_length = H;
    g->length = L;
    g->diameter = D;
    g->radius = r;
    g->gso = gso;
    g->n_linear = N;
    g->pad = 0;
}

// utils::LinSpace / utils::LogSpace (utils.cc:8-31).  The limits arrive as
// floats (Histogram.cu:34-35 reads them with GetFloat) and the width is a
// float quotient widened to double, so edges carry the float truncation.
void make_edges(int type, float lo, float hi, int nbins, double *edges)
{
    if (type == ST_HIST_LIN) {
        const float wf = (hi - lo) / (float)nbins;
        const double w = wf;
        for (int i = 0; i <= nbins; ++i) edges[i] = (double)lo + (double)i * w;
    } else {
        const float bo = (float)std::log10((double)lo);
        const float to = (float)std::log10((double)hi);
        const float wf = (to - bo) / (float)nbins;
        const double w = wf;
        for (int i = 0; i <= nbins; ++i) edges[i] = std::pow(10.0, (double)bo + (double)i * w);
    }
}

}  // namespace st

using namespace st;

extern "C" {

const char *st_last_error(void) { return g_last_error.c_str(); }

const char *st_version(void) { return "supertrack_amd 0.1 (gfx950)"; }

int st_device_count(int *count)
{
    if (!count) return set_error(ST_EINVAL, "st_device_count: null pointer");
    ST_HIP_TRY(hipGetDeviceCount(count));
    return ST_OK;
}

int st_malloc(void **ptr, size_t bytes)
{
    if (!ptr) return set_error(ST_EINVAL, "st_malloc: null pointer");
    *ptr = nullptr;
    if (bytes == 0) return ST_OK;
    hipError_t e = dev_malloc(ptr, bytes);
    if (e != hipSuccess) {
        *ptr = nullptr;
        return set_error(ST_ENOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    return ST_OK;
}

int st_free(void *ptr)
{
    if (ptr) ST_HIP_TRY(hipFree(ptr));
    return ST_OK;
}

// The helpers below are ordered after all device work queued so far on any
// stream (the library's streams are non-blocking, so a plain hipMemcpy on the
// null stream would not wait for them) and complete before returning.
int st_memcpy_htod(void *dst, const void *src, size_t bytes)
{
    if (!bytes) return ST_OK;
    ST_HIP_TRY(hipDeviceSynchronize());
    ST_HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    ST_HIP_TRY(hipDeviceSynchronize());
    return ST_OK;
}

int st_memcpy_dtoh(void *dst, const void *src, size_t bytes)
{
    if (!bytes) return ST_OK;
    ST_HIP_TRY(hipDeviceSynchronize());
    ST_HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return ST_OK;
}

int st_memcpy_dtod(void *dst, const void *src, size_t bytes)
{
    if (!bytes) return ST_OK;
    ST_HIP_TRY(hipDeviceSynchronize());
    ST_HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    ST_HIP_TRY(hipDeviceSynchronize());
    return ST_OK;
}

int st_memset(void *dst, int value, size_t bytes)
{
    if (!bytes) return ST_OK;
    ST_HIP_TRY(hipDeviceSynchronize());
    ST_HIP_TRY(hipMemset(dst, value, bytes));
    ST_HIP_TRY(hipDeviceSynchronize());
    return ST_OK;
}

int st_device_synchronize(void)
{
    ST_HIP_TRY(hipDeviceSynchronize());
    return ST_OK;
}

int st_debug_fail_allocation(int64_t countdown)
{
    if (countdown < 0) return set_error(ST_EINVAL, "st_debug_fail_allocation: countdown < 0");
    g_fail_countdown.store(countdown);
    return ST_OK;
}

int st_shutdown(void)
{
    std::lock_guard<std::mutex> lock(g_stream_mu);
    int cur = 0;
    const bool have_cur = hipGetDevice(&cur) == hipSuccess;
    int rc = ST_OK;
    for (int dev = 0; dev < (int)g_streams.size(); ++dev) {
        if (!g_streams[dev]) continue;
        hipError_t e = hipSetDevice(dev);
        if (e == hipSuccess) e = hipStreamSynchronize(g_streams[dev]);
        if (e == hipSuccess) e = hipStreamDestroy(g_streams[dev]);
        if (e != hipSuccess && rc == ST_OK)
            rc = set_error(ST_EHIP, "st_shutdown: device %d: %s", dev, hipGetErrorString(e));
        g_streams[dev] = nullptr;
    }
    if (have_cur) (void)hipSetDevice(cur);
    return rc;
}

int st_default_stream(void **hip_stream)
{
    if (!hip_stream) return set_error(ST_EINVAL, "st_default_stream: null pointer");
    hipStream_t s = default_stream();
    if (!s) return set_error(ST_EHIP, "st_default_stream: could not create a stream");
    *hip_stream = (void *)s;
    return ST_OK;
}

int st_geometry_make(double half_length_ini, double diameter_ini, st_geometry *out)
{
    if (!out) return set_error(ST_EINVAL, "st_geometry_make: null output");
    if (!(diameter_ini > 0) || !(half_length_ini >= 0) || !std::isfinite(diameter_ini) ||
        !std::isfinite(half_length_ini))
        return set_error(ST_EINVAL, "st_geometry_make: need finite half length >= 0 and diameter > 0 "
                                    "(got %g, %g)", half_length_ini, diameter_ini);
    make_geometry(half_length_ini, diameter_ini, out);
    if ((double)out->n_linear * out->n_linear * (double)out->n_linear > 1.8e19)
        return set_error(ST_ERANGE, "st_geometry_make: %d spheres per axis overflow 64-bit keys",
                         out->n_linear);
    return ST_OK;
}

}  // extern "C"
