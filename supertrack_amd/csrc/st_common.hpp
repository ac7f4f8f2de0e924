// st_common.hpp -- definitions shared by the host and device code of
// libsupertrack_amd: status handling, the Philox4x32-10 block function, the
// shift mapping and the exact bin search.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/supertrack_amd.h"

namespace st {

// ---- status handling (host) -------------------------------------------------

int set_error(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

#define ST_HIP_TRY(expr)                                                               \
    do {                                                                               \
        hipError_t st_hip_err_ = (expr);                                               \
        if (st_hip_err_ != hipSuccess)                                                 \
            return ::st::set_error(ST_EHIP, "%s failed: %s (%s:%d)", #expr,            \
                                   hipGetErrorString(st_hip_err_), __FILE__, __LINE__); \
    } while (0)

// Stream used by objects that were not given one: one per device, created on
// first use and kept for the life of the process.
hipStream_t default_stream();

// Every device and pinned-host allocation of the library goes through these,
// so that st_debug_fail_allocation can make the n-th one fail (with
// hipErrorOutOfMemory, as a full device would).
hipError_t dev_malloc(void **p, size_t bytes);
hipError_t pinned_malloc(void **p, size_t bytes, unsigned int flags);

// Status for a failed HIP call: ST_ENOMEM for an allocation failure.
inline int hip_status(hipError_t e)
{
    return (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? ST_ENOMEM : ST_EHIP;
}

// Returns ST_ENOMEM (with the failing call in the message) if expr fails.
#define ST_ALLOC_TRY(expr)                                                              \
    do {                                                                                \
        hipError_t st_alloc_err_ = (expr);                                              \
        if (st_alloc_err_ != hipSuccess)                                                \
            return ::st::set_error(ST_ENOMEM, "%s failed: %s (%s:%d)", #expr,            \
                                   hipGetErrorString(st_alloc_err_), __FILE__, __LINE__); \
    } while (0)

// ---- Philox4x32-10 (Salmon et al., SC'11) ------------------------------------

struct u32x4 {
    uint32_t v[4];
};

__host__ __device__ inline u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.v[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.v[2];
        u32x4 n;
        n.v[0] = (uint32_t)(p1 >> 32) ^ c.v[1] ^ k0;
        n.v[1] = (uint32_t)p1;
        n.v[2] = (uint32_t)(p0 >> 32) ^ c.v[3] ^ k1;
        n.v[3] = (uint32_t)p0;
        c = n;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

// Counter word 2 of the shift stream ("SHFT").
constexpr uint32_t kShiftStreamTag = 0x53484654u;

// (0, 1] float from one Philox word: ((r >> 8) + 1) * 2^-24, exact in float.
__host__ __device__ inline float u01_open_closed(uint32_t r)
{
    return (float)((r >> 8) + 1u) * (1.0f / 16777216.0f);
}

// Shift of oversample `os` of track `track_index`: the reference maps its
// uniform u to ((double)u * L) - H' (VoxelConstrainedSphereMethod.cu:176-177)
// and uses u = 0.5 when ShiftTrack is false (:125-136, giving exactly 0).
// The z component is the additive ShiftTrackZ extension, 0.0 when off.
struct Shift {
    double x, y, z;
};

__host__ __device__ inline Shift shift_for(uint64_t seed, uint32_t track_index, uint32_t os,
                                           int shift_xy, int shift_z, double length,
                                           double half_length)
{
    u32x4 c = {{os, track_index, kShiftStreamTag, 0u}};
    u32x4 r = philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const float ux = shift_xy ? u01_open_closed(r.v[0]) : 0.5f;
    const float uy = shift_xy ? u01_open_closed(r.v[1]) : 0.5f;
    Shift s;
    s.x = ((double)ux * length) - half_length;
    s.y = ((double)uy * length) - half_length;
    s.z = shift_z ? (((double)u01_open_closed(r.v[2]) * length) - half_length) : 0.0;
    return s;
}

// ---- exact bin search ----------------------------------------------------------
// cub::DeviceHistogram::HistogramRange semantics (Histogram.cu:158): a value
// counts iff edges[0] <= v < edges[nbins]; its bin is upper_bound - 1.  The
// guess g (any int) only speeds the search up; the result does not depend on it.
__host__ __device__ inline int bin_of(const double *__restrict__ edges, int nbins, double v, int g)
{
    if (!(v >= edges[0] && v < edges[nbins])) return -1;
    if (g < 0) g = 0;
    if (g > nbins - 1) g = nbins - 1;
    // Walk at most a few steps from the guess, then fall back to bisection.
    for (int step = 0; step < 4; ++step) {
        if (v < edges[g]) {
            --g;
        } else if (!(v < edges[g + 1])) {
            ++g;
        } else {
            // edges[g] <= v < edges[g+1]; upper_bound semantics also need the
            // last such g when edges repeat, which the strict test gives.
            return g;
        }
        if (g < 0 || g > nbins - 1) break;
    }
    int lo = 0, hi = nbins;  // first index i in [0, nbins] with v < edges[i]
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (v < edges[mid]) hi = mid; else lo = mid + 1;
    }
    return lo - 1;
}

}  // namespace st
