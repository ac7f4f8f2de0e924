// st_device.hpp -- device-only helpers shared by the HIP translation units.
#pragma once

#include "st_internal.hpp"

namespace st {

// 1.5 * 2^52: for |q| < 2^51, (q + kRoundMagic) - kRoundMagic is rint(q)
// (round half to even, the default rounding mode) and the low word of
// q + kRoundMagic holds rint(q) as a two's complement integer.
constexpr double kRoundMagic = 6755399441055744.0;

__device__ __forceinline__ long long rint_to_i64(double q)
{
    const double t = q + kRoundMagic;
    return (long long)(__double_as_longlong(t) - __double_as_longlong(kRoundMagic));
}

// Bin of v with upper_bound semantics over edges[0..nbins] (the bin b with
// edges[b] <= v < edges[b + 1]), or -1 outside [edges[0], edges[nbins]).
// The closed-form guess uses the hardware float log2; the walk from the
// guess compares against the exact FP64 edges, so the guess only sets the
// number of steps.
__device__ __forceinline__ int bin_fast(const double *__restrict__ edges, const BinGuess &g, double v)
{
    if (!(v >= g.e0 && v < g.en)) return -1;
    float t;
    if (g.type == ST_HIST_LIN) t = (float)((v - g.origin) * g.inv_width);
    else t = (__log2f((float)v) - g.log2_origin) * g.log2_inv_width;
    int b = (t > 0.0f) ? (int)fminf(t, (float)(g.nbins - 1)) : 0;
    // v in [e0, en): the walk never leaves [0, nbins - 1]
#pragma unroll 1
    for (int step = 0; step < 4; ++step) {
        if (v < edges[b]) --b;
        else if (!(v < edges[b + 1])) ++b;
        else return b;
    }
    int lo = 0, hi = g.nbins;   // first i with v < edges[i]
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (v < edges[mid]) hi = mid; else lo = mid + 1;
    }
    return lo - 1;
}

}  // namespace st
