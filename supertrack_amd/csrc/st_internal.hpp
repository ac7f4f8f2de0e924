// st_internal.hpp -- internal declarations shared by the translation units of
// libsupertrack_amd.
#pragma once

#include <functional>
#include <vector>

#include "st_common.hpp"
#include "parallel.hpp"

namespace st {

int device_cu_count();
void make_geometry(double half_length_ini, double diameter_ini, st_geometry *g);
void make_edges(int type, float lo, float hi, int nbins, double *edges);

// Parameters of the closed-form bin guess handed to kernels with the edges.
struct BinGuess {
    int type;        // ST_HIST_LIN / ST_HIST_LOG
    int nbins;
    double origin;   // lin: edges[0]; log: float log10(lo) widened
    double inv_width;
    float log2_origin;      // log: origin * log2(10)
    float log2_inv_width;   // log: inv_width / log2(10)
    double e0, en;          // first and last edge
};

__host__ __device__ inline int bin_guess(const BinGuess &g, double v)
{
    double t;
    if (g.type == ST_HIST_LIN) {
        t = (v - g.origin) * g.inv_width;
    } else {
        t = (v > 0.0) ? (log10(v) - g.origin) * g.inv_width : -1.0;
    }
    if (!(t > -1.0)) return 0;            // also catches NaN
    if (t > 2147483000.0) return g.nbins - 1;
    return (int)t;
}

}  // namespace st

struct st_hist;
struct st_vcs;
namespace st {
// Sorts, reduces and bins pairs into h's device counters on `stream`.
int hist_add_pairs(st_hist *h, const st_pairs *pairs, hipStream_t stream);
}  // namespace st

// Histogram object (Histogram.cuh:19-78): bin edges, device counters and the
// per-track sort buffers.
struct st_hist {
    int type = ST_HIST_LIN;
    int nbins = 0;
    float lo = 0, hi = 0;
    std::vector<double> edges_host;
    st::BinGuess guess{};
    double *edges_dev = nullptr;     // nbins + 1
    uint64_t *bins_dev = nullptr;    // nbins counters
    hipStream_t stream = nullptr;
    int64_t capacity = 0;            // pairs capacity of the current track
    bool in_track = false;
    // onesweep sort buffers (pairs capacities above the LDS path's limit)
    int64_t sort_capacity = 0;
    uint64_t *keys_a = nullptr, *keys_b = nullptr;
    uint32_t *vals_a = nullptr, *vals_b = nullptr;
    unsigned long long *os_status = nullptr;   // look-back words [tiles][256 digits]
    unsigned int *digit_counts = nullptr;      // [8 passes][256 digits]
    unsigned long long *os_ticket = nullptr;
    unsigned long long os_ticket_base = 0;
    uint32_t os_epoch = 0;
    std::vector<uint64_t> host_counts;
    // methods with queued oversamples whose pairs this histogram consumed
    std::vector<st_vcs *> producers;
    // the device counters are zero once the work queued on `stream` is done
    // (nothing has been counted since they were last cleared)
    bool bins_zero = true;
    // tracks were ended without a host copy: the counters carry their counts
    // into the next track (st_hist_end_track with host_bins NULL)
    bool carry = false;
};

// VoxelConstrainedSphereMethod object (VoxelConstrainedSphereMethod.hh:43-75).
struct st_vcs {
    st_geometry g{};
    int shift_xy = 1, shift_z = 0;
    int suggested_blocks = 0, suggested_threads = 0;
    hipStream_t stream = nullptr;
    // per-track state (AllocateTrackProcess .. FreeTrackProcess)
    bool in_track = false;
    int64_t n = 0;
    int32_t n_oversamples = 0;
    int32_t oversample = 0;
    int64_t lineal_capacity = 0;
    double *lineal = nullptr;            // e / ((2/3) D) per point, computed once per track
    double *shifts = nullptr;            // 3 * n_oversamples
    int32_t shifts_capacity = 0;
    // decoupled look-back state of the fused scoring kernel
    unsigned long long *status = nullptr;
    int64_t status_capacity = 0;
    unsigned long long *ticket = nullptr;
    uint64_t ticket_base = 0;
    uint32_t epoch = 0;
    // queued oversamples (st_vcs_set_queued): the oversamples of a track are
    // recorded and run as one launch when the track ends
    bool queue_mode = false;     // asked for by the caller
    bool queue_track = false;    // the current track is queued
    bool queue_active = false;   // oversamples are pending
    st_pairs queue_pairs{};      // the caller's pairs of the pending oversamples
    int32_t queue_os0 = 0, queue_count = 0;
    std::vector<std::vector<st_hist *>> queue_consumers;   // per pending oversample
    const double *queue_x = nullptr, *queue_yc = nullptr, *queue_z = nullptr;   // the track's device arrays
    uint64_t *queue_key = nullptr;   // scratch rows of n pairs, one per pending oversample
    double *queue_y = nullptr;
    int64_t queue_key_capacity = 0, queue_y_capacity = 0;
    hipEvent_t queue_event = nullptr;
};

namespace st {
// Runs the pending queued oversamples of m (no-op if none).
int queue_flush(st_vcs *m);
}  // namespace st
