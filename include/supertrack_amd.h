/*
 * supertrack_amd.h -- C ABI of the MI355X-native VoxelConstrainedSphere
 * superposition path.
 *
 * One shared library, libsupertrack_amd.so, exports everything declared here.
 * Signatures use plain pointers and sizes only.  Every call returns ST_OK (0)
 * or a negative ST_E* status; st_last_error() gives the message of the last
 * failure on the calling thread.  The library never aborts the process.
 *
 * Reference interfaces replaced (paths relative to the SuperTrack root):
 *   st_geometry_make            SphericalGeometry::SphericalGeometry       SphericalGeometry.cu:34-58
 *   st_track_*                  Track                                      Track.cuh:7-23, Track.cu:14-126
 *   st_pairs_*                  VolumeEdepPair                             VolumeEdepPair.cuh:5-16, VolumeEdepPair.cu:6-24
 *   st_vcs_create               VoxelConstrainedSphereMethod ctor/ParseInput  VoxelConstrainedSphereMethod.cu:9-35
 *   st_vcs_begin_track          AllocateTrackProcess (+ShiftTrackZOrigin, GenerateRandomXYShift)  :39-83, :125-154
 *   st_vcs_process_oversample   ProcessTrack (FilterInScoringBox, FilterTrackInSphere, ScoreTrackInSphere) :86-101, :156-324
 *   st_vcs_end_track            FreeTrackProcess                           :104-113
 *   st_vcs_destroy              Free + delete                              :116-119, SuperTrackManager.cu:86
 *   st_hist_create              Histogram(nbins, lo, hi, type) / Histogram(INIReader)  Histogram.cu:16-96
 *   st_hist_begin_track         Histogram::AllocateTrackProcess             Histogram.cu:102-112
 *   st_hist_add                 Histogram::SortReduceAndAddToHistogram      Histogram.cu:114-123, 147-171
 *   st_hist_end_track           Histogram::FreeTrackProcess (AccumulateCPUHistogram)  Histogram.cu:125-141, 177-186
 *   st_hist_edges               the bin edges of the CPU TH1D               Histogram.cu:53-78
 *   st_library_*, st_batch_*    no single reference counterpart: the batched form of the per-track
 *                               loop SuperTrackManager.cu:58-83 over a resident track library
 *
 * Threading: objects are not thread-safe; use one st_vcs + st_hist per host
 * thread (the reference uses one per forked worker).  Device work of an
 * object is queued on its stream (st_*_set_stream; by default one stream
 * created by the library per device) and is asynchronous until
 * st_vcs_end_track / st_hist_end_track / st_batch_sync / st_hist_fetch.
 */
#ifndef SUPERTRACK_AMD_H
#define SUPERTRACK_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ST_OK 0
#define ST_EINVAL (-1)   /* bad argument                                  */
#define ST_ENOMEM (-2)   /* device or host allocation failed              */
#define ST_EHIP (-3)     /* a HIP runtime call failed                     */
#define ST_ESTATE (-4)   /* call out of order (e.g. process before begin) */
#define ST_ERANGE (-5)   /* size beyond a documented limit                */

#define ST_HIST_LIN 0
#define ST_HIST_LOG 1

#define ST_F64 0
#define ST_F32 1

/* Message of the last failed call on this thread ("" if none). */
const char *st_last_error(void);
/* Library version string. */
const char *st_version(void);

/* ---- device memory helpers (for callers without their own HIP runtime) ---- */
int st_device_count(int *count);
int st_malloc(void **ptr, size_t bytes);
int st_free(void *ptr);
int st_memcpy_htod(void *dst, const void *src, size_t bytes);
int st_memcpy_dtoh(void *dst, const void *src, size_t bytes);
int st_memcpy_dtod(void *dst, const void *src, size_t bytes);
int st_memset(void *dst, int value, size_t bytes);
int st_device_synchronize(void);
/* The stream that objects use unless given another (one per device, created
 * by the library on first use, non-blocking). */
int st_default_stream(void **hip_stream);
/* Waits for the work on the library's default streams and destroys them, so
 * that nothing of the library outlives the HIP runtime at process exit.  Call
 * it after destroying every object, before the process exits (the Python
 * binding and the supertrack driver do).  A later call that needs a default
 * stream creates a new one.  No counterpart in the reference (its CUDA
 * objects are left to the driver at exit). */
int st_shutdown(void);
/* Test switch for the allocation-failure paths: with countdown = k > 0 the
 * k-th device or pinned-host allocation that the library makes from now on
 * fails as if the device were full (the call reports ST_ENOMEM and releases
 * what it had allocated); 0 turns the switch off.  No counterpart in the
 * reference. */
int st_debug_fail_allocation(int64_t countdown);

/* ---- SphericalGeometry (SphericalGeometry.cuh:8-22) ---- */
typedef struct {
    double half_length;  /* H' after snapping to N whole spheres         */
    double length;       /* L = N * D                                    */
    double diameter;     /* D                                            */
    double radius;       /* r = D / 2                                    */
    double gso;          /* greatest sphere offset, -H' + r              */
    int32_t n_linear;    /* N                                            */
    int32_t pad;
} st_geometry;

int st_geometry_make(double half_length_ini, double diameter_ini, st_geometry *out);

/* ---- Track (device SoA, caller-owned; Track.cuh:18-21) ---- */
typedef struct {
    double *x, *y, *z, *edep;   /* device pointers, nm / eV */
    int64_t n;                  /* number of points          */
    int64_t capacity;
} st_track;

int st_track_alloc(st_track *t, int64_t capacity);
int st_track_upload(st_track *t, const double *x, const double *y, const double *z,
                    const double *edep, int64_t n);
void st_track_free(st_track *t);

/* ---- VolumeEdepPair (device, caller-owned; VolumeEdepPair.cuh:13-15) ---- */
typedef struct {
    uint64_t *volume;        /* sphere keys, device                        */
    double *edep;            /* lineal energy per pair (keV/um), device    */
    int32_t *num_elements;   /* pair count in pinned, device-mapped host memory
                                (the reference's managed numElements): the
                                host reads it after the producing stream has
                                been synchronised (st_vcs_end_track)      */
    int64_t capacity;
    uint64_t key_bound;      /* largest key the producer can write; 0 = unknown
                                (set by st_vcs_process_oversample; the radix
                                sort for more than 12288 pairs runs only the
                                passes this needs) */
    int32_t *count_dev;      /* device copy of the count, written next to
                                num_elements by the producer and read by the
                                histogram kernels                           */
} st_pairs;

/* Both counts start at the capacity (VolumeEdepPair.cu:10-13). */
int st_pairs_alloc(st_pairs *p, int64_t capacity);
void st_pairs_free(st_pairs *p);
/* Sets both counts, for pairs that the caller filled itself (waits for the
 * device first). */
int st_pairs_set_count(st_pairs *p, int32_t count);

/* ---- VoxelConstrainedSphereMethod (the per-track drop-in) ---- */
typedef struct st_vcs st_vcs;

/* shift_xy: [VoxelConstrainedSphere] ShiftTrack; shift_z: the additive
 * ShiftTrackZ key (0 reproduces the reference).  suggested_blocks/threads are
 * the SuggestedCuda* hints (0: library default). */
int st_vcs_create(const st_geometry *g, int32_t shift_xy, int32_t shift_z,
                  int32_t suggested_blocks, int32_t suggested_threads, st_vcs **out);
int st_vcs_set_stream(st_vcs *m, void *hip_stream);
/* Adds H' to track->z in place (as ShiftTrackZOrigin does) and draws the
 * Philox shift table for n_oversamples oversamples of (seed, track_index). */
int st_vcs_begin_track(st_vcs *m, st_track *track, uint64_t seed, uint32_t track_index,
                       int32_t n_oversamples);
/* Writes the (key, lineal energy) pair of every in-sphere point, in ascending
 * point index, to out and the count to *out->num_elements. */
int st_vcs_process_oversample(st_vcs *m, const st_track *track, st_pairs *out);
int st_vcs_end_track(st_vcs *m);
void st_vcs_destroy(st_vcs *m);
/* Queued oversamples (on != 0; off by default).  The reference's per-track
 * loop reads nothing between its per-oversample calls before the track's
 * synchronisation (SuperTrackManager.cu:71-76), so a queued method records
 * st_vcs_process_oversample calls, histograms record their st_hist_add calls
 * on the same pairs, and st_vcs_end_track runs the whole track in one launch
 * (one workgroup per oversample: score, compaction, sort, reduce, bin).
 * Results equal the per-oversample path's bit for bit, and after the track
 * ends the pairs hold the last oversample's, as they would.  Until then the
 * pairs' contents and counts are undefined.  Pending work also runs when it
 * is needed: st_hist_end_track / st_hist_fetch / st_hist_begin_track /
 * st_hist_set_stream of a consuming histogram, st_pairs_free of the pairs, a
 * new track, another pairs object, or turning the mode off.  Tracks of more
 * than 12288 points (or whose queued pair rows would exceed 1 GiB) run
 * oversample by oversample.  No counterpart in the reference. */
int st_vcs_set_queued(st_vcs *m, int32_t on);

/* ---- Histogram ---- */
typedef struct st_hist st_hist;

/* type: ST_HIST_LIN / ST_HIST_LOG; lo/hi are floats on purpose (the
 * reference parses them with GetFloat). */
int st_hist_create(int32_t type, float lo, float hi, int32_t nbins, st_hist **out);
int st_hist_set_stream(st_hist *h, void *hip_stream);
int st_hist_nbins(const st_hist *h, int32_t *nbins);
int st_hist_edges(const st_hist *h, double *edges /* nbins + 1 */);
int st_hist_begin_track(st_hist *h, const st_pairs *pairs);
int st_hist_add(st_hist *h, const st_pairs *pairs);
/* host_bins has nbins + 2 entries in ROOT order (0 underflow, 1..nbins,
 * nbins+1 overflow); the counts are added to entries 1..nbins and the device
 * counters cleared.  host_bins NULL: no transfer; the track's counts stay
 * summed in the device counters, the next st_hist_begin_track keeps them, and
 * st_hist_fetch (or a later st_hist_end_track with host_bins) reads the
 * total. */
int st_hist_end_track(st_hist *h, double *host_bins);
/* Copies the device counters accumulated by st_batch_run (u64, nbins entries)
 * to counts, optionally zeroing them. */
int st_hist_fetch(st_hist *h, uint64_t *counts, int32_t reset);
void st_hist_destroy(st_hist *h);

/* ---- Batched form: resident track library + one fused persistent pass ---- */
typedef struct st_library st_library;

typedef struct {
    int64_t n_tracks;
    int64_t n_points;
    int64_t device_bytes;        /* HBM held by the library                    */
    int64_t fast_tracks;         /* tracks scored by the fused kernel          */
    int64_t exact_tracks;        /* tracks routed to the per-segment exact path */
    int32_t precision;           /* ST_F64 / ST_F32                            */
    int32_t n_geometries;
} st_library_info;

/* Host SoA arrays in each track's own point order; offsets has n_tracks + 1
 * entries.  geom_of_track may be NULL (all tracks use geoms[0]). */
int st_library_create(const double *x, const double *y, const double *z, const double *edep,
                      const int64_t *offsets, int64_t n_tracks,
                      const st_geometry *geoms, int32_t n_geoms, const int32_t *geom_of_track,
                      int32_t precision, st_library **out);
/* Streamed construction of the same library, for libraries whose host copy
 * would not fit in memory at once: st_library_begin sizes the device arrays
 * for n_tracks / n_points in total; st_library_append lays out and uploads
 * the next tracks in library order (host arrays as for st_library_create,
 * indexed by this call's offsets; geom_of_track may be NULL); after the last
 * append, st_library_finish publishes the track table.  st_batch_create
 * accepts only a finished library.  st_library_create is begin, appends of
 * bounded size, finish.  Appending more tracks or points than begin sized, or
 * finishing with fewer, is ST_EINVAL. */
int st_library_begin(int64_t n_tracks, int64_t n_points, const st_geometry *geoms, int32_t n_geoms,
                     int32_t precision, st_library **out);
int st_library_append(st_library *lib, const double *x, const double *y, const double *z, const double *edep,
                      const int64_t *offsets, int64_t n_tracks, const int32_t *geom_of_track);
int st_library_finish(st_library *lib);
int st_library_info_get(const st_library *lib, st_library_info *info);
void st_library_destroy(st_library *lib);

typedef struct {
    int32_t n_oversamples;
    int32_t shift_xy;
    int32_t shift_z;
    uint32_t track_index_base;    /* shift counter of track t = base + t      */
    uint64_t seed;
    int32_t use_graph;            /* capture the launch sequence once, replay */
    int32_t autotune;             /* pick launch parameters by timing          */
    int32_t block_threads;        /* 0 = default / autotuned                   */
    int32_t table_slots;          /* LDS hash slots per workgroup: 1280 or 1472
                                     (one point per thread), 1536, 2048 or
                                     4096; 0 (with blocks_per_cu and
                                     points_per_thread 0) = 1280 at seven or
                                     1472 at six workgroups per CU (ST_F32)
                                     where they fit, else 1536, or 1280
                                     where that fits more workgroups          */
    int32_t blocks_per_cu;        /* persistent workgroups per CU; 0 = 7 or 6
                                     (ST_F32, with the tables above), else 5
                                     (ST_F32), 4 (ST_F64), capped by
                                     occupancy                                 */
    int32_t points_per_thread;    /* 1 or 2 per iteration; 0 = 1 (ST_F32 at five
                                     workgroups per CU, else 2), 2 (ST_F64)    */
    int32_t hist_copies;          /* LDS copies of the histograms per workgroup
                                     (1, 2 or 4; wave w counts into copy
                                     w % copies); 0 = 1 or autotuned          */
} st_batch_opts;

typedef struct st_batch st_batch;

int st_batch_create(st_library *lib, const st_batch_opts *opts, st_hist *const *hists,
                    int32_t n_hists, st_batch **out);
int st_batch_set_stream(st_batch *b, void *hip_stream);
/* One pass: every track x every oversample; counts are added to the device
 * counters of each histogram (read them with st_hist_fetch). */
int st_batch_run(st_batch *b);
int st_batch_sync(st_batch *b);
/* Device time of the last st_batch_run in ms: ms2[0] the fused kernel
 * (with its reset), ms2[1] the exact per-segment path. */
int st_batch_times(st_batch *b, float *ms2);
/* Launch parameters in use (5 entries): block_threads, table_slots,
 * blocks_per_cu, points_per_thread, hist_copies (0: the histograms count
 * straight into the device counters, too large for LDS). */
int st_batch_params(const st_batch *b, int32_t *params5);
/* Counts of the last st_batch_run's fused kernel: stats3[0] points scored
 * (points of the in-voxel block ranges of the segments that pass the x/y
 * cull), stats3[1] of those whose FP32 decision was within the error bound
 * and was redone from the FP64 side copy (ST_F32), stats3[2] points inside a
 * sphere.  Waits for the run.  No counterpart in the reference: measurement. */
int st_batch_stats(st_batch *b, int64_t *stats3);
/* Conservation totals of the last st_batch_run's fused kernel (waits for the
 * run), all modulo 2^64: balance4[0] the fixed-point energy words inserted
 * into the sphere tables, [1] those of the entries retired from them (binned
 * or outside every histogram), [2] table entries created, [3] entries
 * retired.  A pass that loses no entry has [0] == [1] and [2] == [3];
 * st_batch_sync returns ST_ESTATE otherwise.  [0] and [1] are the same in
 * every pass over the same work; [2] and [3] may vary by a few between
 * passes (an entry kept across a flush can merge into one that a new point
 * made first: one more created and retired).  No counterpart in the
 * reference: a check on the fused kernel. */
int st_batch_balance(st_batch *b, uint64_t *balance4);
/* Test hook: record (segment, key, sum) of every scored sphere of the next
 * st_batch_run into host-visible arrays of the given capacity (segment =
 * track * n_oversamples + oversample).  capacity 0 disables recording. */
int st_batch_record(st_batch *b, int64_t capacity);
int st_batch_record_fetch(st_batch *b, int64_t *segments, uint64_t *keys, double *sums,
                          int64_t *count);
void st_batch_destroy(st_batch *b);

/* Allocation-free pass.  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The same pass as st_batch_create +
 * st_batch_run over a finished library, adding counts to each histogram's
 * device counters, but it allocates nothing: its control block (segment
 * ticket, error flag, run statistics) is a 256-byte device array of the
 * library, one per device, held until the process ends.  Launch shape as
 * st_batch_create without autotune (opts->autotune must be 0; use_graph is
 * ignored).  Limit: the library may hold no track routed to the exact
 * per-segment path (st_library_info_get: exact_tracks == 0), else ST_ERANGE.
 * hip_stream NULL = the library's default stream.  stats7 NULL: the call only
 * queues the pass and returns.  Otherwise it waits for the pass, writes
 * stats7[0..2] as st_batch_stats and [3..6] as st_batch_balance, and returns
 * ST_ESTATE if the kernel reported an error or the conservation totals differ.
 * Calls on one device share that device's control block: calls that overlap
 * in time on one device (other streams or host threads) race and must not be
 * issued. */
int st_batch_run_static(st_library *lib, const st_batch_opts *opts, st_hist *const *hists, int32_t n_hists,
                        void *hip_stream, int64_t *stats7);

/* ---- synthetic track library (deterministic, Philox-driven) ---- */
typedef struct {
    double voxel_half_length;   /* x0, y0 ~ U(-h, h); primary enters at z = -2 * entry_half_length */
    double entry_half_length;
    double length_min_nm;       /* primary length ~ U[min, max]                   */
    double length_max_nm;
    double spacing_nm;          /* mean event spacing (length-limited mode)       */
    double p_cluster;           /* delta-ray cluster probability                  */
    double p_cluster_max;       /* > p_cluster: p ~ U[p_cluster, p_cluster_max]   */
    double cluster_mean;        /* geometric cluster size mean                    */
    double walk_sigma_nm;       /* Gaussian random walk step sigma                */
    double edep_median_ev;
    double edep_sigma_ln;
    double edep_min_ev;
    double edep_max_ev;
    double spacing_min_nm;      /* count mode: spacing clamp                      */
    double spacing_max_nm;
    int64_t n_min;              /* count mode when n_min > 0: n ~ logU[n_min, n_max] */
    int64_t n_max;
    uint32_t stream_id;         /* Philox key word 1                              */
    uint32_t pad;
} st_synth_spec;

/* Points of tracks [t0, t1) (counts[t - t0]). */
int st_synth_counts(const st_synth_spec *spec, int64_t t0, int64_t t1, int64_t *counts);
/* Fills tracks [t0, t1) at offsets[t - t0] - offsets[0] of the output arrays. */
int st_synth_fill(const st_synth_spec *spec, int64_t t0, int64_t t1, const int64_t *offsets,
                  double *x, double *y, double *z, double *edep);

#ifdef __cplusplus
}
#endif
#endif
