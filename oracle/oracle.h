/*
 * oracle.h -- FP64 CPU restatement of SuperTrack's VoxelConstrainedSphere
 * superposition path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * implementation in supertrack_amd/ is compared against.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product never links, imports or calls it.
 *
 * Every function follows the reference source cited next to it (paths are
 * relative to the SuperTrack repository root).  Floating point is compiled
 * with -ffp-contract=off and no fast-math, so each expression rounds exactly
 * as written, in the order written.
 *
 * Deviations from the reference that are fixed by specification rather than
 * by the reference's code:
 *   - Random shifts come from Philox4x32-10 (Salmon et al., SC'11), not from
 *     cuRAND's XORWOW stream: counter = {oversample, track_index, 0x53484654,
 *     0}, key = {seed low 32 bits, seed high 32 bits}; word 0 -> x, word 1 ->
 *     y, word 2 -> z.  A word r maps to the float u = ((r >> 8) + 1) * 2^-24,
 *     which lies in (0, 1] like curandGenerateUniform's output.
 *   - The optional z shift (shift_z) is an additive extension; with it off the
 *     z shift is exactly 0.0 and the path reduces to the reference.
 *   - Per-sphere sums use the canonical order: ascending point index, starting
 *     from the first value (the reference's order is set by atomics and is not
 *     reproducible run to run).
 */
#ifndef SUPERTRACK_ORACLE_H
#define SUPERTRACK_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    double half_length;   /* H' (after snapping to an integer number of spheres) */
    double length;        /* L = N * D                                             */
    double diameter;      /* D                                                     */
    double radius;        /* r = D / 2                                             */
    double gso;           /* greatest sphere offset = -H' + r                      */
    int32_t n_linear;     /* N                                                     */
    int32_t pad;
} orc_geometry;

/* SphericalGeometry::SphericalGeometry (SphericalGeometry.cu:34-58). */
void orc_geometry_make(double half_length_ini, double diameter_ini, orc_geometry *g);

/* Philox4x32-10 block function (Salmon et al. 2011, Random123 philox4x32). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Uniform float in (0,1] from one Philox word. */
float orc_u01_open_closed(uint32_t r);

/* Shift table for one track: out[3*k+0..2] = (sx, sy, sz) of oversample k.
 * VoxelConstrainedSphereMethod.cu:61-83 (generation), :176-177 (mapping). */
void orc_shift_table(uint64_t seed, uint32_t track_index, int32_t n_oversamples,
                     int32_t shift_xy, int32_t shift_z, const orc_geometry *g,
                     double *out);

/* utils::LinSpace / utils::LogSpace with the float-typed limits that
 * Histogram reads through INIReader::GetFloat (utils.cc:8-31,
 * Histogram.cu:33-36,53-72).  is_log: 0 = "lin", 1 = "log".  edges has
 * nbins + 1 entries. */
void orc_edges(int32_t is_log, float lo, float hi, int32_t nbins, double *edges);

/* cub::DeviceHistogram::HistogramRange bin selection (Histogram.cu:158):
 * returns upper_bound(edges, v) - 1 when edges[0] <= v < edges[nbins],
 * otherwise -1 (NaN is never counted). */
int32_t orc_bin(const double *edges, int32_t nbins, double v);

/* 1 if an edge adjacent to v (positions lower_bound(v) - 1 and lower_bound(v))
 * lies within rtol * |v| of it, else 0 (NaN: 0).  Not a reference function:
 * the allowance when a histogram filled from sums with relative error below
 * rtol is compared with one filled from the oracle's sums. */
int32_t orc_near_edge(const double *edges, int32_t nbins, double v, double rtol);

/* ShiftTrackZOrigin (VoxelConstrainedSphereMethod.cu:138-154): z[i] += H'. */
void orc_track_zshift(double *z, int64_t n, double half_length);

/* One oversample of one track (VoxelConstrainedSphereMethod.cu:156-324) given
 * the z-shifted track (z' = z + H'):
 *   pair_keys/pair_y (capacity n, may be NULL): the (sphere key, lineal energy)
 *     pairs of every in-sphere point, in ascending point index;
 *   keys/sums (capacity n, may be NULL): the unique keys ascending with their
 *     canonical sums (Histogram.cu:150-153).
 * Returns the number of unique keys; *n_pairs receives the pair count. */
int64_t orc_score_oversample(const double *x, const double *y, const double *zp,
                             const double *e, int64_t n, const orc_geometry *g,
                             double sx, double sy, double sz,
                             uint64_t *pair_keys, double *pair_y, int64_t *n_pairs,
                             uint64_t *keys, double *sums);

/* Histogram::SortReduceAndAddToHistogram's reduce step alone
 * (Histogram.cu:150-153): unique keys ascending with canonical sums of the
 * given pairs.  Returns the number of unique keys. */
int64_t orc_reduce_pairs(const uint64_t *pair_keys, const double *pair_y, int64_t m,
                         uint64_t *keys, double *sums);

/* Whole-run driver (SuperTrackManager.cu:58-83 loop shape): every track of a
 * library (SoA arrays + offsets, tracks in their own point order) with
 * n_oversamples oversamples each, OpenMP over (track, oversample) segments,
 * thread-private u64 bins merged at the end.  Track t uses geometry
 * geoms[geom_of_track[t]] (geom_of_track may be NULL: geometry 0) and shift
 * counter track_index = track_index_base + t.  Only tracks with
 * t % track_stride == 0 are processed.  Up to two histograms (lin and/or log)
 * are filled from the same sums.  If near_counts is not NULL, near_counts[h]
 * is increased by the number of sums for which orc_near_edge(histogram h,
 * near_rtol) holds.  Returns the wall seconds of the scoring loop and writes
 * the thread count used to *threads_used. */
double orc_run(const double *x, const double *y, const double *z, const double *e,
               const int64_t *offsets, int64_t n_tracks, int64_t track_stride,
               const orc_geometry *geoms, const int32_t *geom_of_track,
               int32_t n_oversamples, uint64_t seed, uint32_t track_index_base,
               int32_t shift_xy, int32_t shift_z,
               const double *edges0, int32_t nbins0, uint64_t *bins0,
               const double *edges1, int32_t nbins1, uint64_t *bins1,
               int32_t n_threads, int32_t *threads_used,
               double near_rtol, uint64_t *near_counts);

#ifdef __cplusplus
}
#endif
#endif
