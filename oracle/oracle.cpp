// oracle.cpp -- FP64 CPU restatement of the VoxelConstrainedSphere path.
//
// TEST INFRASTRUCTURE ONLY: the checker for supertrack_amd's HIP path (see
// oracle.h for the rules).  Build: oracle/Makefile (g++ -O2 -fopenmp
// -ffp-contract=off).  Every expression below keeps the reference kernel's
// operand order so that it rounds the same way.
#include "oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {

// SphericalGeometry.cu:34-58.  TMath::Floor is std::floor.
void orc_geometry_make(double half_length_ini, double diameter_ini, orc_geometry *g)
{
    double H = half_length_ini;
    double D = diameter_ini;
    double r = D / 2;
    int N = (int)std::floor((H * 2) / D);       // :47
    double L = N * D;                            // :50
    H = L / 2;                                   // :51
    double gso = -H + r;                         // :54
    if (r > H) {                                 // :57-63
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

// Philox4x32-10: Salmon, Moraes, Dror, Shaw, "Parallel Random Numbers: As
// Easy as 1, 2, 3" (SC'11), section 3.3: multipliers 0xD2511F53/0xCD9E8D57,
// Weyl key increments 0x9E3779B9/0xBB67AE85, ten rounds.
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

float orc_u01_open_closed(uint32_t r)
{
    return (float)((r >> 8) + 1u) * (1.0f / 16777216.0f);
}

// VoxelConstrainedSphereMethod.cu:61-83 draws two floats per oversample (or
// sets them to 0.5 when ShiftTrack is false, :125-136); :176-177 maps them to
// ((double)u * L) - H.  The z shift is the additive extension (0.0 when off).
void orc_shift_table(uint64_t seed, uint32_t track_index, int32_t n_oversamples,
                     int32_t shift_xy, int32_t shift_z, const orc_geometry *g,
                     double *out)
{
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (int32_t k = 0; k < n_oversamples; ++k) {
        const uint32_t ctr[4] = {(uint32_t)k, track_index, 0x53484654u, 0u};
        uint32_t r[4];
        orc_philox4x32_10(ctr, key, r);
        float ux = shift_xy ? orc_u01_open_closed(r[0]) : 0.5f;
        float uy = shift_xy ? orc_u01_open_closed(r[1]) : 0.5f;
        out[3 * k + 0] = ((double)ux * g->length) - g->half_length;
        out[3 * k + 1] = ((double)uy * g->length) - g->half_length;
        out[3 * k + 2] = shift_z ? (((double)orc_u01_open_closed(r[2]) * g->length) - g->half_length) : 0.0;
    }
}

// utils.cc:8-31.  bottom/top are floats (INIReader::GetFloat, Histogram.cu:34-35)
// and the width quotient is evaluated in float before widening to double.
void orc_edges(int32_t is_log, float lo, float hi, int32_t nbins, double *edges)
{
    if (!is_log) {
        double width = (hi - lo) / (float)nbins;                 // utils.cc:11
        for (int i = 0; i <= nbins; ++i) edges[i] = lo + (i * width);   // :15
    } else {
        float bottom_order_mag = (float)std::log10((double)lo);  // :21
        float top_order_mag = (float)std::log10((double)hi);     // :22
        double width = (top_order_mag - bottom_order_mag) / (float)nbins;  // :25
        for (int i = 0; i <= nbins; ++i)
            edges[i] = std::pow(10.0, (double)bottom_order_mag + (i * width));  // :29
    }
}

// Histogram.cu:158 (cub::DeviceHistogram::HistogramRange): a sample counts
// when edges[0] <= v < edges[nbins]; its bin is upper_bound(edges) - 1.
int32_t orc_bin(const double *edges, int32_t nbins, double v)
{
    if (!(v >= edges[0] && v < edges[nbins])) return -1;
    const double *ub = std::upper_bound(edges, edges + nbins, v);
    return (int32_t)(ub - edges) - 1;
}

// Tolerance bookkeeping for comparing a histogram filled from approximate
// sums with one filled from these sums: v may move to a neighbouring bin only
// if an edge adjacent to it lies within rtol * |v| (the edges at positions
// lower_bound(v) - 1 and lower_bound(v), as numpy's searchsorted 'left').
int32_t orc_near_edge(const double *edges, int32_t nbins, double v, double rtol)
{
    if (!(v == v)) return 0;
    const int64_t i = std::lower_bound(edges, edges + nbins + 1, v) - edges;
    for (int64_t j = i - 1; j <= i; ++j)
        if (j >= 0 && j <= nbins && std::fabs(v - edges[j]) <= rtol * std::fabs(v)) return 1;
    return 0;
}

// VoxelConstrainedSphereMethod.cu:138-154.
void orc_track_zshift(double *z, int64_t n, double half_length)
{
    for (int64_t i = 0; i < n; ++i) z[i] = z[i] + half_length;
}

namespace {

struct KeyIdx {
    uint64_t key;
    int64_t idx;   // position in the pair list (ascending point index)
};

// One oversample: FilterInScoringBox (:156-225), FilterTrackInSphere
// (:227-293) and ScoreTrackInSphere (:295-324), then the reduce-by-key of
// Histogram.cu:150-153 in canonical order.  pk/py receive the pairs.
void score_pairs(const double *x, const double *y, const double *zp, const double *e,
                 int64_t n, const orc_geometry &g, double sx, double sy, double sz,
                 std::vector<uint64_t> &pk, std::vector<double> &py)
{
    const double H = g.half_length;
    const double D = g.diameter;
    const double gso = g.gso;
    const double r2 = g.radius * g.radius;                         // :241
    const double lineal_denominator = (2. / 3.) * D;               // :302
    const uint64_t N = (uint64_t)g.n_linear;
    pk.clear();
    py.clear();
    for (int64_t i = 0; i < n; ++i) {
        double xs = x[i] + sx;                                     // :183
        double ys = y[i] + sy;                                     // :184
        double zs = zp[i] + sz;                                    // extension; == zp[i] when sz == 0
        if (!(std::fabs(xs) < H && std::fabs(ys) < H && std::fabs(zs) < H)) continue;  // :198
        double ax = xs - gso;                                      // :249-251
        double ay = ys - gso;
        double az = zs - gso;
        double dx = ((double)std::llrint(ax / D) * D) - ax;        // :253-255
        double dy = ((double)std::llrint(ay / D) * D) - ay;
        double dz = ((double)std::llrint(az / D) * D) - az;
        double dist = ((dx * dx) + (dy * dy)) + (dz * dz);         // :258
        if (!(dist <= r2)) continue;                               // :271
        uint64_t ix = (uint64_t)std::llrint((xs - gso) / D);       // :313-315
        uint64_t iy = (uint64_t)std::llrint((ys - gso) / D);
        uint64_t iz = (uint64_t)std::llrint((zs - gso) / D);
        uint64_t key = ix + iy * N + iz * N * N;                   // :318
        pk.push_back(key);
        py.push_back(e[i] / lineal_denominator);                   // :322
    }
}

// Stable sort by key, then left-to-right sums from the first value.
int64_t reduce_canonical(const std::vector<uint64_t> &pk, const std::vector<double> &py,
                         std::vector<KeyIdx> &scratch, uint64_t *keys, double *sums)
{
    const int64_t m = (int64_t)pk.size();
    scratch.resize(m);
    for (int64_t i = 0; i < m; ++i) scratch[i] = KeyIdx{pk[i], i};
    std::stable_sort(scratch.begin(), scratch.end(),
                     [](const KeyIdx &a, const KeyIdx &b) { return a.key < b.key; });
    int64_t u = 0;
    for (int64_t i = 0; i < m;) {
        uint64_t k = scratch[i].key;
        double s = py[scratch[i].idx];
        int64_t j = i + 1;
        for (; j < m && scratch[j].key == k; ++j) s = s + py[scratch[j].idx];
        if (keys) keys[u] = k;
        if (sums) sums[u] = s;
        ++u;
        i = j;
    }
    return u;
}

}  // namespace

int64_t orc_score_oversample(const double *x, const double *y, const double *zp,
                             const double *e, int64_t n, const orc_geometry *g,
                             double sx, double sy, double sz,
                             uint64_t *pair_keys, double *pair_y, int64_t *n_pairs,
                             uint64_t *keys, double *sums)
{
    std::vector<uint64_t> pk;
    std::vector<double> py;
    std::vector<KeyIdx> scratch;
    score_pairs(x, y, zp, e, n, *g, sx, sy, sz, pk, py);
    const int64_t m = (int64_t)pk.size();
    if (n_pairs) *n_pairs = m;
    // memcpy from an empty vector's data() (possibly null) is undefined even for 0 bytes
    if (pair_keys && m) std::memcpy(pair_keys, pk.data(), m * sizeof(uint64_t));
    if (pair_y && m) std::memcpy(pair_y, py.data(), m * sizeof(double));
    return reduce_canonical(pk, py, scratch, keys, sums);
}

int64_t orc_reduce_pairs(const uint64_t *pair_keys, const double *pair_y, int64_t m,
                         uint64_t *keys, double *sums)
{
    std::vector<uint64_t> pk(pair_keys, pair_keys + m);
    std::vector<double> py(pair_y, pair_y + m);
    std::vector<KeyIdx> scratch;
    return reduce_canonical(pk, py, scratch, keys, sums);
}

double orc_run(const double *x, const double *y, const double *z, const double *e,
               const int64_t *offsets, int64_t n_tracks, int64_t track_stride,
               const orc_geometry *geoms, const int32_t *geom_of_track,
               int32_t n_oversamples, uint64_t seed, uint32_t track_index_base,
               int32_t shift_xy, int32_t shift_z,
               const double *edges0, int32_t nbins0, uint64_t *bins0,
               const double *edges1, int32_t nbins1, uint64_t *bins1,
               int32_t n_threads, int32_t *threads_used,
               double near_rtol, uint64_t *near_counts)
{
    if (track_stride < 1) track_stride = 1;
    std::vector<int64_t> tracks;
    for (int64_t t = 0; t < n_tracks; t += track_stride) tracks.push_back(t);
    const int64_t n_seg = (int64_t)tracks.size() * n_oversamples;
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
    int used = omp_get_max_threads();
#else
    int used = 1;
#endif
    if (threads_used) *threads_used = used;
    std::vector<std::vector<uint64_t>> priv0(used, std::vector<uint64_t>(bins0 ? nbins0 : 0, 0));
    std::vector<std::vector<uint64_t>> priv1(used, std::vector<uint64_t>(bins1 ? nbins1 : 0, 0));
    std::vector<uint64_t> near_priv(2 * (size_t)used, 0);

    auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel
    {
#ifdef _OPENMP
        const int tid = omp_get_thread_num();
#else
        const int tid = 0;
#endif
        std::vector<uint64_t> pk, keys;
        std::vector<double> py, sums, zp, shifts;
        std::vector<KeyIdx> scratch;
        int64_t cur_track = -1;
#pragma omp for schedule(dynamic, 1)
        for (int64_t s = 0; s < n_seg; ++s) {
            const int64_t t = tracks[s / n_oversamples];
            const int32_t os = (int32_t)(s % n_oversamples);
            const orc_geometry &g = geoms[geom_of_track ? geom_of_track[t] : 0];
            const int64_t b = offsets[t], n = offsets[t + 1] - offsets[t];
            if (t != cur_track) {
                // Per-track set-up: z' = z + H' (:45) and the shift table (:49).
                zp.assign(z + b, z + b + n);
                orc_track_zshift(zp.data(), n, g.half_length);
                shifts.resize(3 * (size_t)n_oversamples);
                orc_shift_table(seed, track_index_base + (uint32_t)t, n_oversamples,
                                shift_xy, shift_z, &g, shifts.data());
                cur_track = t;
            }
            score_pairs(x + b, y + b, zp.data(), e + b, n, g, shifts[3 * os],
                        shifts[3 * os + 1], shifts[3 * os + 2], pk, py);
            keys.resize(pk.size());
            sums.resize(pk.size());
            int64_t u = reduce_canonical(pk, py, scratch, keys.data(), sums.data());
            for (int64_t i = 0; i < u; ++i) {
                if (bins0) {
                    int32_t bi = orc_bin(edges0, nbins0, sums[i]);
                    if (bi >= 0) priv0[tid][bi]++;
                    if (near_counts && orc_near_edge(edges0, nbins0, sums[i], near_rtol)) near_priv[2 * tid]++;
                }
                if (bins1) {
                    int32_t bi = orc_bin(edges1, nbins1, sums[i]);
                    if (bi >= 0) priv1[tid][bi]++;
                    if (near_counts && orc_near_edge(edges1, nbins1, sums[i], near_rtol)) near_priv[2 * tid + 1]++;
                }
            }
        }
    }
    auto t1 = std::chrono::steady_clock::now();
    for (int i = 0; i < used; ++i) {
        if (bins0) for (int32_t j = 0; j < nbins0; ++j) bins0[j] += priv0[i][j];
        if (bins1) for (int32_t j = 0; j < nbins1; ++j) bins1[j] += priv1[i][j];
        if (near_counts) {
            near_counts[0] += near_priv[2 * i];
            near_counts[1] += near_priv[2 * i + 1];
        }
    }
    return std::chrono::duration<double>(t1 - t0).count();
}

}  // extern "C"
