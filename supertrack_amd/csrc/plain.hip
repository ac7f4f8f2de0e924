// plain.hip -- the per-track drop-in path on gfx950: Track / VolumeEdepPair
// buffers, VoxelConstrainedSphereMethod (shift table, z offset, one fused
// shift -> box -> sphere -> key kernel per oversample) and Histogram
// (stable radix sort -> left-to-right reduce-by-key -> bin).
//
// Results are bit-identical to the FP64 restatement of the reference kernels
// (VoxelConstrainedSphereMethod.cu:156-324, Histogram.cu:147-171) with
// per-sphere sums taken in ascending point index: compaction preserves point
// order and the sort is stable.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <unordered_map>

#include "st_common.hpp"
#include "st_internal.hpp"
#include "st_device.hpp"

namespace st {
namespace {

// ---------------------------------------------------------------------------
// exact nearest lattice index
// ---------------------------------------------------------------------------
// The reference takes llrint(a / D) with a correctly rounded division.  a * invD
// differs from fl(a / D) by at most 3.0001 * 2^-53 * |a / D|, so the two round
// to the same integer unless q lies within that distance of a half-integer;
// only then is the division performed.  Either way the result equals
// llrint(fl(a / D)) exactly.
__device__ __forceinline__ double nearest_index(double a, double D, double invD)
{
    const double q = a * invD;
    double k = rint(q);
    const double f = fabs(q - k);   // exact: q and k are within a factor of 2 or k == 0
    if (f > 0.5 - (fabs(q) * 1e-15 + 1e-300)) k = rint(a / D);
    return k;
}

// ---------------------------------------------------------------------------
// per-track kernels
// ---------------------------------------------------------------------------
// Per-track set-up in one launch: ShiftTrackZOrigin's z += H'
// (VoxelConstrainedSphereMethod.cu:138-154), the lineal energy e / ((2/3) D)
// of each point (:302,322), which depends only on the point and so is taken
// once per track instead of once per oversample (the same IEEE value), and
// GenerateRandomXYShift's shift table (:61-83) with Philox.
struct SetupArgs {
    double *z;
    const double *e;
    double *lin;
    int64_t n;
    double half_length, denominator, length;
    double *shifts;   // 3 per oversample
    int n_os;
    uint64_t seed;
    uint32_t track_index;
    int shift_xy, shift_z;
};

__global__ void track_setup_kernel(SetupArgs a)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t first = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (int64_t i = first; i < a.n; i += stride) {
        a.z[i] = a.z[i] + a.half_length;
        a.lin[i] = a.e[i] / a.denominator;
    }
    for (int64_t k = first; k < a.n_os; k += stride) {
        const Shift s = shift_for(a.seed, a.track_index, (uint32_t)k, a.shift_xy, a.shift_z, a.length, a.half_length);
        a.shifts[3 * k + 0] = s.x;
        a.shifts[3 * k + 1] = s.y;
        a.shifts[3 * k + 2] = s.z;
    }
}

// ---------------------------------------------------------------------------
// fused scoring kernel with order-preserving decoupled look-back compaction
// ---------------------------------------------------------------------------
constexpr int kScoreBlock = 256;
constexpr int kScoreRows = 4;
constexpr int kScoreTile = kScoreBlock * kScoreRows;
constexpr int kWaves = kScoreBlock / 64;

// Geometry of the membership test, in the reference's units.
struct ScoreGeom {
    double H, D, invD, gso, r2;
    uint64_t N;
};

// FilterInScoringBox (:183-198), FilterTrackInSphere (:249-271) and the key of
// ScoreTrackInSphere (:313-318) for one point of a track whose z already
// carries +H': true if the shifted point lies in the voxel and in its nearest
// sphere, whose key is then written to key.
__device__ __forceinline__ bool score_point(double x, double y, double z, double sx, double sy, double sz,
                                            const ScoreGeom &g, uint64_t &key)
{
    const double xs = x + sx;
    const double ys = y + sy;
    const double zs = z + sz;
    if (!(fabs(xs) < g.H && fabs(ys) < g.H && fabs(zs) < g.H)) return false;
    const double ax = xs - g.gso;
    const double ay = ys - g.gso;
    const double az = zs - g.gso;
    const double kx = nearest_index(ax, g.D, g.invD);
    const double ky = nearest_index(ay, g.D, g.invD);
    const double kz = nearest_index(az, g.D, g.invD);
    const double dx = (kx * g.D) - ax;
    const double dy = (ky * g.D) - ay;
    const double dz = (kz * g.D) - az;
    const double dist = ((dx * dx) + (dy * dy)) + (dz * dz);
    if (!(dist <= g.r2)) return false;
    const uint64_t ix = (uint64_t)(int64_t)kx;
    const uint64_t iy = (uint64_t)(int64_t)ky;
    const uint64_t iz = (uint64_t)(int64_t)kz;
    key = ix + iy * g.N + iz * g.N * g.N;
    return true;
}

struct ScoreArgs {
    const double *x, *y, *z, *lin;
    int64_t n;
    ScoreGeom g;
    const double *shifts;
    int os;
    uint64_t *out_key;
    double *out_y;
    int32_t *out_count;        // device copy of the pair count (read by the histogram kernels)
    int32_t *out_count_host;   // host-visible copy (VolumeEdepPair::numElements)
    unsigned long long *status;
    unsigned long long *ticket;
    unsigned long long ticket_base;
    uint32_t epoch;
    int num_tiles;
};

// status word: [63:62] flag (1 aggregate, 2 inclusive prefix), [61:32] epoch,
// [31:0] count.  Epochs make stale words from earlier calls unreadable, so the
// array is never cleared.
__device__ __forceinline__ unsigned long long pack_status(unsigned long long flag, uint32_t epoch,
                                                          uint32_t v)
{
    return (flag << 62) | ((unsigned long long)(epoch & 0x3FFFFFFFu) << 32) | v;
}

__global__ __launch_bounds__(kScoreBlock) void score_kernel(ScoreArgs a)
{
    __shared__ int s_tile;
    __shared__ uint32_t s_counts[kScoreRows][kWaves];
    __shared__ uint32_t s_excl;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    if (tid == 0)
        s_tile = (int)(atomicAdd(a.ticket, 1ull) - a.ticket_base);
    __syncthreads();
    const int tile = s_tile;
    const int64_t base = (int64_t)tile * kScoreTile;
    const double sx = a.shifts[3 * a.os + 0];
    const double sy = a.shifts[3 * a.os + 1];
    const double sz = a.shifts[3 * a.os + 2];

    bool flag[kScoreRows];
    uint64_t key[kScoreRows];
    double yv[kScoreRows];
#pragma unroll
    for (int r = 0; r < kScoreRows; ++r) {
        flag[r] = false;
        key[r] = 0;
        yv[r] = 0.0;
        const int64_t i = base + (int64_t)r * kScoreBlock + tid;
        if (i < a.n && score_point(a.x[i], a.y[i], a.z[i], sx, sy, sz, a.g, key[r])) {
            yv[r] = a.lin[i];   // ScoreTrackInSphere's value (:322), computed once per track
            flag[r] = true;
        }
    }

    // ranks inside the tile (ascending point index = row-major order)
    uint32_t rank[kScoreRows];
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
    for (int r = 0; r < kScoreRows; ++r) {
        const uint64_t m = __ballot(flag[r]);
        rank[r] = (uint32_t)__popcll(m & lt_mask);
        if (lane == 0) s_counts[r][wave] = (uint32_t)__popcll(m);
    }
    __syncthreads();
    uint32_t tile_total = 0;
    uint32_t off[kScoreRows];
#pragma unroll
    for (int r = 0; r < kScoreRows; ++r) {
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            if (w == wave) off[r] = tile_total;
            tile_total += s_counts[r][w];
        }
    }

    // decoupled look-back by wave 0: each step reads the status words of up to
    // 64 predecessors at once (lane l: tile - 1 - l - done), sums the counts up
    // to the nearest inclusive prefix, and waits only if one of those words is
    // not yet published
    if (wave == 0) {
        uint32_t excl = 0;
        const uint32_t ep = a.epoch & 0x3FFFFFFFu;
        if (tile == 0) {
            if (lane == 0)
                __hip_atomic_store(&a.status[0], pack_status(2, a.epoch, tile_total), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        } else {
            if (lane == 0)
                __hip_atomic_store(&a.status[tile], pack_status(1, a.epoch, tile_total), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            int hi = tile - 1;   // nearest predecessor not yet summed
            while (true) {
                const int j = hi - lane;
                unsigned long long st = 0;
                if (j >= 0) st = __hip_atomic_load(&a.status[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const bool ready = j >= 0 && ((uint32_t)(st >> 32) & 0x3FFFFFFFu) == ep && (st >> 62) != 0;
                const uint64_t incl = __ballot(ready && (st >> 62) == 2);
                const uint64_t waiting = __ballot(j >= 0 && !ready);
                // lanes summed this step: up to the first inclusive prefix, or all
                const uint64_t span = incl ? ((2ull << (__ffsll((long long)incl) - 1)) - 1ull) : ~0ull;
                if (waiting & span) {
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                uint32_t v = (((span >> lane) & 1ull) && j >= 0) ? (uint32_t)st : 0u;
#pragma unroll
                for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
                excl += v;
                if (incl) break;
                hi -= 64;
            }
            if (lane == 0)
                __hip_atomic_store(&a.status[tile], pack_status(2, a.epoch, excl + tile_total), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        }
        if (lane == 0) {
            s_excl = excl;
            if (tile == a.num_tiles - 1) {
                *a.out_count = (int32_t)(excl + tile_total);
                *a.out_count_host = (int32_t)(excl + tile_total);
            }
        }
    }
    __syncthreads();
    const uint32_t excl = s_excl;
#pragma unroll
    for (int r = 0; r < kScoreRows; ++r) {
        if (flag[r]) {
            const uint32_t pos = excl + off[r] + rank[r];
            a.out_key[pos] = key[r];
            a.out_y[pos] = yv[r];
        }
    }
}

// ---------------------------------------------------------------------------
// Histogram: stable LSD radix sort -> left-to-right reduce-by-key -> bin
// ---------------------------------------------------------------------------
// Small inputs (m <= kLdsCap pairs) are grouped by one 1024-thread workgroup
// in LDS: one stable counting-sort pass on a 10-bit hash of the key puts every
// key's pairs in one bucket, in pair order; the reduce then sums each key's
// pairs left to right within its bucket, comparing keys staged in LDS in
// sorted order.  Larger inputs take the onesweep kernels below (8-bit digits
// over the full keys).
constexpr int kSortThreads = 1024;
constexpr int kSortWaves = kSortThreads / 64;
constexpr int kLdsCap = 12288;
constexpr int kHashBits = 10;
constexpr int kBuckets = 1 << kHashBits;
// per (bucket, wave) counts, stored wave-major ([wave][bucket]: random
// buckets of one wave fall in different LDS banks); thread t scans bucket t
constexpr size_t kCntBytes = (size_t)kBuckets * kSortWaves * sizeof(uint16_t);
// layout [P hash][P index][counts][Q hash][Q index]: the sorted result lands
// in Q (hashes and pair indices), and P + counts is one contiguous free region
constexpr size_t kSortBytes = (size_t)kLdsCap * 4 * sizeof(uint16_t) + kCntBytes;
// free region after the sort: keys and energies in sorted order fit up to
// kStage pairs
constexpr int kStage = (int)((kLdsCap * 2 * sizeof(uint16_t) + kCntBytes) / (2 * sizeof(uint64_t)));
// histograms of up to kBinCap bins are counted in LDS and added to the global
// counters once per launch: the lineal energies of one track fall in few
// bins, and global atomics on one address serialise in L2
constexpr int kBinCap = 4096;
constexpr size_t kLdsBytes = kSortBytes + (size_t)kBinCap * sizeof(uint32_t);
static_assert(kBuckets == kSortThreads, "one bucket row per thread in the scan");

__device__ __forceinline__ uint16_t key_bucket(uint64_t k, int bits)
{
    return (uint16_t)((k * 0x9E3779B97F4A7C15ull) >> (64 - bits));
}
struct HistArgs {
    const uint64_t *keys;
    const double *y;
    const int32_t *count;
    const double *edges;
    BinGuess guess;
    unsigned long long *bins;
};

// Exclusive scan of one value per thread over a 1024-thread workgroup.
__device__ __forceinline__ uint32_t block_exclusive_scan_1024(uint32_t v, uint32_t *s_wave)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    uint32_t incl = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t t = __shfl_up(incl, d, 64);
        if (lane >= d) incl += t;
    }
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    if (wave == 0) {
        const uint32_t w = (lane < kSortWaves) ? s_wave[lane] : 0u;
        uint32_t wi = w;
#pragma unroll
        for (int d = 1; d < kSortWaves; d <<= 1) {
            const uint32_t t = __shfl_up(wi, d, 64);
            if (lane >= d) wi += t;
        }
        if (lane < kSortWaves) s_wave[lane] = wi - w;
    }
    __syncthreads();
    const uint32_t r = s_wave[wave] + incl - v;
    __syncthreads();
    return r;
}

__device__ __forceinline__ void bin_sum(const HistArgs &a, double s)
{
    const int b = bin_fast(a.edges, a.guess, s);
    if (b >= 0) atomicAdd(&a.bins[b], 1ull);
}

// Histograms of up to kEdgeRegs * kSortThreads - 1 bins keep their edges in
// LDS when the free region has room (each sum's bin then costs no global trip)
constexpr int kEdgeRegs = 4;

// Lanes of the wave whose valid flag is set and whose hash equals this lane's.
__device__ __forceinline__ uint64_t hash_peers(uint32_t h, bool valid)
{
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < kHashBits; ++b) {
        const bool bit = (h >> b) & 1u;
        const uint64_t bb = __ballot(bit);
        peers &= bit ? bb : ~bb;
    }
    return peers;
}

// One workgroup: group m <= kLdsCap pairs by key bucket in LDS, reduce, bin.
__device__ void lds_sort_reduce_bin(const HistArgs &a, int m, unsigned char *smem)
{
    uint16_t *ph = reinterpret_cast<uint16_t *>(smem);
    // the P index slot (ph + kLdsCap) is not used by the one-pass sort; with
    // P hash and the counts it is the staging region after the scatter
    uint16_t *cnt = ph + 2 * kLdsCap;   // [kSortWaves][kBuckets]
    uint16_t *qh = cnt + kBuckets * kSortWaves;
    uint16_t *qv = qh + kLdsCap;
    uint32_t *sbins = reinterpret_cast<uint32_t *>(smem + kSortBytes);
    __shared__ uint32_t s_wave[kSortWaves];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int nbins = a.guess.nbins;
    const bool lds_bins = nbins <= kBinCap;
    if (lds_bins)
        for (int b = tid; b < nbins; b += kSortThreads) sbins[b] = 0;
    // edges and energies staged in the free region after the sort, if both fit
    const bool staged = m <= kStage;
    const bool lds_edges = staged && nbins < kEdgeRegs * kSortThreads &&
                           (size_t)m * 16 + (size_t)(nbins + 1) * 8 <= (size_t)kStage * 16;
    double er[kEdgeRegs];
    if (lds_edges) {
#pragma unroll
        for (int r = 0; r < kEdgeRegs; ++r) {
            const int e = tid + r * kSortThreads;
            er[r] = e <= nbins ? a.edges[e] : 0.0;
        }
    }
    for (int i = tid; i < m; i += kSortThreads) ph[i] = key_bucket(a.keys[i], kHashBits);
    // wave w owns pairs [w_lo, w_hi) and the counts cnt[w x kBuckets + *]
    const int per_wave = (m + kSortWaves - 1) / kSortWaves;
    const int w_lo = min(wave * per_wave, m);
    const int w_hi = min(w_lo + per_wave, m);
    const uint64_t lt_mask = (1ull << lane) - 1ull;
    // zero the counts with contiguous 16-byte stores (conflict-free)
    for (int k = tid; k < (int)(kCntBytes / sizeof(uint4)); k += kSortThreads)
        reinterpret_cast<uint4 *>(cnt)[k] = make_uint4(0u, 0u, 0u, 0u);
    __syncthreads();
    // count (order does not matter here): one LDS atomic per pair on the
    // 32-bit word holding its 16-bit (wave, bucket) counter; a counter stays
    // below 2^16 (at most per_wave pairs), so no carry reaches the other half
    {
        unsigned int *cnt32 = reinterpret_cast<unsigned int *>(cnt + wave * kBuckets);
        for (int i = w_lo + lane; i < w_hi; i += 64) {
            const unsigned h = ph[i];
            atomicAdd(&cnt32[h >> 1], 1u << (16 * (h & 1u)));
        }
    }
    __syncthreads();
    // exclusive scan, bucket-major: all of bucket b's pairs (waves in order)
    // precede bucket b + 1's
    {
        uint32_t run = 0;
        uint32_t local[kSortWaves];
#pragma unroll
        for (int w = 0; w < kSortWaves; ++w) {
            local[w] = run;
            run += cnt[w * kBuckets + tid];
        }
        const uint32_t base = block_exclusive_scan_1024(run, s_wave);
#pragma unroll
        for (int w = 0; w < kSortWaves; ++w) cnt[w * kBuckets + tid] = (uint16_t)(base + local[w]);
    }
    __syncthreads();
    // scatter in pair order: rank among lower lanes of the same hash, then
    // the lowest lane moves the wave's cursor past the step's peers
    for (int base = w_lo; base < w_hi; base += 64) {
        const int i = base + lane;
        const bool valid = i < w_hi;
        const uint32_t h = valid ? ph[i] : 0u;
        const uint64_t peers = hash_peers(h, valid);
        uint32_t cur = 0;
        if (valid) cur = cnt[wave * kBuckets + h];
        if (valid) {
            const uint32_t pos = cur + (uint32_t)__popcll(peers & lt_mask);
            qh[pos] = (uint16_t)h;
            qv[pos] = (uint16_t)i;
            if ((peers & lt_mask) == 0) cnt[wave * kBuckets + h] = (uint16_t)(cur + __popcll(peers));
        }
    }
    __syncthreads();
    const uint16_t *src_h = qh;
    const uint16_t *src_v = qv;
    // keys and energies in sorted order, staged in the free region (P pair +
    // counts) when they fit, edges after them; else read from global memory
    uint64_t *sk = reinterpret_cast<uint64_t *>(smem);
    double *sy = reinterpret_cast<double *>(sk + m);
    double *se = sy + m;
    if (staged) {
        for (int i = tid; i < m; i += kSortThreads) {
            const uint16_t v = src_v[i];
            sk[i] = a.keys[v];
            sy[i] = a.y[v];
        }
        if (lds_edges) {
#pragma unroll
            for (int r = 0; r < kEdgeRegs; ++r) {
                const int e = tid + r * kSortThreads;
                if (e <= nbins) se[e] = er[r];
            }
        }
        __syncthreads();
    }
    auto add_bin = [&](double v) {
        const int b = lds_edges ? bin_fast(se, a.guess, v) : bin_fast(a.edges, a.guess, v);
        if (b >= 0) {
            if (lds_bins) atomicAdd(&sbins[b], 1u);
            else atomicAdd(&a.bins[b], 1ull);
        }
    };
    // reduce-by-key left to right (Histogram.cu:153), bin (:158): the pair at
    // sorted position i starts its key's sum if no earlier pair of its bucket
    // has the same key; the sum then takes the bucket's pairs of that key in
    // order (ascending pair index: the sort is stable).  The staged loop reads
    // LDS only (kChunk pairs at a time, so only the FP64 adds stay serial);
    // the other reads global memory only: a loop that mixes the two gets
    // generic loads, whose waits also cover the global ones.
    constexpr int kChunk = 8;
    if (staged) {
#pragma unroll 1
        for (int i = tid; i < m; i += kSortThreads) {
            const uint16_t hb = src_h[i];
            const uint64_t kk = sk[i];
            bool first = true;
            bool live = true;
#pragma unroll 1
            for (int j = i - 1; live; j -= kChunk) {
                uint16_t hh[kChunk];
                uint64_t kv[kChunk];
#pragma unroll
                for (int u = 0; u < kChunk; ++u) {
                    const int jj = max(j - u, 0);
                    hh[u] = src_h[jj];
                    kv[u] = sk[jj];
                }
#pragma unroll
                for (int u = 0; u < kChunk; ++u) {
                    live = live && (j - u >= 0) && (hh[u] == hb);
                    if (live && kv[u] == kk) {
                        first = false;
                        live = false;
                    }
                }
            }
            if (!first) continue;
            double acc = sy[i];
            live = true;
#pragma unroll 1
            for (int j = i + 1; live; j += kChunk) {
                uint16_t hh[kChunk];
                uint64_t kv[kChunk];
                double yv[kChunk];
#pragma unroll
                for (int u = 0; u < kChunk; ++u) {
                    const int jj = min(j + u, m - 1);
                    hh[u] = src_h[jj];
                    kv[u] = sk[jj];
                    yv[u] = sy[jj];
                }
#pragma unroll
                for (int u = 0; u < kChunk; ++u) {
                    live = live && (j + u < m) && (hh[u] == hb);
                    if (live && kv[u] == kk) acc = acc + yv[u];
                }
            }
            add_bin(acc);
        }
    } else {
#pragma unroll 1
        for (int i = tid; i < m; i += kSortThreads) {
            const uint16_t hb = src_h[i];
            const uint64_t kk = a.keys[src_v[i]];
            bool first = true;
#pragma unroll 1
            for (int j = i - 1; j >= 0 && src_h[j] == hb; --j) {
                if (a.keys[src_v[j]] == kk) {
                    first = false;
                    break;
                }
            }
            if (!first) continue;
            double acc = a.y[src_v[i]];
#pragma unroll 1
            for (int j = i + 1; j < m && src_h[j] == hb; ++j)
                if (a.keys[src_v[j]] == kk) acc = acc + a.y[src_v[j]];
            add_bin(acc);
        }
    }
    if (!lds_bins) return;
    __syncthreads();
    for (int b = tid; b < nbins; b += kSortThreads) {
        const uint32_t c = sbins[b];
        if (c) atomicAdd(&a.bins[b], (unsigned long long)c);
    }
}

__global__ __launch_bounds__(kSortThreads) void hist_lds_kernel(HistArgs a)
{
    extern __shared__ unsigned char smem[];
    const int m = *a.count;
    if (m <= 0 || m > kLdsCap) return;
    lds_sort_reduce_bin(a, m, smem);
}

// ---------------------------------------------------------------------------
// queued oversamples: one workgroup per oversample, score -> sort -> bin
// ---------------------------------------------------------------------------
// A track of at most kLdsCap points whose oversamples were queued is run as one
// launch: workgroup k scores every point for oversample os0 + k (tiles of
// kSortThreads points, compacted in point order by wave ballots and a
// running offset), writes its pairs to its own scratch row (the launch's last
// oversample to the caller's pairs and counts, which the per-oversample path
// would have left there), and sorts, reduces and bins them into every
// histogram that consumed them, exactly as hist_lds_kernel does.
constexpr int kMaxQueuedHists = 4;

struct QueuedArgs {
    const double *x, *y, *z, *lin;
    int n;
    ScoreGeom g;
    const double *shifts;       // 3 per oversample of the track
    int os0;                    // oversample of workgroup 0
    int n_os;                   // workgroups
    uint64_t *row_key;          // scratch rows of n pairs, one per workgroup
    double *row_y;
    int last_to_caller;         // the last workgroup writes to the caller's pairs
    uint64_t *last_key;
    double *last_y;
    int32_t *last_count;        // device copy of the count
    int32_t *last_count_host;   // host-visible copy (VolumeEdepPair::numElements)
    HistArgs hist[kMaxQueuedHists];
    int n_hists;
};

__global__ __launch_bounds__(kSortThreads) void queued_track_kernel(QueuedArgs a)
{
    extern __shared__ unsigned char smem[];
    __shared__ uint32_t s_wcount[kSortWaves];
    const int k = blockIdx.x;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const bool to_caller = a.last_to_caller && k == a.n_os - 1;
    uint64_t *okey = to_caller ? a.last_key : a.row_key + (int64_t)k * a.n;
    double *oy = to_caller ? a.last_y : a.row_y + (int64_t)k * a.n;
    const int os = a.os0 + k;
    const double sx = a.shifts[3 * os + 0];
    const double sy = a.shifts[3 * os + 1];
    const double sz = a.shifts[3 * os + 2];
    const uint64_t lt_mask = (1ull << lane) - 1ull;
    uint32_t base = 0;
    for (int t0 = 0; t0 < a.n; t0 += kSortThreads) {
        const int i = t0 + tid;
        uint64_t key = 0;
        const bool flag = i < a.n && score_point(a.x[i], a.y[i], a.z[i], sx, sy, sz, a.g, key);
        const uint64_t m = __ballot(flag);
        if (lane == 0) s_wcount[wave] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t off = base, total = 0;
#pragma unroll
        for (int w = 0; w < kSortWaves; ++w) {
            const uint32_t c = s_wcount[w];
            off += (w < wave) ? c : 0u;
            total += c;
        }
        if (flag) {
            const uint32_t pos = off + (uint32_t)__popcll(m & lt_mask);
            okey[pos] = key;
            oy[pos] = a.lin[i];
        }
        base += total;
        __syncthreads();   // s_wcount is rewritten by the next tile
    }
    const int m = (int)base;
    if (to_caller && tid == 0) {
        *a.last_count = m;
        *a.last_count_host = m;
    }
    // the pairs written above are read back by other threads of the workgroup
    __syncthreads();
    if (m <= 0) return;
    for (int h = 0; h < a.n_hists; ++h) {
        HistArgs ha = a.hist[h];
        ha.keys = okey;
        ha.y = oy;
        lds_sort_reduce_bin(ha, m, smem);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Histogram, m > kLdsCap: onesweep LSD radix sort over global buffers
// ---------------------------------------------------------------------------
// os_digits_kernel  all passes' digit counts of all pairs (one read)
// os_pass_kernel    per 8-bit pass: each 2048-element tile ranks its digits
//                   stably (rows of 256 in order, 8 wave ballots per row),
//                   takes per-digit offsets from earlier tiles by a decoupled
//                   look-back (status words tagged with an epoch, never
//                   cleared) and scatters
// os_reduce_kernel  left-to-right sums of each run, exact binning; clears the
//                   digit counts for the next call
constexpr int kGBits = 8;
constexpr int kGDigits = 1 << kGBits;
constexpr int kMaxPasses = 8;
constexpr int kOsThreads = 256;
constexpr int kOsWaves = kOsThreads / 64;
constexpr int kOsRows = 8;
constexpr int kOsTile = kOsThreads * kOsRows;

__device__ __forceinline__ uint64_t same_digit_mask(bool valid, uint32_t d)
{
    uint64_t same = __ballot(valid);
#pragma unroll
    for (int bit = 0; bit < kGBits; ++bit) {
        const uint64_t bb = __ballot((d >> bit) & 1u);
        same &= ((d >> bit) & 1u) ? bb : ~bb;
    }
    return same;
}

struct OsDigitsArgs {
    const uint64_t *keys;
    const int32_t *count;
    unsigned int *digit_counts;   // [kMaxPasses][kGDigits]
    int passes;
};

__global__ __launch_bounds__(kOsThreads) void os_digits_kernel(OsDigitsArgs a)
{
    __shared__ unsigned int s_c[kMaxPasses * kGDigits];
    const int m = *a.count;
    const int64_t base = (int64_t)blockIdx.x * kOsTile;
    if (m <= kLdsCap || base >= m) return;
    for (int i = threadIdx.x; i < a.passes * kGDigits; i += kOsThreads) s_c[i] = 0;
    __syncthreads();
    for (int r = 0; r < kOsRows; ++r) {
        const int64_t i = base + (int64_t)r * kOsThreads + threadIdx.x;
        if (i < m) {
            const uint64_t k = a.keys[i];
            for (int p = 0; p < a.passes; ++p) atomicAdd(&s_c[p * kGDigits + ((k >> (p * kGBits)) & 255)], 1u);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < a.passes * kGDigits; i += kOsThreads)
        if (s_c[i]) atomicAdd(&a.digit_counts[i], s_c[i]);
}

struct OsPassArgs {
    const uint64_t *src_k;        // null on pass 0: the pairs' keys, identity payload
    const uint32_t *src_v;
    const uint64_t *pair_keys;
    uint64_t *dst_k;
    uint32_t *dst_v;
    const int32_t *count;
    const unsigned int *digit_counts;   // this pass's row
    unsigned long long *status;         // [tiles][kGDigits]
    unsigned long long *ticket;
    unsigned long long ticket_base;
    uint32_t epoch;
    int shift;
};

__global__ __launch_bounds__(kOsThreads) void os_pass_kernel(OsPassArgs a)
{
    __shared__ uint16_t s_cnt[kOsRows][kOsWaves][kGDigits];
    __shared__ uint32_t s_off[kGDigits];
    __shared__ uint32_t s_wsum[kOsWaves];
    __shared__ int s_tile;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    // every workgroup takes its ticket (the host advances the base by the grid size)
    if (tid == 0) s_tile = (int)(atomicAdd(a.ticket, 1ull) - a.ticket_base);
    for (int i = tid; i < kOsRows * kOsWaves * kGDigits; i += kOsThreads) (&s_cnt[0][0][0])[i] = 0;
    __syncthreads();
    const int m = *a.count;
    const int tile = s_tile;
    const int64_t base = (int64_t)tile * kOsTile;
    // small inputs were sorted in LDS; tiles past the data publish nothing
    // (tickets are handed out in order, so no later tile holds data)
    if (m <= kLdsCap || base >= m) return;
    // exclusive scan of the pass's digit counts: start of each digit's output
    uint32_t dbase;
    {
        const uint32_t c = a.digit_counts[tid];
        uint32_t incl = c;
#pragma unroll
        for (int dd = 1; dd < 64; dd <<= 1) {
            const uint32_t t = __shfl_up(incl, dd, 64);
            if (lane >= dd) incl += t;
        }
        if (lane == 63) s_wsum[wave] = incl;
        __syncthreads();
        uint32_t wb = 0;
        for (int w = 0; w < wave; ++w) wb += s_wsum[w];
        dbase = wb + incl - c;
    }
    uint64_t kk[kOsRows];
    uint32_t vv[kOsRows];
#pragma unroll
    for (int r = 0; r < kOsRows; ++r) {
        const int64_t i = base + (int64_t)r * kOsThreads + tid;
        kk[r] = 0;
        vv[r] = 0;
        if (i < m) {
            kk[r] = a.src_k ? a.src_k[i] : a.pair_keys[i];
            vv[r] = a.src_k ? a.src_v[i] : (uint32_t)i;
        }
    }
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    uint32_t rank[kOsRows];
#pragma unroll
    for (int r = 0; r < kOsRows; ++r) {
        const int64_t i = base + (int64_t)r * kOsThreads + tid;
        const bool valid = i < m;
        const uint32_t d = (uint32_t)((kk[r] >> a.shift) & (kGDigits - 1));
        const uint64_t same = same_digit_mask(valid, d);
        rank[r] = (uint32_t)__popcll(same & lt_mask);
        if (valid && rank[r] == 0) s_cnt[r][wave][d] = (uint16_t)__popcll(same);
    }
    __syncthreads();
    // thread d: exclusive prefix of digit d over (row, wave) in tile order,
    // then the look-back for digit d
    {
        const int d = tid;
        uint32_t run = 0;
        for (int r = 0; r < kOsRows; ++r)
            for (int w = 0; w < kOsWaves; ++w) {
                const uint32_t c = s_cnt[r][w][d];
                s_cnt[r][w][d] = (uint16_t)run;
                run += c;
            }
        const uint32_t ep = a.epoch & 0x3FFFFFFFu;
        unsigned long long *st = a.status + (size_t)tile * kGDigits + d;
        auto pack = [&](unsigned long long flag, uint32_t v) {
            return (flag << 62) | ((unsigned long long)ep << 32) | v;
        };
        uint32_t excl = 0;
        if (tile == 0) {
            __hip_atomic_store(st, pack(2, run), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            __hip_atomic_store(st, pack(1, run), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            int j = tile - 1;
            while (true) {
                const unsigned long long s =
                    __hip_atomic_load(a.status + (size_t)j * kGDigits + d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const uint32_t sep = (uint32_t)(s >> 32) & 0x3FFFFFFFu;
                const unsigned long long fl = s >> 62;
                if (sep != ep || fl == 0) {
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                excl += (uint32_t)s;
                if (fl == 2) break;
                --j;
            }
            __hip_atomic_store(st, pack(2, excl + run), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        s_off[d] = dbase + excl;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kOsRows; ++r) {
        const int64_t i = base + (int64_t)r * kOsThreads + tid;
        if (i < m) {
            const uint32_t d = (uint32_t)((kk[r] >> a.shift) & (kGDigits - 1));
            const uint32_t pos = s_off[d] + s_cnt[r][wave][d] + rank[r];
            a.dst_k[pos] = kk[r];
            a.dst_v[pos] = vv[r];
        }
    }
}

struct OsReduceArgs {
    HistArgs h;
    const uint64_t *keys;     // sorted keys (null: no pass ran; use the pairs' keys)
    const uint32_t *vals;
    unsigned int *digit_counts;
    int passes;
};

__global__ __launch_bounds__(256) void os_reduce_kernel(OsReduceArgs a)
{
    const int m = *a.h.count;
    if (blockIdx.x == 0)
        for (int i = threadIdx.x; i < kMaxPasses * kGDigits; i += blockDim.x) a.digit_counts[i] = 0;
    if (m <= kLdsCap) return;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
        const uint64_t kk = a.keys ? a.keys[i] : a.h.keys[i];
        if (i > 0 && (a.keys ? a.keys[i - 1] : a.h.keys[i - 1]) == kk) continue;
        double s = a.h.y[a.keys ? a.vals[i] : (uint32_t)i];
        for (int64_t j = i + 1; j < m; ++j) {
            const uint64_t kj = a.keys ? a.keys[j] : a.h.keys[j];
            if (kj != kk) break;
            s = s + a.h.y[a.keys ? a.vals[j] : (uint32_t)j];
        }
        bin_sum(a.h, s);
    }
}

__global__ void zero_u64_kernel(unsigned long long *p, int64_t n)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0;
}

ScoreGeom score_geom(const st_geometry &g)
{
    ScoreGeom s;
    s.H = g.half_length;
    s.D = g.diameter;
    s.invD = 1.0 / g.diameter;
    s.gso = g.gso;
    s.r2 = g.radius * g.radius;
    s.N = (uint64_t)g.n_linear;
    return s;
}

inline int grid_for(int64_t n, int threads, int cap = 65535)
{
    int64_t g = (n + threads - 1) / threads;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

template <typename T>
int grow(T **p, int64_t *cap, int64_t need)
{
    if (need <= *cap && *p) return ST_OK;
    if (*p) {
        (void)hipFree(*p);
        *p = nullptr;
        *cap = 0;
    }
    if (need <= 0) return ST_OK;
    hipError_t e = dev_malloc((void **)p, (size_t)need * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return set_error(ST_ENOMEM, "hipMalloc(%lld x %zu B) failed: %s", (long long)need, sizeof(T),
                         hipGetErrorString(e));
    }
    *cap = need;
    return ST_OK;
}


// The LDS kernels use more than 64 KiB of dynamic LDS: allowed once per device.
int ensure_sort_attributes()
{
    constexpr int kDevices = 64;
    static std::mutex mu;
    static bool done[kDevices] = {};
    int dev = 0;
    ST_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kDevices) return set_error(ST_ERANGE, "device %d beyond the %d supported", dev, kDevices);
    std::lock_guard<std::mutex> lock(mu);
    if (done[dev]) return ST_OK;
    ST_HIP_TRY(hipFuncSetAttribute((const void *)hist_lds_kernel,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsBytes));
    ST_HIP_TRY(hipFuncSetAttribute((const void *)queued_track_kernel,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsBytes));
    done[dev] = true;
    return ST_OK;
}

// Pairs with queued oversamples pending, by the pairs' key array, and the
// method that will produce them.
std::mutex g_queue_mu;
std::unordered_map<const void *, st_vcs *> g_queue;

st_vcs *queue_producer(const void *pairs_key)
{
    std::lock_guard<std::mutex> lock(g_queue_mu);
    auto it = g_queue.find(pairs_key);
    return it == g_queue.end() ? nullptr : it->second;
}

// Largest scratch of queued pair rows per method (n points x oversamples x
// 16 B); tracks that would need more run oversample by oversample.
constexpr int64_t kQueueScratchBytes = (int64_t)1 << 30;

// Onesweep buffers for `cap` pairs: ping-pong keys/payloads, look-back status
// words per (tile, digit), the digit counts and the tile ticket.
int ensure_sort_buffers(st_hist *h, int64_t cap)
{
    // each allocated on its own, so that a failure between them leaves no
    // half-made state behind for the next call
    if (!h->digit_counts) {
        ST_ALLOC_TRY(dev_malloc((void **)&h->digit_counts, (size_t)kMaxPasses * kGDigits * sizeof(unsigned int)));
        ST_HIP_TRY(hipMemset(h->digit_counts, 0, (size_t)kMaxPasses * kGDigits * sizeof(unsigned int)));
    }
    if (!h->os_ticket) {
        ST_ALLOC_TRY(dev_malloc((void **)&h->os_ticket, sizeof(unsigned long long)));
        ST_HIP_TRY(hipMemset(h->os_ticket, 0, sizeof(unsigned long long)));
        h->os_ticket_base = 0;
    }
    if (cap > h->sort_capacity) {
        (void)hipFree(h->keys_a); (void)hipFree(h->keys_b); (void)hipFree(h->vals_a); (void)hipFree(h->vals_b);
        (void)hipFree(h->os_status);
        h->keys_a = h->keys_b = nullptr;
        h->vals_a = h->vals_b = nullptr;
        h->os_status = nullptr;
        h->sort_capacity = 0;
        const int64_t tiles = (cap + kOsTile - 1) / kOsTile;
        ST_ALLOC_TRY(dev_malloc((void **)&h->keys_a, (size_t)cap * sizeof(uint64_t)));
        ST_ALLOC_TRY(dev_malloc((void **)&h->keys_b, (size_t)cap * sizeof(uint64_t)));
        ST_ALLOC_TRY(dev_malloc((void **)&h->vals_a, (size_t)cap * sizeof(uint32_t)));
        ST_ALLOC_TRY(dev_malloc((void **)&h->vals_b, (size_t)cap * sizeof(uint32_t)));
        ST_ALLOC_TRY(dev_malloc((void **)&h->os_status, (size_t)tiles * kGDigits * sizeof(unsigned long long)));
        ST_HIP_TRY(hipMemset(h->os_status, 0, (size_t)tiles * kGDigits * sizeof(unsigned long long)));
        h->sort_capacity = cap;
    }
    return ST_OK;
}

}  // namespace

// Sort -> reduce -> bin the pairs into h->bins_dev on `stream` (no zeroing).
int hist_add_pairs(st_hist *h, const st_pairs *pairs, hipStream_t stream)
{
    int rc = ensure_sort_attributes();
    if (rc) return rc;
    HistArgs ha;
    ha.keys = pairs->volume;
    ha.y = pairs->edep;
    ha.count = pairs->count_dev;
    ha.edges = h->edges_dev;
    ha.guess = h->guess;
    ha.bins = (unsigned long long *)h->bins_dev;
    h->bins_zero = false;
    // small inputs (decided on the device) take the one-workgroup LDS sort
    hist_lds_kernel<<<1, kSortThreads, kLdsBytes, stream>>>(ha);
    ST_HIP_TRY(hipGetLastError());
    if (pairs->capacity <= kLdsCap) return ST_OK;
    rc = ensure_sort_buffers(h, pairs->capacity);
    if (rc) return rc;
    // passes for the largest key the producer can write (all 64 bits if unknown)
    const uint64_t kb = pairs->key_bound ? pairs->key_bound : ~0ull;
    const int bits = 64 - __builtin_clzll(kb | 1ull);
    const int passes = (bits + kGBits - 1) / kGBits;
    const int tiles = (int)((pairs->capacity + kOsTile - 1) / kOsTile);
    OsDigitsArgs da{pairs->volume, pairs->count_dev, h->digit_counts, passes};
    os_digits_kernel<<<tiles, kOsThreads, 0, stream>>>(da);
    const uint64_t *src_k = nullptr;
    const uint32_t *src_v = nullptr;
    uint64_t *bk[2] = {h->keys_a, h->keys_b};
    uint32_t *bv[2] = {h->vals_a, h->vals_b};
    for (int p = 0; p < passes; ++p) {
        OsPassArgs pa;
        pa.src_k = src_k;
        pa.src_v = src_v;
        pa.pair_keys = pairs->volume;
        pa.dst_k = bk[p & 1];
        pa.dst_v = bv[p & 1];
        pa.count = pairs->count_dev;
        pa.digit_counts = h->digit_counts + p * kGDigits;
        pa.status = h->os_status;
        pa.ticket = h->os_ticket;
        pa.ticket_base = h->os_ticket_base;
        pa.epoch = ++h->os_epoch;
        if ((pa.epoch & 0x3FFFFFFFu) == 0) pa.epoch = ++h->os_epoch;
        pa.shift = p * kGBits;
        os_pass_kernel<<<tiles, kOsThreads, 0, stream>>>(pa);
        h->os_ticket_base += (unsigned long long)tiles;
        src_k = bk[p & 1];
        src_v = bv[p & 1];
    }
    OsReduceArgs ra{ha, src_k, src_v, h->digit_counts, passes};
    os_reduce_kernel<<<grid_for(pairs->capacity, 256, 2048), 256, 0, stream>>>(ra);
    ST_HIP_TRY(hipGetLastError());
    return ST_OK;
}

int queue_flush(st_vcs *m)
{
    if (!m->queue_active) return ST_OK;
    const int cnt = m->queue_count;
    const int64_t n = m->n;
    int rc = ensure_sort_attributes();
    if (!rc) rc = grow(&m->queue_key, &m->queue_key_capacity, std::max<int64_t>((int64_t)cnt * n, 1));
    if (!rc) rc = grow(&m->queue_y, &m->queue_y_capacity, std::max<int64_t>((int64_t)cnt * n, 1));
    std::vector<st_hist *> consumers;
    if (!rc) {
        QueuedArgs a{};
        a.x = m->queue_x;
        a.y = m->queue_yc;
        a.z = m->queue_z;
        a.lin = m->lineal;
        a.n = (int)n;
        a.g = score_geom(m->g);
        a.shifts = m->shifts;
        a.last_key = m->queue_pairs.volume;
        a.last_y = m->queue_pairs.edep;
        a.last_count = m->queue_pairs.count_dev;
        a.last_count_host = m->queue_pairs.num_elements;
        // runs of consecutive oversamples consumed by the same histograms,
        // up to kMaxQueuedHists histograms per launch
        for (int k0 = 0; k0 < cnt && !rc;) {
            int k1 = k0 + 1;
            while (k1 < cnt && m->queue_consumers[k1] == m->queue_consumers[k0]) ++k1;
            const std::vector<st_hist *> &cons = m->queue_consumers[k0];
            const bool has_last = k1 == cnt;
            for (size_t c0 = 0; c0 < std::max<size_t>(cons.size(), has_last ? 1 : 0) && !rc;
                 c0 += kMaxQueuedHists) {
                a.os0 = m->queue_os0 + k0;
                a.n_os = k1 - k0;
                a.row_key = m->queue_key + (int64_t)k0 * n;
                a.row_y = m->queue_y + (int64_t)k0 * n;
                a.last_to_caller = has_last && c0 == 0;
                a.n_hists = 0;
                for (size_t c = c0; c < cons.size() && c < c0 + kMaxQueuedHists; ++c) {
                    st_hist *h = cons[c];
                    HistArgs &ha = a.hist[a.n_hists++];
                    ha.keys = nullptr;
                    ha.y = nullptr;
                    ha.count = nullptr;
                    ha.edges = h->edges_dev;
                    ha.guess = h->guess;
                    ha.bins = (unsigned long long *)h->bins_dev;
                    if (std::find(consumers.begin(), consumers.end(), h) == consumers.end()) consumers.push_back(h);
                    h->bins_zero = false;
                }
                queued_track_kernel<<<a.n_os, kSortThreads, kLdsBytes, m->stream>>>(a);
                if (hipGetLastError() != hipSuccess) rc = set_error(ST_EHIP, "queued_track_kernel launch failed");
            }
            k0 = k1;
        }
    }
    // histograms on other streams wait for these launches
    for (st_hist *h : consumers) {
        if (rc || h->stream == m->stream) continue;
        if (!m->queue_event && hipEventCreateWithFlags(&m->queue_event, hipEventDisableTiming) != hipSuccess) {
            m->queue_event = nullptr;
            rc = set_error(ST_EHIP, "queued oversamples: hipEventCreate failed");
            break;
        }
        if (hipEventRecord(m->queue_event, m->stream) != hipSuccess ||
            hipStreamWaitEvent(h->stream, m->queue_event, 0) != hipSuccess)
            rc = set_error(ST_EHIP, "queued oversamples: stream ordering failed");
    }
    // the group ends whether or not it ran
    {
        std::lock_guard<std::mutex> lock(g_queue_mu);
        auto it = g_queue.find(m->queue_pairs.volume);
        if (it != g_queue.end() && it->second == m) g_queue.erase(it);
    }
    for (const auto &cons : m->queue_consumers)
        for (st_hist *h : cons) h->producers.erase(std::remove(h->producers.begin(), h->producers.end(), m),
                                                   h->producers.end());
    m->queue_active = false;
    m->queue_count = 0;
    m->queue_consumers.clear();
    return rc;
}

// Flushes every method whose queued oversamples h consumes.
int queue_flush_producers(st_hist *h)
{
    int rc = ST_OK;
    while (!h->producers.empty()) {
        st_vcs *m = h->producers.back();
        const int r = queue_flush(m);
        if (r && !rc) rc = r;
        h->producers.erase(std::remove(h->producers.begin(), h->producers.end(), m), h->producers.end());
    }
    return rc;
}

}  // namespace st

using namespace st;

extern "C" {

// ---------------------------------------------------------------------------
// Track / VolumeEdepPair
// ---------------------------------------------------------------------------
int st_track_alloc(st_track *t, int64_t capacity)
{
    if (!t || capacity < 0) return set_error(ST_EINVAL, "st_track_alloc: bad arguments");
    std::memset(t, 0, sizeof *t);
    // Track::AllocateEmptyTrack (Track.cu:112-118); one spare slot like the
    // reference's n + 1 allocation (Track.cu:16).
    const size_t bytes = (size_t)(capacity + 1) * sizeof(double);
    double **arrs[4] = {&t->x, &t->y, &t->z, &t->edep};
    for (double **p : arrs) {
        hipError_t e = dev_malloc((void **)p, bytes);
        if (e != hipSuccess) {
            st_track_free(t);
            return set_error(ST_ENOMEM, "st_track_alloc: hipMalloc(%zu) failed: %s", bytes,
                             hipGetErrorString(e));
        }
    }
    t->capacity = capacity;
    t->n = 0;
    return ST_OK;
}

int st_track_upload(st_track *t, const double *x, const double *y, const double *z,
                    const double *edep, int64_t n)
{
    if (!t || n < 0 || (n > 0 && (!x || !y || !z || !edep)))
        return set_error(ST_EINVAL, "st_track_upload: bad arguments");
    if (!t->x || n > t->capacity) {
        st_track_free(t);
        int rc = st_track_alloc(t, n);
        if (rc) return rc;
    }
    const size_t bytes = (size_t)n * sizeof(double);
    if (bytes) {
        ST_HIP_TRY(hipDeviceSynchronize());
        ST_HIP_TRY(hipMemcpy(t->x, x, bytes, hipMemcpyHostToDevice));
        ST_HIP_TRY(hipMemcpy(t->y, y, bytes, hipMemcpyHostToDevice));
        ST_HIP_TRY(hipMemcpy(t->z, z, bytes, hipMemcpyHostToDevice));
        ST_HIP_TRY(hipMemcpy(t->edep, edep, bytes, hipMemcpyHostToDevice));
        ST_HIP_TRY(hipDeviceSynchronize());
    }
    t->n = n;
    return ST_OK;
}

void st_track_free(st_track *t)
{
    if (!t) return;
    (void)hipFree(t->x);
    (void)hipFree(t->y);
    (void)hipFree(t->z);
    (void)hipFree(t->edep);
    t->x = t->y = t->z = t->edep = nullptr;
    t->n = t->capacity = 0;
}

// VolumeEdepPair::Allocate (VolumeEdepPair.cu:6-17): numElements starts at
// the capacity and the arrays start zeroed.  numElements is managed memory
// in the reference, read on the host after the track's synchronisation
// (Histogram.cu:105); here it is pinned, device-mapped host memory written by
// the producer next to a device copy that the histogram kernels read.
int st_pairs_alloc(st_pairs *p, int64_t capacity)
{
    if (!p || capacity < 0 || capacity > 0x7FFFFFFF)
        return set_error(ST_EINVAL, "st_pairs_alloc: capacity must be in [0, 2^31)");
    std::memset(p, 0, sizeof *p);
    const size_t n = (size_t)(capacity > 0 ? capacity : 1);
    hipError_t e1 = dev_malloc((void **)&p->volume, n * sizeof(uint64_t));
    hipError_t e2 = dev_malloc((void **)&p->edep, n * sizeof(double));
    hipError_t e3 = dev_malloc((void **)&p->count_dev, sizeof(int32_t));
    hipError_t e4 = pinned_malloc((void **)&p->num_elements, sizeof(int32_t),
                                  hipHostMallocMapped | hipHostMallocCoherent);
    if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess || e4 != hipSuccess) {
        if (e4 != hipSuccess) p->num_elements = nullptr;
        st_pairs_free(p);
        return set_error(ST_ENOMEM, "st_pairs_alloc: device or pinned host allocation failed");
    }
    // kernels store through the same address (unified addressing)
    void *mapped = nullptr;
    if (hipHostGetDevicePointer(&mapped, p->num_elements, 0) != hipSuccess || mapped != p->num_elements) {
        st_pairs_free(p);
        return set_error(ST_EHIP, "st_pairs_alloc: pinned count is not mapped at its host address");
    }
    ST_HIP_TRY(hipMemset(p->volume, 0, n * sizeof(uint64_t)));
    ST_HIP_TRY(hipMemset(p->edep, 0, n * sizeof(double)));
    const int32_t c = (int32_t)capacity;
    ST_HIP_TRY(hipMemcpy(p->count_dev, &c, sizeof c, hipMemcpyHostToDevice));
    *p->num_elements = c;
    p->capacity = capacity;
    p->key_bound = 0;
    return ST_OK;
}

void st_pairs_free(st_pairs *p)
{
    if (!p) return;
    if (p->volume) {
        st_vcs *m = queue_producer(p->volume);
        if (m) (void)queue_flush(m);
    }
    if (p->volume || p->edep || p->count_dev || p->num_elements) (void)hipDeviceSynchronize();
    (void)hipFree(p->volume);
    (void)hipFree(p->edep);
    (void)hipFree(p->count_dev);
    if (p->num_elements) (void)hipHostFree(p->num_elements);
    p->volume = nullptr;
    p->edep = nullptr;
    p->count_dev = nullptr;
    p->num_elements = nullptr;
    p->capacity = 0;
}

int st_pairs_set_count(st_pairs *p, int32_t count)
{
    if (!p || !p->count_dev || !p->num_elements || count < 0 || count > p->capacity)
        return set_error(ST_EINVAL, "st_pairs_set_count: bad arguments");
    ST_HIP_TRY(hipDeviceSynchronize());
    ST_HIP_TRY(hipMemcpy(p->count_dev, &count, sizeof count, hipMemcpyHostToDevice));
    *p->num_elements = count;
    return ST_OK;
}

// ---------------------------------------------------------------------------
// VoxelConstrainedSphereMethod
// ---------------------------------------------------------------------------
int st_vcs_create(const st_geometry *g, int32_t shift_xy, int32_t shift_z,
                  int32_t suggested_blocks, int32_t suggested_threads, st_vcs **out)
{
    if (!g || !out) return set_error(ST_EINVAL, "st_vcs_create: null argument");
    *out = nullptr;
    if (!(g->diameter > 0) || g->n_linear < 1)
        return set_error(ST_EINVAL, "st_vcs_create: invalid geometry");
    st_vcs *m = new (std::nothrow) st_vcs;
    if (!m) return set_error(ST_ENOMEM, "st_vcs_create: out of host memory");
    m->g = *g;
    m->shift_xy = shift_xy ? 1 : 0;
    m->shift_z = shift_z ? 1 : 0;
    m->suggested_blocks = suggested_blocks;
    m->suggested_threads = suggested_threads;
    m->stream = default_stream();
    hipError_t e = dev_malloc((void **)&m->ticket, sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemset(m->ticket, 0, sizeof(unsigned long long));
    if (e != hipSuccess) {
        st_vcs_destroy(m);
        return set_error(hip_status(e), "st_vcs_create: %s", hipGetErrorString(e));
    }
    *out = m;
    return ST_OK;
}

int st_vcs_set_stream(st_vcs *m, void *hip_stream)
{
    if (!m) return set_error(ST_EINVAL, "st_vcs_set_stream: null method");
    int rc = queue_flush(m);
    if (rc) return rc;
    m->stream = hip_stream ? (hipStream_t)hip_stream : default_stream();
    return ST_OK;
}

int st_vcs_set_queued(st_vcs *m, int32_t on)
{
    if (!m) return set_error(ST_EINVAL, "st_vcs_set_queued: null method");
    int rc = queue_flush(m);
    if (rc) return rc;
    m->queue_mode = on != 0;
    if (!m->queue_mode) m->queue_track = false;
    return ST_OK;
}

int st_vcs_begin_track(st_vcs *m, st_track *track, uint64_t seed, uint32_t track_index,
                       int32_t n_oversamples)
{
    if (!m || !track) return set_error(ST_EINVAL, "st_vcs_begin_track: null argument");
    if (n_oversamples < 0) return set_error(ST_EINVAL, "st_vcs_begin_track: n_oversamples < 0");
    if (track->n < 0 || track->n > ((int64_t)1 << 31) - 1)
        return set_error(ST_ERANGE, "st_vcs_begin_track: track of %lld points (limit 2^31 - 1)",
                         (long long)track->n);
    if (track->n > 0 && (!track->x || !track->y || !track->z || !track->edep))
        return set_error(ST_EINVAL, "st_vcs_begin_track: track has null arrays");
    const int64_t n = track->n;
    int rc = queue_flush(m);   // oversamples left pending by a track that was not ended
    if (rc) return rc;
    m->in_track = false;       // until this track is set up
    rc = grow(&m->lineal, &m->lineal_capacity, std::max<int64_t>(n, 1));
    if (rc) return rc;
    int64_t shifts_cap = m->shifts_capacity;
    rc = grow(&m->shifts, &shifts_cap, std::max<int64_t>(3 * (int64_t)n_oversamples, 3));
    if (rc) return rc;
    m->shifts_capacity = (int32_t)shifts_cap;
    const int64_t tiles = std::max<int64_t>(1, (n + kScoreTile - 1) / kScoreTile);
    if (tiles > m->status_capacity) {
        rc = grow(&m->status, &m->status_capacity, tiles);
        if (rc) return rc;
        ST_HIP_TRY(hipMemsetAsync(m->status, 0, (size_t)tiles * sizeof(unsigned long long), m->stream));
    }
    // ShiftTrackZOrigin, once per track (VoxelConstrainedSphereMethod.cu:45),
    // the lineal energies and the shift table
    if (n > 0 || n_oversamples > 0) {
        SetupArgs sa;
        sa.z = track->z;
        sa.e = track->edep;
        sa.lin = m->lineal;
        sa.n = n;
        sa.half_length = m->g.half_length;
        sa.denominator = (2. / 3.) * m->g.diameter;
        sa.length = m->g.length;
        sa.shifts = m->shifts;
        sa.n_os = n_oversamples;
        sa.seed = seed;
        sa.track_index = track_index;
        sa.shift_xy = m->shift_xy;
        sa.shift_z = m->shift_z;
        track_setup_kernel<<<grid_for(std::max<int64_t>(n, n_oversamples), 256, 4096), 256, 0, m->stream>>>(sa);
    }
    ST_HIP_TRY(hipGetLastError());
    m->n = n;
    m->n_oversamples = n_oversamples;
    m->oversample = 0;
    m->in_track = true;
    // queued: tracks whose pairs fit the one-workgroup sort and whose pair
    // rows fit the scratch bound; others run oversample by oversample
    m->queue_track = m->queue_mode && n >= 1 && n <= kLdsCap &&
                     n * (int64_t)std::max(n_oversamples, 1) * 16 <= kQueueScratchBytes;
    return ST_OK;
}

int st_vcs_process_oversample(st_vcs *m, const st_track *track, st_pairs *out)
{
    if (!m || !track || !out) return set_error(ST_EINVAL, "st_vcs_process_oversample: null argument");
    if (!m->in_track) return set_error(ST_ESTATE, "st_vcs_process_oversample before st_vcs_begin_track");
    if (m->oversample >= m->n_oversamples)
        return set_error(ST_ESTATE, "st_vcs_process_oversample: all %d oversamples already processed",
                         m->n_oversamples);
    if (track->n != m->n)
        return set_error(ST_EINVAL, "st_vcs_process_oversample: track has %lld points, begun with %lld",
                         (long long)track->n, (long long)m->n);
    if (out->capacity < m->n || !out->volume || !out->edep || !out->num_elements || !out->count_dev)
        return set_error(ST_EINVAL, "st_vcs_process_oversample: pairs capacity %lld < %lld points",
                         (long long)out->capacity, (long long)m->n);
    // largest key this method can write: every index is at most N
    {
        const unsigned long long N = (unsigned long long)m->g.n_linear;
        const long double kb = (long double)N + (long double)N * N + (long double)N * N * N;
        out->key_bound = kb >= 1.8e19L ? 0ull : (N + N * N + N * N * N);
    }
    if (m->queue_track) {
        // record the oversample; it runs when the track ends (or earlier, if
        // the pending pairs are needed)
        if (m->queue_active && (m->queue_pairs.volume != out->volume || m->queue_x != track->x ||
                                m->queue_yc != track->y || m->queue_z != track->z)) {
            int rc = queue_flush(m);
            if (rc) return rc;
        }
        if (!m->queue_active) {
            st_vcs *other = queue_producer(out->volume);
            if (other && other != m) {
                int rc = queue_flush(other);
                if (rc) return rc;
            }
            m->queue_active = true;
            m->queue_pairs = *out;
            m->queue_x = track->x;
            m->queue_yc = track->y;
            m->queue_z = track->z;
            m->queue_os0 = m->oversample;
            m->queue_count = 0;
            m->queue_consumers.clear();
            std::lock_guard<std::mutex> lock(g_queue_mu);
            g_queue[out->volume] = m;
        }
        m->queue_count++;
        m->queue_consumers.emplace_back();
        m->oversample++;
        return ST_OK;
    }
    const int64_t n = m->n;
    const int tiles = (int)std::max<int64_t>(1, (n + kScoreTile - 1) / kScoreTile);
    ScoreArgs a;
    a.x = track->x;
    a.y = track->y;
    a.z = track->z;
    a.lin = m->lineal;
    a.n = n;
    a.g = score_geom(m->g);
    a.shifts = m->shifts;
    a.os = m->oversample;
    a.out_key = out->volume;
    a.out_y = out->edep;
    a.out_count = out->count_dev;
    a.out_count_host = out->num_elements;
    a.status = m->status;
    a.ticket = m->ticket;
    a.ticket_base = m->ticket_base;
    a.epoch = ++m->epoch;
    if ((a.epoch & 0x3FFFFFFFu) == 0) a.epoch = ++m->epoch;
    a.num_tiles = tiles;
    score_kernel<<<tiles, kScoreBlock, 0, m->stream>>>(a);
    ST_HIP_TRY(hipGetLastError());
    m->ticket_base += (uint64_t)tiles;
    m->oversample++;
    return ST_OK;
}

int st_vcs_end_track(st_vcs *m)
{
    if (!m) return set_error(ST_EINVAL, "st_vcs_end_track: null method");
    const int rc = queue_flush(m);
    ST_HIP_TRY(hipStreamSynchronize(m->stream));
    m->in_track = false;
    return rc;
}

void st_vcs_destroy(st_vcs *m)
{
    if (!m) return;
    (void)queue_flush(m);
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    (void)hipFree(m->queue_key);
    (void)hipFree(m->queue_y);
    if (m->queue_event) (void)hipEventDestroy(m->queue_event);
    (void)hipFree(m->lineal);
    (void)hipFree(m->shifts);
    (void)hipFree(m->status);
    (void)hipFree(m->ticket);
    delete m;
}

// ---------------------------------------------------------------------------
// Histogram
// ---------------------------------------------------------------------------
int st_hist_create(int32_t type, float lo, float hi, int32_t nbins, st_hist **out)
{
    if (!out) return set_error(ST_EINVAL, "st_hist_create: null output");
    *out = nullptr;
    if (type != ST_HIST_LIN && type != ST_HIST_LOG)
        return set_error(ST_EINVAL, "Unknown histogram type %d: available types are lin (0) and log (1)",
                         type);
    if (nbins < 1 || nbins > (1 << 24))
        return set_error(ST_EINVAL, "st_hist_create: nbins must be in [1, 2^24] (got %d)", nbins);
    st_hist *h = new (std::nothrow) st_hist;
    if (!h) return set_error(ST_ENOMEM, "st_hist_create: out of host memory");
    h->type = type;
    h->nbins = nbins;
    h->lo = lo;
    h->hi = hi;
    h->edges_host.resize((size_t)nbins + 1);
    make_edges(type, lo, hi, nbins, h->edges_host.data());
    h->guess.type = type;
    h->guess.nbins = nbins;
    if (type == ST_HIST_LIN) {
        const float wf = (hi - lo) / (float)nbins;
        h->guess.origin = (double)lo;
        h->guess.inv_width = 1.0 / (double)wf;
    } else {
        const float bo = (float)std::log10((double)lo);
        const float to = (float)std::log10((double)hi);
        const float wf = (to - bo) / (float)nbins;
        h->guess.origin = (double)bo;
        h->guess.inv_width = 1.0 / (double)wf;
    }
    if (!std::isfinite(h->guess.inv_width)) h->guess.inv_width = 0.0;
    if (!std::isfinite(h->guess.origin)) h->guess.origin = 0.0;
    h->guess.log2_origin = (float)(h->guess.origin * 3.321928094887362);
    h->guess.log2_inv_width = (float)(h->guess.inv_width / 3.321928094887362);
    h->guess.e0 = h->edges_host[0];
    h->guess.en = h->edges_host[(size_t)nbins];
    h->stream = default_stream();
    hipError_t e = dev_malloc((void **)&h->edges_dev, ((size_t)nbins + 1) * sizeof(double));
    if (e == hipSuccess) e = dev_malloc((void **)&h->bins_dev, (size_t)nbins * sizeof(uint64_t));
    if (e == hipSuccess)
        e = hipMemcpy(h->edges_dev, h->edges_host.data(), ((size_t)nbins + 1) * sizeof(double),
                      hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(h->bins_dev, 0, (size_t)nbins * sizeof(uint64_t));
    if (e != hipSuccess) {
        st_hist_destroy(h);
        return set_error(hip_status(e), "st_hist_create: %s", hipGetErrorString(e));
    }
    h->host_counts.resize(nbins);
    *out = h;
    return ST_OK;
}

int st_hist_set_stream(st_hist *h, void *hip_stream)
{
    if (!h) return set_error(ST_EINVAL, "st_hist_set_stream: null histogram");
    int rc = queue_flush_producers(h);
    if (rc) return rc;
    hipStream_t next = hip_stream ? (hipStream_t)hip_stream : default_stream();
    // work still queued on the old stream (a clear of the counters) is done
    // before the new stream can touch them
    if (next != h->stream && h->stream) ST_HIP_TRY(hipStreamSynchronize(h->stream));
    h->stream = next;
    return ST_OK;
}

int st_hist_nbins(const st_hist *h, int32_t *nbins)
{
    if (!h || !nbins) return set_error(ST_EINVAL, "st_hist_nbins: null argument");
    *nbins = h->nbins;
    return ST_OK;
}

int st_hist_edges(const st_hist *h, double *edges)
{
    if (!h || !edges) return set_error(ST_EINVAL, "st_hist_edges: null argument");
    std::memcpy(edges, h->edges_host.data(), h->edges_host.size() * sizeof(double));
    return ST_OK;
}

// Histogram::AllocateTrackProcess (Histogram.cu:102-112): size the sort
// buffers for the pairs' capacity.  Buffers are kept and only grown, instead
// of being allocated and freed per track.
int st_hist_begin_track(st_hist *h, const st_pairs *pairs)
{
    if (!h || !pairs) return set_error(ST_EINVAL, "st_hist_begin_track: null argument");
    if (pairs->capacity < 0 || pairs->capacity > 0x7FFFFFFF)
        return set_error(ST_EINVAL, "st_hist_begin_track: bad pairs capacity");
    int rc = queue_flush_producers(h);
    if (rc) return rc;
    h->capacity = pairs->capacity;
    if (h->capacity > kLdsCap) {
        rc = ensure_sort_buffers(h, h->capacity);
        if (rc) return rc;
    }
    // the per-track counters start at zero (Histogram.cu:139-140); they
    // already are unless something was counted since they were cleared, and
    // they keep the counts carried from tracks ended without a host copy
    if (!h->bins_zero && !h->carry) {
        zero_u64_kernel<<<grid_for(h->nbins, 256, 1024), 256, 0, h->stream>>>(
            (unsigned long long *)h->bins_dev, h->nbins);
        ST_HIP_TRY(hipGetLastError());
        h->bins_zero = true;
    }
    h->in_track = true;
    return ST_OK;
}

// Histogram::SortReduceAndAddToHistogram (Histogram.cu:114-123,147-171).
int st_hist_add(st_hist *h, const st_pairs *pairs)
{
    if (!h || !pairs) return set_error(ST_EINVAL, "st_hist_add: null argument");
    if (!h->in_track) return set_error(ST_ESTATE, "st_hist_add before st_hist_begin_track");
    if (pairs->capacity > h->capacity)
        return set_error(ST_EINVAL, "st_hist_add: pairs capacity %lld above the begun %lld",
                         (long long)pairs->capacity, (long long)h->capacity);
    // pairs whose oversample is queued: this histogram consumes it when the
    // producer runs its queue
    if (pairs->volume) {
        st_vcs *m = queue_producer(pairs->volume);
        if (m && m->queue_active && m->queue_count > 0 && m->queue_pairs.volume == pairs->volume) {
            m->queue_consumers.back().push_back(h);
            if (std::find(h->producers.begin(), h->producers.end(), m) == h->producers.end())
                h->producers.push_back(m);
            return ST_OK;
        }
    }
    return hist_add_pairs(h, pairs, h->stream);
}

// Histogram::FreeTrackProcess -> AccumulateCPUHistogram (Histogram.cu:125-141,
// 177-186): host_bins[i + 1] += count[i].
int st_hist_end_track(st_hist *h, double *host_bins)
{
    if (!h) return set_error(ST_EINVAL, "st_hist_end_track: null histogram");
    if (!h->in_track) return set_error(ST_ESTATE, "st_hist_end_track before st_hist_begin_track");
    int rc = queue_flush_producers(h);
    if (rc) return rc;
    if (!host_bins) {
        // the counts stay summed in the device counters
        h->carry = true;
        h->in_track = false;
        return ST_OK;
    }
    ST_HIP_TRY(hipMemcpyAsync(h->host_counts.data(), h->bins_dev, (size_t)h->nbins * sizeof(uint64_t),
                              hipMemcpyDeviceToHost, h->stream));
    ST_HIP_TRY(hipStreamSynchronize(h->stream));
    for (int i = 0; i < h->nbins; ++i) host_bins[i + 1] += (double)h->host_counts[i];
    zero_u64_kernel<<<grid_for(h->nbins, 256, 1024), 256, 0, h->stream>>>(
        (unsigned long long *)h->bins_dev, h->nbins);
    ST_HIP_TRY(hipGetLastError());
    h->bins_zero = true;
    h->carry = false;
    h->in_track = false;
    return ST_OK;
}

int st_hist_fetch(st_hist *h, uint64_t *counts, int32_t reset)
{
    if (!h || !counts) return set_error(ST_EINVAL, "st_hist_fetch: null argument");
    int rc = queue_flush_producers(h);
    if (rc) return rc;
    ST_HIP_TRY(hipMemcpyAsync(counts, h->bins_dev, (size_t)h->nbins * sizeof(uint64_t),
                              hipMemcpyDeviceToHost, h->stream));
    ST_HIP_TRY(hipStreamSynchronize(h->stream));
    if (reset) {
        zero_u64_kernel<<<grid_for(h->nbins, 256, 1024), 256, 0, h->stream>>>(
            (unsigned long long *)h->bins_dev, h->nbins);
        ST_HIP_TRY(hipGetLastError());
        ST_HIP_TRY(hipStreamSynchronize(h->stream));
        h->bins_zero = true;
        h->carry = false;
    }
    return ST_OK;
}

void st_hist_destroy(st_hist *h)
{
    if (!h) return;
    (void)queue_flush_producers(h);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    (void)hipFree(h->edges_dev);
    (void)hipFree(h->bins_dev);
    (void)hipFree(h->keys_a);
    (void)hipFree(h->keys_b);
    (void)hipFree(h->vals_a);
    (void)hipFree(h->vals_b);
    (void)hipFree(h->os_status);
    (void)hipFree(h->digit_counts);
    (void)hipFree(h->os_ticket);
    delete h;
}

}  // extern "C"
