// batch_kernel.hpp -- device side of the batched form (included by
// batch.hip only): the fused persistent kernel, its argument and control
// structures, and the per-pass reset kernel.  Kept apart from the host code
// so that the kernel's source hash (bench.py) names the device code alone.
//
// Each workgroup takes (track, oversample) segments from a global ticket
// (largest tracks first, all oversamples of a track together so the track
// stays in L2/MALL).  Per segment it culls by the track's x/y bounds, derives
// the box of sphere indices the segment can reach (32-bit table keys local to
// it), finds the in-voxel block range with two 64-way searches over block
// starts, and streams the range in chunks of 256 points.  In-sphere points
// are accumulated per sphere in an LDS open-addressing table as fixed-point
// integers (exact and independent of order, so results are bit-reproducible
// and do not depend on launch parameters).  When the table fills past a
// quarter, every entry below the layer of the last point scored is complete:
// those are converted to lineal energy and binned into LDS-privatised
// histograms; the rest are kept.
#pragma once

#include <type_traits>

#include "st_common.hpp"
#include "st_internal.hpp"
#include "st_device.hpp"
#include "layout.hpp"

namespace st {
namespace {

constexpr int kBT = 256;                    // threads per workgroup
constexpr int kMaxGeoms = 16;
constexpr int kMaxHists = 2;

constexpr double kTwo24 = 16777216.0;
constexpr double kTwo40 = 1099511627776.0;
constexpr double kTwo32 = 4294967296.0;
constexpr double kTwoM24 = 1.0 / 16777216.0;
constexpr double kTwoM64 = 1.0 / 18446744073709551616.0;
constexpr double kTwoM32 = 1.0 / 4294967296.0;


struct GeomDev {
    double H, L, D, invD, r2, gso, denom, denom_inv;
    double yscale;      // denom_inv * 2^32 (ST_F32 fixed point)
    double qc;          // gso / D: lattice units of the voxel centre offset
    double r2q;         // r^2 / D^2
    // ST_F64 thresholds in lattice units (as BlockF32's, with the FP64
    // rounding bound dl of the geometry: the data carry no FP32 rounding)
    double fm, mrgq;
    unsigned long long N, N2;
};


struct HistDev {
    const double *edges;
    unsigned long long *bins;
    BinGuess guess;
    int lds_offset;     // in u32 units from the histogram LDS base
};

// Per-pass control block in device memory, cleared by batch_reset_kernel
// before every pass (a kernel, not a memset: the pair is captured in a graph).
//   stats[0] points scored, [1] exact FP64 rechecks, [2] in-sphere points,
//   [3] fixed-point energy inserted into the tables, [4] fixed-point energy of
//   the entries retired (binned or outside every histogram), [5] table entries
//   created, [6] entries retired.  [3]/[4] are sums modulo 2^64 (of the ST_F64
//   hi word plus the lo word times an odd constant); a pass that loses no
//   entry has [3] == [4] and [5] == [6].
constexpr int kStats = 7;
struct BatchCtl {
    unsigned long long ticket;
    unsigned long long rec_count;
    // 1: a table filled up, 2: a flush found more entries than its list bound,
    // 4: an in-sphere point's key outside its segment's key box
    int error;
    int pad;
    unsigned long long stats[kStats];
};
constexpr unsigned long long kBalanceMul = 0x9E3779B97F4A7C15ull;

struct BatchArgs {
    // library
    const TrackDesc *tracks;
    const double *x64, *y64, *z64, *e64;
    const float4 *p32;          // ST_F32: (dx, dy, dz, edep) per point
    const double *zfirst;
    const BlockF32 *blocks;
    const GeomDev *geoms;
    // work
    const int32_t *order;
    int64_t n_seg;
    int32_t n_os;
    int32_t shift_xy, shift_z;
    uint32_t track_index_base;
    uint64_t seed;
    BatchCtl *ctl;
    // histograms
    HistDev hist[kMaxHists];
    int n_hists;
    int hist_lds_words;     // 0: bins go straight to the device counters
    int hist_copies;        // LDS copies of the histograms (wave w counts into copy w % copies)

    // recording (test hook)
    const struct RecArgs *rec;     // null: off
};

// Recording of every binned sphere (test hook), in device memory so that the
// kernel keeps one pointer, not five, through its loops.
struct RecArgs {
    int64_t cap;
    int64_t *seg;                  // the segment's ticket
    unsigned long long *key;       // the local key
    unsigned long long *box;       // 3 words per record: kx0 | ky0 << 32, kz0 | nx << 32, ny
    double *sum;
};

__device__ __forceinline__ int uniform_i32(int v)
{
    return __builtin_amdgcn_readfirstlane(v);
}

__device__ __forceinline__ unsigned uniform_u32(unsigned v)
{
    return (unsigned)__builtin_amdgcn_readfirstlane((int)v);
}

__device__ __forceinline__ long long uniform_i64(long long v)
{
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)v);
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)((unsigned long long)v >> 32));
    return (long long)(((unsigned long long)hi << 32) | lo);
}

// A record through the scalar cache (read-only during the kernel, address
// uniform across the workgroup): its fields land in scalar registers.
template <typename T>
__device__ __forceinline__ T load_uniform(const T *p)
{
    static_assert(sizeof(T) % 8 == 0 && alignof(T) >= 8, "records of 8-byte words");
    typedef const __attribute__((address_space(4))) unsigned long long *ConstPtr;
    const ConstPtr q = (ConstPtr)(const void *)p;
    unsigned long long w[sizeof(T) / 8];
#pragma unroll
    for (size_t i = 0; i < sizeof(T) / 8; ++i) w[i] = q[i];
    T r;
    __builtin_memcpy(&r, w, sizeof(T));
    return r;
}

// Block record through the scalar cache (read-only during the kernel, index
// uniform across the workgroup).
__device__ __forceinline__ BlockF32 load_block(const BlockF32 *p, int64_t idx)
{
    typedef const __attribute__((address_space(4))) double *ConstPtr;
    const ConstPtr q = (ConstPtr)(const void *)(p + idx);
    BlockF32 r;
    r.oqx = q[0];
    r.oqy = q[1];
    r.oqz = q[2];
    r.dl = q[3];
    r.fm = q[4];
    r.mrgq = q[5];
    return r;
}

// the reference's key: kx + ky * N + kz * N^2 (mod 2^64, its unsigned sum)
__device__ __forceinline__ unsigned long long lattice_key(unsigned kx, unsigned ky, unsigned kz, unsigned N)
{
    const unsigned long long kyz = (unsigned long long)kz * N + ky;
    return kyz * N + kx;
}

// The sphere table holds 32-bit keys local to a slab of the segment's box:
//   key = ((kz - kz0) * ny + (ky - ky0)) * nx + (kx - kx0)
// over the box [kx0, kx0 + nx) x [ky0, ky0 + ny) x [kz0, kz0 + nz) of sphere
// indices that the segment's points can reach (from the track's bounds and
// the shift), with nx * ny * nz <= 2^32 - 2 so that every key is below
// kEmpty32.  Within a segment the map is one-to-one and keeps the layer-major
// order of the reference's key, so "key below the first layer still to come"
// is the same test in both.  A segment whose box exceeds 2^32 - 2 spheres is
// scored in slabs of layers, each with its own box and a full flush at its end.
constexpr unsigned kEmpty32 = 0xFFFFFFFFu;
constexpr int kLayerUnsure = -0x7FFFFFFF;   // a layer to recompute exactly (|layer| < 2^30 otherwise)
constexpr unsigned long long kMaxKeys = 0xFFFFFFFEull;   // keys per slab box

struct SlabBox {
    unsigned kx0, ky0, kz0;   // global index of local (0, 0, 0)
    unsigned nx, ny, nz;      // extent (nz: this slab's layers)
    unsigned nxy;             // nx * ny
};


// ---------------------------------------------------------------------------
// exact FP64 membership (the reference's operations, rounding as written)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double nearest_index(double a, double D, double invD)
{
    const double q = a * invD;
    double k = rint(q);
    const double f = fabs(q - k);
    if (f > 0.5 - (fabs(q) * 1e-15 + 1e-300)) k = rint(a / D);
    return k;
}

// What the exact path needs of the current segment and slab.  The kernel
// keeps it in LDS: the path is rare, and its constants then hold no
// registers through the point loop.
struct ExactParams {
    double H, gso, D, invD, r2;       // geometry
    double sx, sy, sz;                // shift
    unsigned N;
    unsigned kx0, ky0, nx, ny;        // the segment's key box in x and y
    unsigned kz_lo, nz_all, multi;    // its layers, and whether it is scored in slabs
    unsigned kz0, nz;                 // the current slab's layers
};

// On success kx, ky, kz are the reference's sphere indices, in [0, N]: an
// index equals N only for a point within rounding of the voxel's upper face.
// z: the stored coordinate (H' is added here, as the reference adds it once).
__device__ __forceinline__ bool exact_member(double x, double y, double z, const ExactParams &g, unsigned &ix,
                                             unsigned &iy, unsigned &iz)
{
    const double xs = x + g.sx;
    const double ys = y + g.sy;
    const double zs = (z + g.H) + g.sz;
    if (!(fabs(xs) < g.H && fabs(ys) < g.H && fabs(zs) < g.H)) return false;
    const double ax = xs - g.gso, ay = ys - g.gso, az = zs - g.gso;
    const double kx = nearest_index(ax, g.D, g.invD);
    const double ky = nearest_index(ay, g.D, g.invD);
    const double kz = nearest_index(az, g.D, g.invD);
    const double dx = (kx * g.D) - ax;
    const double dy = (ky * g.D) - ay;
    const double dz = (kz * g.D) - az;
    const double dist = ((dx * dx) + (dy * dy)) + (dz * dz);
    if (!(dist <= g.r2)) return false;
    // in-voxel indices lie in [0, N] with N < 2^31
    ix = (unsigned)(int)kx;
    iy = (unsigned)(int)ky;
    iz = (unsigned)(int)kz;
    return true;
}

// ---------------------------------------------------------------------------
// fixed-point accumulation
// ---------------------------------------------------------------------------
// ST_F64: y = hi * 2^-24 + lo * 2^-64 with hi = floor(y * 2^24); both parts
// exact (lo rounded at 2^-65), sums exact while they stay in int64.
// ST_F32: y = acc * 2^-32.
template <int MODE>
struct Acc;

template <>
struct Acc<ST_F64> {
    static constexpr int words = 2;
    // y * 2^24 < 2^51 (kMaxPointY64)
    __device__ static void split(double y, long long &a0, long long &a1)
    {
        const double t = y * kTwo24;
        const double fh = floor(t);
        a0 = rint_to_i64(fh);
        a1 = rint_to_i64((t - fh) * kTwo40);
    }
    __device__ static double value(unsigned long long a0, unsigned long long a1)
    {
        return (double)(long long)a0 * kTwoM24 + (double)(long long)a1 * kTwoM64;
    }
    // conservation word of an entry or a point (see BatchCtl)
    __device__ static unsigned long long balance(unsigned long long a0, unsigned long long a1)
    {
        return a0 + a1 * kBalanceMul;
    }
};

template <>
struct Acc<ST_F32> {
    static constexpr int words = 1;
    __device__ static double value(unsigned long long a0, unsigned long long)
    {
        return (double)(long long)a0 * kTwoM32;
    }
    __device__ static unsigned long long balance(unsigned long long a0, unsigned long long) { return a0; }
};

// Home slot in [0, SLOTS) of a key: multiplicative (Fibonacci) hash scaled
// to the table size (measured faster than the same hash in FP64).
template <int SLOTS>
__device__ __forceinline__ unsigned int slot_hash(unsigned key)
{
    return (unsigned)(((unsigned long long)(key * 0x9E3779B1u) * SLOTS) >> 32);
}

template <int SLOTS>
__device__ __forceinline__ unsigned int next_slot(unsigned int h)
{
    return (h + 1 == (unsigned)SLOTS) ? 0u : h + 1;
}

// ---------------------------------------------------------------------------
// 64-way lower bound by one wave: first i in [0, n) with pred(i), else n.
// ---------------------------------------------------------------------------
template <typename Pred>
__device__ int wave_lower_bound(int n, Pred pred)
{
    const int lane = threadIdx.x & 63;
    int lo = 0, hi = n;
    while (lo < hi) {
        const int len = hi - lo;
        const int step = (len + 63) / 64;
        const int i = lo + lane * step;
        const bool p = (i < hi) ? pred(i) : true;
        const unsigned long long mask = __ballot(p);
        if (mask == 0ull) {
            lo = lo + 63 * step + 1;
        } else {
            const int f = __ffsll((long long)mask) - 1;
            const int nhi = lo + f * step;
            const int nlo = (f == 0) ? lo : lo + (f - 1) * step + 1;
            lo = nlo;
            hi = min(nhi, hi);
        }
    }
    return lo;
}

// Iterations between occupancy checks: two where the table holds the
// largest survivor set plus two iterations' points at one point per thread
// (1472 and 1536 slots), else one.
constexpr int check_every(int slots, int ppt)
{
    return ((slots == 1472 || slots == 1536) && ppt == 1) ? 2 : 1;
}

// Capacity of each wave's part of the slot list.  A wave lists at most the
// survivors it re-inserted in the last flush (no more than the entries it
// walked: kLP rows of 64) or, at an occupancy check that did not flush, fewer
// than the flush threshold; then at most the points of the iterations until
// the next check.
constexpr int list_region(int slots, int ppt)
{
    const int flush_at = slots / 4;
    const int checks = check_every(slots, ppt);
    const int window = (int)kMaxWindow;
    const int max_entries = (flush_at - 1 > window ? flush_at - 1 : window) + checks * ppt * kBlk;
    const int rows = (max_entries + kBT - 1) / kBT;
    const int walked = 64 * rows < window ? 64 * rows : window;
    return (flush_at - 1 > walked ? flush_at - 1 : walked) + checks * ppt * 64;
}

// ---------------------------------------------------------------------------
// the persistent fused kernel
// ---------------------------------------------------------------------------
// kPPT: library blocks (points per thread) loaded together per iteration.
// Waves per SIMD the kernels are compiled for: more waves hide more of the
// LDS-atomic and load latency, and every step up measured faster despite the
// registers spilled (outside the point path).  ST_F32 at one point per
// thread: 1536 slots five (96 registers per lane), 1472 slots six (80; 26 KB
// of LDS per workgroup with 1000 histogram bins), 1280 slots seven (72;
// 22.7 KB).  ST_F64 with 1536 slots at two points per thread: four (128, no
// spills; 39 KB of LDS).
template <int MODE, int SLOTS, int kPPT>
constexpr int waves_per_eu()
{
    if (MODE == ST_F64) return (SLOTS == 1536 && kPPT == 2) ? 4 : 1;
    if (kPPT != 1) return 1;
    return SLOTS == 1280 ? 7 : SLOTS == 1472 ? 6 : SLOTS == 1536 ? 5 : 1;
}

template <int MODE, int SLOTS, int kPPT>
__global__ __launch_bounds__(kBT) __attribute__((amdgpu_waves_per_eu(waves_per_eu<MODE, SLOTS, kPPT>())))
void batch_kernel(BatchArgs a)
{
    static_assert(SLOTS == 1280 || SLOTS == 1472 || SLOTS == 1536 || SLOTS == 2048 || SLOTS == 4096, "table sizes");
    // A flush starts once the table is a quarter full: a low load factor
    // keeps the linear-probe chains (dependent LDS round trips, paid by the
    // whole wave for its longest lane) short; a quarter was fastest on MI355X
    // (an eighth flushes too often, a half probes too long).
    constexpr int kFlushAt = SLOTS / 4;
    // iterations between occupancy checks (each check is a workgroup barrier)
    constexpr int kCheck = check_every(SLOTS, kPPT);
    // occupancy never exceeds max(flush threshold, survivors) + the points of
    // the iterations between two checks, kept under 90 % of the table
    static_assert((kFlushAt > kMaxWindow ? kFlushAt : kMaxWindow) + kCheck * kPPT * kBlk <= SLOTS * 9 / 10,
                  "the table must hold the survivors of a flush plus the points between two checks");
    // entries at a flush: those kept by the previous flush (at most kMaxWindow)
    // or below the threshold, plus the points of the iterations since the
    // last check
    constexpr int kMaxEntries = (kFlushAt - 1 > kMaxWindow ? kFlushAt - 1 : (int)kMaxWindow) + kCheck * kPPT * kBlk;
    constexpr int kLP = (kMaxEntries + kBT - 1) / kBT;   // list entries per thread
    constexpr int kWords = Acc<MODE>::words;
    extern __shared__ unsigned long long lds[];
    unsigned long long *tacc0 = lds;
    unsigned long long *tacc1 = tacc0 + SLOTS;  // used when kWords == 2
    unsigned int *tkey = reinterpret_cast<unsigned int *>(tacc0 + SLOTS * kWords);
    // slots in use, in insertion order: one part of kRegion entries per wave
    constexpr int kRegion = list_region(SLOTS, kPPT);
    unsigned short *tlist = reinterpret_cast<unsigned short *>(tkey + SLOTS);
    unsigned int *lhist = reinterpret_cast<unsigned int *>(tlist + (kBT / 64) * kRegion);

    __shared__ long long s_seg;
    __shared__ int s_range[2];
    // the exact path's constants (read only by that rare path)
    __shared__ ExactParams s_ex;
    // the histogram descriptors, read by the flushes from LDS (they then
    // hold no scalar registers through the point loop)
    __shared__ HistDev s_hist[kMaxHists];
    // entries each wave listed since the last flush, and a lower bound on the
    // layer of every later point, published at each occupancy check into
    // alternating halves
    __shared__ int s_wcnt[2][kBT / 64];
    __shared__ int s_layer[2];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave_first = uniform_i32(tid & ~63);   // first thread of this wave
    // per-thread statistics, summed at exit
    unsigned int n_flagged = 0, n_in = 0;
    // points scored by the workgroup (uniform): the lengths of its ranges
    unsigned long long n_scored = 0;
    // conservation: energy words into and out of the table, entries created
    // and retired (a lost entry breaks one of the two equalities)
    unsigned long long bal_in = 0, bal_out = 0;
    // per wave (uniform): entries it listed from new points, entries retired
    unsigned long long n_created = 0, n_retired = 0;
    for (int s = tid; s < SLOTS; s += kBT) {
        tkey[s] = kEmpty32;
        tacc0[s] = 0;
        if (kWords == 2) tacc1[s] = 0;
    }
    for (int w = tid; w < a.hist_lds_words * a.hist_copies; w += kBT) lhist[w] = 0;
    if (tid < kMaxHists) s_hist[tid] = a.hist[tid];
    // this wave's copy of the privatised histograms (more copies, fewer
    // waves contending for one hot bin's LDS word)
    // (the wave index is uniform: the offset stays in a scalar register)
    unsigned int *const lhist_w = lhist + uniform_i32(((tid >> 6) % a.hist_copies) * a.hist_lds_words);
    __syncthreads();
    // upper bound on the counts added to any LDS bin since the last spill
    long long binned = 0;
    // entries this wave listed since the last flush (its part of the list),
    // and the half of s_wcnt the next check uses (both uniform)
    int my_cnt = 0;
    int par = 0;
    // every wave's list count at the last check (uniform)
    int wcnt[kBT / 64] = {};
    // this wave's part of the list (a uniform offset)
    unsigned short *const tlist_w = tlist + (tid >> 6) * kRegion;
    // the current slab's key box (uniform)
    SlabBox box{0u, 0u, 0u, 1u, 1u, 1u, 1u};

    // Appends the slots that lanes newly occupied (slot >= 0) to this wave's
    // part of the list (positions from its own count: no atomics).
    // new_entries: the slots hold new entries (not survivors of a flush)
    auto append = [&](const int *slot, auto count_tag, bool new_entries) {
        constexpr int kN = decltype(count_tag)::value;
        unsigned long long m[kN];
        int cnt = 0;
#pragma unroll
        for (int j = 0; j < kN; ++j) {
            m[j] = __ballot(slot[j] >= 0);
            cnt += __popcll(m[j]);
        }
        if (cnt == 0) return;
        int base = my_cnt;
        my_cnt += cnt;
        if (new_entries) n_created += (unsigned long long)cnt;
        if (my_cnt > kRegion) {   // cannot happen (the bound above); reported, never written past
            if (lane == 0) atomicOr(&a.ctl->error, 2);
        }
#pragma unroll
        for (int j = 0; j < kN; ++j) {
            const int below = __builtin_amdgcn_mbcnt_hi((unsigned)(m[j] >> 32),
                                                        __builtin_amdgcn_mbcnt_lo((unsigned)m[j], 0u));
            if (slot[j] >= 0 && base + below < kRegion) tlist_w[base + below] = (unsigned short)slot[j];
            base += __popcll(m[j]);
        }
    };

    // Flushes entries with key < bound (all entries when bound == kEmpty32):
    // complete sums are converted to lineal energy and binned; the rest are
    // re-inserted into the cleared table.  Threads walk the dense list of
    // used slots, so the work follows the entries, not the table size, and
    // the bin edges of two entries are loaded before either is compared.
    // after_check: called right after an occupancy check, whose barrier
    // already ordered every wave's appends and whose counts are in wcnt.
    auto flush = [&](unsigned bound, bool after_check) __attribute__((always_inline)) {
        // raised issue priority while the workgroup's waves are tied together
        // by the flush barriers (measured: about 1 % on cfg 4)
        __builtin_amdgcn_s_setprio(2);
        if (!after_check) {
            // the counts as an occupancy check publishes them
            if (lane == 0) s_wcnt[par][tid >> 6] = my_cnt;
            __syncthreads();
#pragma unroll
            for (int w = 0; w < kBT / 64; ++w) wcnt[w] = s_wcnt[par][w];
            par ^= 1;
        }
        // list position idx in [0, n) lies in wave w's part at idx - start[w]
        int start[kBT / 64];
        int n = 0;
#pragma unroll
        for (int w = 0; w < kBT / 64; ++w) {
            start[w] = n;
            n += min(wcnt[w], kRegion);
        }
        if (n > kLP * kBT) {   // cannot happen (the bound above); reported, never read past the list
            if (tid == 0) atomicOr(&a.ctl->error, 2);
            n = kLP * kBT;
        }
        unsigned kk[kLP];
        unsigned long long k0[kLP], k1[kLP];
#pragma unroll
        for (int r = 0; r < kLP; ++r) kk[r] = kEmpty32;
#pragma unroll
        for (int r0 = 0; r0 < kLP; r0 += 2) {
            // rows past the list: nothing to read, convert or bin (n is uniform)
            if (r0 * kBT >= n) break;
            double vv[2];
            bool em[2];
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int r = r0 + u;
                em[u] = false;
                vv[u] = 0.0;
                if (r >= kLP) continue;
                const int idx = tid + r * kBT;
                if (idx < n) {
                    int part = 0;
#pragma unroll
                    for (int w = 1; w < kBT / 64; ++w) part += (idx >= start[w]) ? 1 : 0;
                    int first = start[0];
#pragma unroll
                    for (int w = 1; w < kBT / 64; ++w) first = (part == w) ? start[w] : first;
                    const int sl = min((int)tlist[part * kRegion + (idx - first)], SLOTS - 1);
                    // read and clear in one exchange each (the slot is this
                    // thread's alone until the barrier)
                    const unsigned k = atomicExch(&tkey[sl], kEmpty32);
                    const unsigned long long v0 = atomicExch(&tacc0[sl], 0ull);
                    const unsigned long long v1 = (kWords == 2) ? atomicExch(&tacc1[sl], 0ull) : 0ull;
                    em[u] = k < bound;
                    if (!em[u]) kk[r] = k;
                    k0[r] = v0;
                    k1[r] = v1;
                    vv[u] = Acc<MODE>::value(v0, v1);
                    if (em[u]) bal_out += Acc<MODE>::balance(v0, v1);
                    if (em[u] && a.rec != nullptr) {
                        const RecArgs rc = load_uniform(a.rec);
                        const unsigned long long p = atomicAdd(&a.ctl->rec_count, 1ull);
                        if ((int64_t)p < rc.cap) {
                            // test hook only: the ticket (the host maps it to
                            // the segment id)
                            rc.seg[p] = (int64_t)s_seg;
                            rc.key[p] = k;
                            rc.box[3 * p] = s_ex.kx0 | ((unsigned long long)s_ex.ky0 << 32);
                            rc.box[3 * p + 1] = s_ex.kz0 | ((unsigned long long)s_ex.nx << 32);
                            rc.box[3 * p + 2] = s_ex.ny;
                            rc.sum[p] = vv[u];
                        }
                    }
                }
            }
#pragma unroll
            for (int h = 0; h < kMaxHists; ++h) {
                if (h >= a.n_hists) break;
                const HistDev &hd = s_hist[h];
                int bb[2];
                double lo[2], hi[2];
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    // closed-form guess, clamped into [0, nbins - 1]: always a valid load
                    const double v = vv[u];
                    float t;
                    if (hd.guess.type == ST_HIST_LIN) t = (float)((v - hd.guess.origin) * hd.guess.inv_width);
                    else t = (__log2f((float)v) - hd.guess.log2_origin) * hd.guess.log2_inv_width;
                    bb[u] = (t > 0.0f) ? (int)fminf(t, (float)(hd.guess.nbins - 1)) : 0;
                    lo[u] = hd.edges[bb[u]];
                    hi[u] = hd.edges[bb[u] + 1];
                }
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    if (!em[u]) continue;
                    const double v = vv[u];
                    int b = bb[u];
                    if (!(v >= lo[u] && v < hi[u])) b = bin_fast(hd.edges, hd.guess, v);
                    if (b >= 0) {
                        if (a.hist_lds_words) atomicAdd(&lhist_w[hd.lds_offset + b], 1u);
                        else atomicAdd(&hd.bins[b], 1ull);
                    }
                }
            }
        }
        // every entry is read and cleared before any is re-inserted
        __syncthreads();
        my_cnt = 0;
        // entries walked by this wave's lanes (rows of 64 list positions
        // from wave_first, n uniform: scalar arithmetic)
        unsigned long long walked = 0;
#pragma unroll
        for (int r = 0; r < kLP; ++r) walked += (unsigned long long)min(max(n - (r * kBT + wave_first), 0), 64);
        // The few incomplete entries (a z-window of spheres) are re-inserted
        // into the cleared table; most waves have none and skip the step.
        // Waves that have none go back to their points at once: a new point
        // may insert a survivor's key first, and the survivor then adds its
        // sums to that entry (listed by the point's wave).
        bool keep = false;
#pragma unroll
        for (int r = 0; r < kLP; ++r) keep |= kk[r] != kEmpty32;
        if (__ballot(keep) != 0ull) {
            // first probes of every row at once (independent exchanges, one
            // wait), then the rare collisions slot by slot
            int placed[kLP];
            unsigned prev[kLP];
#pragma unroll
            for (int r = 0; r < kLP; ++r) {
                placed[r] = -1;
                prev[r] = kEmpty32;
                if (kk[r] != kEmpty32) {
                    placed[r] = (int)slot_hash<SLOTS>(kk[r]);
                    prev[r] = atomicCAS(&tkey[placed[r]], kEmpty32, kk[r]);
                }
            }
#pragma unroll
            for (int r = 0; r < kLP; ++r) {
                if (kk[r] == kEmpty32) continue;
                unsigned int h = (unsigned)placed[r];
#pragma unroll 1
                for (int probe = 1; probe < SLOTS && prev[r] != kEmpty32 && prev[r] != kk[r]; ++probe) {
                    h = next_slot<SLOTS>(h);
                    prev[r] = atomicCAS(&tkey[h], kEmpty32, kk[r]);
                }
                // listed again only where it made a new entry
                placed[r] = prev[r] == kEmpty32 ? (int)h : -1;
                // (a full table, excluded by its size, would drop the sums,
                // which the conservation check reports)
                if (prev[r] == kEmpty32 || prev[r] == kk[r]) {
                    atomicAdd(&tacc0[h], k0[r]);
                    if (kWords == 2) atomicAdd(&tacc1[h], k1[r]);
                }
            }
            unsigned long long kept = 0;
#pragma unroll
            for (int r = 0; r < kLP; ++r) kept += __popcll(__ballot(placed[r] >= 0));
            walked -= kept;
            append(placed, std::integral_constant<int, kLP>(), false);
        }
        n_retired += walked;
        __builtin_amdgcn_s_setprio(0);
    };

    for (;;) {
        // segment set-up (ticket, shift, block range: barriers) at raised
        // priority as well, back to normal for the point loop
        __builtin_amdgcn_s_setprio(2);
        if (tid == 0) s_seg = (long long)atomicAdd(&a.ctl->ticket, 1ull);
        __syncthreads();
        const int64_t seg = uniform_i64(s_seg);
        __syncthreads();
        if (seg >= a.n_seg) break;
        const int32_t t = a.order[seg / a.n_os];
        const int32_t os = (int32_t)(seg % a.n_os);
        const TrackDesc td = load_uniform(a.tracks + t);
        const GeomDev g = load_uniform(a.geoms + td.geom);
        const Shift sh = shift_for(a.seed, a.track_index_base + (uint32_t)t, (uint32_t)os, a.shift_xy,
                                   a.shift_z, g.L, g.H);
        // x/y bounds cull: fl(x + s) is monotone in x
        if (!((td.xmax + sh.x) > -g.H && (td.xmin + sh.x) < g.H && (td.ymax + sh.y) > -g.H &&
              (td.ymin + sh.y) < g.H))
            continue;
        // The segment's box of sphere indices.  fl(fl(c + s) - gso) / D is
        // monotone in the coordinate c, so rint of it lies between floor and
        // ceil of the bounds' quotients, and the exact path's nearest_index
        // (which may round a half-way quotient the other way) within one more.
        const unsigned N = (unsigned)g.N;
        const double fN1 = (double)(N - 1u);
        const double zmin = load_uniform(a.zfirst + td.block0);
        const double bx0 = fmax(0.0, floor(((td.xmin + sh.x) - g.gso) * g.invD) - 1.0);
        const double bx1 = fmin(fN1, ceil(((td.xmax + sh.x) - g.gso) * g.invD) + 1.0);
        const double by0 = fmax(0.0, floor(((td.ymin + sh.y) - g.gso) * g.invD) - 1.0);
        const double by1 = fmin(fN1, ceil(((td.ymax + sh.y) - g.gso) * g.invD) + 1.0);
        const double bz0 = fmax(0.0, floor((((zmin + g.H) + sh.z) - g.gso) * g.invD) - 1.0);
        const double bz1 = fmin(fN1, ceil((((td.zmax + g.H) + sh.z) - g.gso) * g.invD) + 1.0);
        if (!(bx0 <= bx1 && by0 <= by1 && bz0 <= bz1)) continue;
        // (uniform values computed by the vector unit: moved to scalar registers)
        const unsigned kx0 = uniform_u32((unsigned)bx0), ky0 = uniform_u32((unsigned)by0);
        const unsigned kz_lo = uniform_u32((unsigned)bz0), kz_hi = uniform_u32((unsigned)bz1);
        const unsigned nx = uniform_u32((unsigned)bx1) - kx0 + 1u, ny = uniform_u32((unsigned)by1) - ky0 + 1u;
        const unsigned nz_all = kz_hi - kz_lo + 1u;
        const unsigned long long nxy = (unsigned long long)nx * ny;
        if (nxy > kMaxKeys) {   // excluded by the library layout; reported, never scored wrong
            if (tid == 0) atomicOr(&a.ctl->error, 4);
            continue;
        }
        // layers per slab: the whole segment unless its box exceeds kMaxKeys keys
        const unsigned layers = (nxy * nz_all <= kMaxKeys) ? nz_all : (unsigned)(kMaxKeys / nxy);
        const bool multi = layers < nz_all;
        // the shift in lattice units
        const double sqx = sh.x * g.invD, sqy = sh.y * g.invD, sqz = sh.z * g.invD;
        // Exact membership of one point and its key in the current slab's
        // box.  An index equal to N (a point within rounding of the voxel's
        // upper face) is carried into the next index, giving the indices of
        // the sphere whose key the reference's sum kx + N ky + N^2 kz names;
        // both then accumulate into that one entry, as in the reference.
        // (all from s_ex, the exact path's constants in LDS)
        auto exact_local = [&](double x, double y, double z, unsigned &key) -> bool {
            const ExactParams &e = s_ex;
            unsigned gx, gy, gz;
            if (!exact_member(x, y, z, e, gx, gy, gz)) return false;
            const bool alias = max(max(gx, gy), gz) >= e.N;
            if (gx >= e.N) {
                gx -= e.N;
                ++gy;
            }
            if (gy >= e.N) {
                gy -= e.N;
                ++gz;
            }
            const unsigned lx = gx - e.kx0, ly = gy - e.ky0, lz = gz - e.kz0;
            // every in-sphere point lies in the segment's box by construction;
            // only a carried index can leave it (or, across slabs, move to
            // another slab): reported, never scored wrong
            if (lx >= e.nx || ly >= e.ny || gz - e.kz_lo >= e.nz_all || (alias && e.multi))
                atomicOr(&a.ctl->error, 4);
            if (!(lx < e.nx && ly < e.ny && lz < e.nz)) return false;
            key = (lz * e.ny + ly) * e.nx + lx;
            return true;
        };
        // (read after the block-range barrier below; rewritten only after
        // the next segment's ticket barriers, or, for a slab's layers, after
        // the previous slab's last flush)
        if (tid == 0) {
            s_ex.H = g.H;
            s_ex.gso = g.gso;
            s_ex.D = g.D;
            s_ex.invD = g.invD;
            s_ex.r2 = g.r2;
            s_ex.sx = sh.x;
            s_ex.sy = sh.y;
            s_ex.sz = sh.z;
            s_ex.N = N;
            s_ex.kx0 = kx0;
            s_ex.ky0 = ky0;
            s_ex.nx = nx;
            s_ex.ny = ny;
            s_ex.kz_lo = kz_lo;
            s_ex.nz_all = nz_all;
            s_ex.multi = multi ? 1u : 0u;
        }
        for (unsigned k0 = kz_lo;;) {
            const unsigned k1 = min(kz_hi, k0 + layers - 1u);
            box = SlabBox{kx0, ky0, k0, nx, ny, k1 - k0 + 1u, (unsigned)nxy};
            if (tid == 0) {
                s_ex.kz0 = k0;
                s_ex.nz = box.nz;
            }
            // lattice-unit offsets that make the rounded coordinates local indices
            const double ox = (MODE == ST_F64 ? sqx - g.qc : sqx) - (double)kx0;
            const double oy = (MODE == ST_F64 ? sqy - g.qc : sqy) - (double)ky0;
            const double oz = (MODE == ST_F64 ? (sqz + g.H * g.invD) - g.qc : sqz) - (double)k0;
            // a point of layer k in [k0, k1] has zs - gso within half a diameter
            // of k D: the slab's points lie strictly inside (zlo, zhi)
            const double zlo = fmax(-g.H, g.gso + ((double)k0 - 1.0) * g.D);
            const double zhi = fmin(g.H, g.gso + ((double)k1 + 1.0) * g.D);
            // block range: blocks whose z span can reach inside (zlo, zhi)
            __builtin_amdgcn_s_setprio(2);
            // (two waves search at once: the ends of the range)
            if (tid < 128) {
                const double *zf = a.zfirst + td.block0;
                const int nb = td.nblocks;
                auto zs_of = [&](double z) { return (z + g.H) + sh.z; };
                if (tid < 64) {
                    const int r_lo = wave_lower_bound(nb, [&](int b) {
                        const double zhb = (b + 1 < nb) ? zf[b + 1] : td.zmax;
                        return zs_of(zhb) > zlo;
                    });
                    if (tid == 0) s_range[0] = r_lo;
                } else {
                    const int r_hi = wave_lower_bound(nb, [&](int b) { return !(zs_of(zf[b]) < zhi); });
                    if (tid == 64) s_range[1] = r_hi;
                }
            }
            __syncthreads();
            __builtin_amdgcn_s_setprio(0);
            const int b_lo = uniform_i32(s_range[0]), b_hi = uniform_i32(s_range[1]);
            // the next slab's set-up writes s_range after every wave's read
            if (multi) __syncthreads();
            if (b_lo < b_hi) {
                const int64_t p_end = td.offset + min((int64_t)b_hi * kBlk, (int64_t)td.n);
                // point indices relative to the track's first point (td.n <=
                // 2^22: 32-bit lane offsets from uniform base addresses)
                const int r_end = (int)(p_end - td.offset);
                binned += (long long)(b_hi - b_lo) * kBlk;
                n_scored += (unsigned long long)(r_end - b_lo * kBlk);
                // Software pipeline: the points of the next iteration are loaded
                // before the current ones are scored.  Lanes past the end of the
                // range load a valid point and are masked off.
                using PointT = typename std::conditional<MODE == ST_F64, double4, float4>::type;
                PointT buf0[kPPT];
                auto fetch = [&](int b, PointT *dst) {
#pragma unroll
                    for (int j = 0; j < kPPT; ++j) {
                        // past the range's last point: that point again
                        const unsigned r = (b + j < b_hi) ? min((unsigned)((b + j) * kBlk + tid), (unsigned)r_end - 1u)
                                                          : (unsigned)(b_lo * kBlk);
                        if constexpr (MODE == ST_F64) {
                            const unsigned o8 = r * 8u;
                            dst[j] = make_double4(*(const double *)((const char *)(a.x64 + td.offset) + o8),
                                                  *(const double *)((const char *)(a.y64 + td.offset) + o8),
                                                  *(const double *)((const char *)(a.z64 + td.offset) + o8),
                                                  *(const double *)((const char *)(a.e64 + td.offset) + o8));
                        } else {
                            dst[j] = *(const float4 *)((const char *)(a.p32 + td.offset) + r * 16u);
                        }
                    }
                };
                PointT buf1[kPPT];
                fetch(b_lo, buf0);
                // One iteration: rows b .. b + kPPT - 1, their points in buf0
                // (buffer 0) or buf1 (buffer 1); then, if it is a checking
                // iteration (a compile-time choice: the check and its flush are
                // inlined once per loop), the occupancy check.
                auto iteration = [&](const int b, auto buffer, auto checking) __attribute__((always_inline)) {
                    PointT(&cur)[kPPT] = decltype(buffer)::value ? buf1 : buf0;
                    const int64_t c = td.offset + (int64_t)b * kBlk;
                    // the iteration's last point (thread kBT - 1, last row):
                    // its local layer, or kLayerUnsure if its rounding was
                    // within the error bound
                    int last_layer = 0;
                    int fresh[kPPT];
#pragma unroll
                    for (int j = 0; j < kPPT; ++j) fresh[j] = -1;
#pragma unroll
                    for (int j = 0; j < kPPT; ++j) {
                        if (b + j >= b_hi) break;
                        const int ri = (b + j) * kBlk + tid;   // relative to the track's first point
                        const bool valid = ri < r_end;
                        bool in = false;
                        unsigned key = 0;
                        long long v0 = 0, v1 = 0;
                        // lattice units relative to the slab's box corner
                        double qx, qy, qz, fm, mrgq;
                        if constexpr (MODE == ST_F64) {
                            // straight from the FP64 data; decisions within the
                            // rounding bound are redone exactly
                            qx = fma(cur[j].x, g.invD, ox);
                            qy = fma(cur[j].y, g.invD, oy);
                            qz = fma(cur[j].z, g.invD, oz);
                            fm = g.fm;
                            mrgq = g.mrgq;
                        } else {
                            // block-uniform terms (scalar loads); per point
                            // |q - reference quotient| <= dl / D
                            const BlockF32 bk = load_block(a.blocks, td.block0 + b + j);
                            qx = fma((double)cur[j].x, g.invD, bk.oqx + ox);
                            qy = fma((double)cur[j].y, g.invD, bk.oqy + oy);
                            qz = fma((double)cur[j].z, g.invD, bk.oqz + oz);
                            fm = bk.fm;
                            mrgq = bk.mrgq;
                        }
                        const double tx = qx + kRoundMagic, ty = qy + kRoundMagic, tz = qz + kRoundMagic;
                        const double ux = (tx - kRoundMagic) - qx, uy = (ty - kRoundMagic) - qy,
                                     uz = (tz - kRoundMagic) - qz;
                        const double distq = fma(ux, ux, fma(uy, uy, uz * uz));
                        // low words of t: the local indices (|q| < 2^30, checked per track)
                        const unsigned lx = (unsigned)__double2loint(tx), ly = (unsigned)__double2loint(ty),
                                       lz = (unsigned)__double2loint(tz);
                        // inside the slab's box, which lies inside the voxel
                        const bool inbox = (lx < box.nx) & (ly < box.ny) & (lz < box.nz);
                        const bool near =
                            (fmax(fmax(fabs(ux), fabs(uy)), fabs(uz)) >= fm) | (fabs(distq - g.r2q) <= mrgq);
                        in = valid & inbox & !near & (distq < g.r2q);
                        key = (lz * box.ny + ly) * box.nx + lx;
                        if (j == kPPT - 1) last_layer = near ? kLayerUnsure : (int)lz;
                        if (valid & near) {
                            // any decision within the error bound is redone exactly
                            // from FP64, unless clearly outside (an index beyond the
                            // box by more than one)
                            const bool cand =
                                (lx + 1u <= box.nx + 1u) & (ly + 1u <= box.ny + 1u) & (lz + 1u <= box.nz + 1u);
                            // (statistic of the FP32 path: decisions sent to FP64)
                            if constexpr (MODE == ST_F32) n_flagged += cand;
                            if constexpr (MODE == ST_F64)
                                in = cand && exact_local(cur[j].x, cur[j].y, cur[j].z, key);
                            else
                                in = cand && exact_local(a.x64[td.offset + ri], a.y64[td.offset + ri],
                                                         a.z64[td.offset + ri], key);
                        }
                        if constexpr (MODE == ST_F64) {
                            if (in) Acc<ST_F64>::split(cur[j].w / g.denom, v0, v1);
                        } else {
                            v0 = rint_to_i64((double)cur[j].w * g.yscale);
                        }
                        n_in += in;
                        if (in) {
                            // linear probing; the loop is kept rolled: nearly every
                            // probe ends at its first slot, and an unrolled loop pays
                            // exec-mask merges for every possible exit
                            unsigned int h = slot_hash<SLOTS>(key);
                            unsigned prev = atomicCAS(&tkey[h], kEmpty32, key);
#pragma unroll 1
                            for (int probe = 1; probe < SLOTS && prev != kEmpty32 && prev != key; ++probe) {
                                h = next_slot<SLOTS>(h);
                                prev = atomicCAS(&tkey[h], kEmpty32, key);
                            }
                            if (prev == kEmpty32 || prev == key) {
                                if (prev == kEmpty32) fresh[j] = (int)h;
                                atomicAdd(&tacc0[h], (unsigned long long)v0);
                                if (kWords == 2) atomicAdd(&tacc1[h], (unsigned long long)v1);
                                bal_in += Acc<MODE>::balance((unsigned long long)v0, (unsigned long long)v1);
                            } else {
                                atomicOr(&a.ctl->error, 1);
                            }
                        }
                    }
                    append(fresh, std::integral_constant<int, kPPT>(), true);
                    const int64_t next = c + (int64_t)kPPT * kBlk;
                    if (!decltype(checking)::value || next >= p_end) return;
                    // Every wave must take the same decision.  Each publishes its own
                    // count into this check's half of s_wcnt and sums all of them
                    // after one barrier: a wave that runs on and appends more writes
                    // only its count for the next check, into the other half, so the
                    // counts read here cannot change under a slower wave (which would
                    // otherwise flush alone).
                    if (lane == 0) s_wcnt[par][tid >> 6] = my_cnt;
                    if (tid == kBT - 1) s_layer[par] = last_layer;
                    __syncthreads();
                    int occ = 0;
#pragma unroll
                    for (int w = 0; w < kBT / 64; ++w) {
                        wcnt[w] = s_wcnt[par][w];
                        occ += wcnt[w];
                    }
                    const int lb = uniform_i32(s_layer[par]);
                    par ^= 1;
                    if (occ >= kFlushAt) {
                        // The exact layer index is monotone in z, so every later
                        // point lies in a layer >= that of this iteration's last
                        // point (the block is full: the range goes on); a key's
                        // layer is never below its point's: entries in the slab's
                        // layers below lb are complete.
                        int lb_exact = lb;
                        if (lb == kLayerUnsure) {
                            // rare: the exact layer, as the reference rounds it
                            const double zz = load_uniform(a.z64 + (next - 1));
                            const ExactParams &e = s_ex;
                            lb_exact = uniform_i32((int)nearest_index(((zz + e.H) + e.sz) - e.gso, e.D, e.invD)) -
                                       (int)box.kz0;
                        }
                        unsigned bound = 0u;
                        if (lb_exact >= (int)box.nz) bound = kEmpty32;
                        else if (lb_exact > 0) bound = (unsigned)lb_exact * box.nxy;
                        flush(bound, true);
                    }
                };
                using Buf0 = std::integral_constant<int, 0>;
                using Buf1 = std::integral_constant<int, 1>;
                if constexpr (kCheck == 2) {
                    // two iterations per trip, their points loaded into
                    // alternating buffers (no register copies), the check after
                    // the second (every second iteration from the range's start)
                    for (int b = b_lo; b < b_hi; b += 2 * kPPT) {
                        if (b + kPPT < b_hi) fetch(b + kPPT, buf1);
                        iteration(b, Buf0(), std::false_type());
                        if (b + kPPT >= b_hi) break;
                        if (b + 2 * kPPT < b_hi) fetch(b + 2 * kPPT, buf0);
                        iteration(b + kPPT, Buf1(), std::true_type());
                    }
                } else {
                    static_assert(kCheck == 1, "occupancy checks every one or two iterations");
                    // one iteration per trip: the current points copied to
                    // buffer 1, the next ones loaded into buffer 0
                    for (int b = b_lo; b < b_hi; b += kPPT) {
#pragma unroll
                        for (int j = 0; j < kPPT; ++j) buf1[j] = buf0[j];
                        if (b + kPPT < b_hi) fetch(b + kPPT, buf0);
                        iteration(b, Buf1(), std::true_type());
                    }
                }
                flush(kEmpty32, false);
            }
            if (k1 >= kz_hi) break;
            k0 = k1 + 1u;
        }
        if (a.hist_lds_words && binned > 0x40000000ll) {
            // keep the 32-bit LDS counters far from overflow
#pragma unroll
            for (int h = 0; h < kMaxHists; ++h) {
                if (h >= a.n_hists) continue;
                const int nb = a.hist[h].guess.nbins;
                for (int b = tid; b < nb; b += kBT) {
                    unsigned long long v = 0;
                    for (int cpy = 0; cpy < a.hist_copies; ++cpy) {
                        unsigned int *w = &lhist[cpy * a.hist_lds_words + a.hist[h].lds_offset + b];
                        v += *w;
                        *w = 0;
                    }
                    if (v) atomicAdd(&a.hist[h].bins[b], v);
                }
            }
            __syncthreads();
            binned = 0;
        }
    }
    {
        // per-lane counters summed over the wave; n_created / n_retired are
        // already per wave, n_scored per workgroup
        unsigned long long c[4] = {n_flagged, n_in, bal_in, bal_out};
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
#pragma unroll
            for (int k = 0; k < 4; ++k) c[k] += __shfl_xor(c[k], d, 64);
        }
        if (lane == 0) {
            if (tid == 0) atomicAdd(&a.ctl->stats[0], n_scored);
#pragma unroll
            for (int k = 0; k < 4; ++k) atomicAdd(&a.ctl->stats[k + 1], c[k]);
            atomicAdd(&a.ctl->stats[5], n_created);
            atomicAdd(&a.ctl->stats[6], n_retired);
        }
    }
    __syncthreads();
    if (a.hist_lds_words == 0) return;
#pragma unroll
    for (int h = 0; h < kMaxHists; ++h) {
        if (h >= a.n_hists) continue;
        const int nb = a.hist[h].guess.nbins;
        for (int b = tid; b < nb; b += kBT) {
            unsigned long long v = 0;
            for (int cpy = 0; cpy < a.hist_copies; ++cpy) v += lhist[cpy * a.hist_lds_words + a.hist[h].lds_offset + b];
            if (v) atomicAdd(&a.hist[h].bins[b], v);
        }
    }
}

__global__ void batch_reset_kernel(BatchCtl *ctl)
{
    if (threadIdx.x == 0) {
        ctl->ticket = 0;
        ctl->rec_count = 0;
        ctl->error = 0;
        ctl->pad = 0;
    }
    if (threadIdx.x < kStats) ctl->stats[threadIdx.x] = 0;
}

}  // namespace
}  // namespace st
