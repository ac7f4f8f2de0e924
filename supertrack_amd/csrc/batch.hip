// batch.hip -- the batched form of the VoxelConstrainedSphere path: a resident
// track library and one persistent fused kernel that scores every
// (track, oversample) segment: shift -> voxel -> sphere -> key -> per-sphere
// sum -> lineal energy -> bin.
//
// Library layout (HBM), built once by st_library_create:
//   * each track's points sorted by z (a rigid shift keeps that order, so for
//     any oversample the points inside the voxel in z form one contiguous run
//     and the sphere layer index iz never decreases along the run);
//   * FP64 SoA x, y, z, edep in that order (the data in ST_F64 mode, the exact
//     side copy in ST_F32 mode);
//   * ST_F32 mode: FP32 (dx, dy, dz, edep) per point, coordinates relative to a
//     per-256-point block origin (FP64) and a per-block bound on the FP32
//     rounding error;
//   * per block: z of its first point (the search index); per track: offsets,
//     x/y bounds, geometry, eligibility.
//
// The fused kernel itself is in batch_kernel.hpp; this file holds the
// library construction, launch shapes, autotuner, exact path and the C ABI.
// Tracks whose densest z-window two diameters wide could touch more spheres
// than a flush may keep (768), with non-finite data, beyond the fixed-point
// headroom or too wide for 32-bit sphere keys are scored by the exact
// per-segment path.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <mutex>
#include <new>
#include <numeric>
#include <type_traits>
#include <vector>

#include "st_common.hpp"
#include "st_internal.hpp"
#include "st_device.hpp"
#include "layout.hpp"

#include "batch_kernel.hpp"

namespace st {
namespace {

template <int MODE, int SLOTS, int PPT>
const void *kernel_ptr()
{
    return (const void *)batch_kernel<MODE, SLOTS, PPT>;
}

template <int MODE>
const void *pick_kernel_mode(int slots, int ppt)
{
    // 1280 and 1472 slots hold the largest survivor set plus the points
    // between two occupancy checks at one point per thread only
    if (slots == 1280) return ppt == 2 ? nullptr : kernel_ptr<MODE, 1280, 1>();
    if (slots == 1472) return ppt == 2 ? nullptr : kernel_ptr<MODE, 1472, 1>();
    if (slots == 1536) return ppt == 2 ? kernel_ptr<MODE, 1536, 2>() : kernel_ptr<MODE, 1536, 1>();
    if (slots == 2048) return ppt == 2 ? kernel_ptr<MODE, 2048, 2>() : kernel_ptr<MODE, 2048, 1>();
    return ppt == 2 ? kernel_ptr<MODE, 4096, 2>() : kernel_ptr<MODE, 4096, 1>();
}

const void *pick_kernel(int mode, int slots, int ppt)
{
    return mode == ST_F64 ? pick_kernel_mode<ST_F64>(slots, ppt) : pick_kernel_mode<ST_F32>(slots, ppt);
}

size_t lds_bytes(int mode, int slots, int ppt, int hist_words)
{
    const int words = (mode == ST_F64) ? 2 : 1;
    // per slot: the sums and the 32-bit key; the waves' parts of the slot
    // list; the privatised bins
    return (size_t)slots * (8 * words + 4) + (size_t)(kBT / 64) * list_region(slots, ppt) * 2 +
           (size_t)hist_words * 4;
}

template <typename T>
int dev_alloc(T **dst, size_t n, int64_t *bytes)
{
    *dst = nullptr;
    if (n == 0) return ST_OK;
    hipError_t e = dev_malloc((void **)dst, n * sizeof(T));
    if (e != hipSuccess) {
        *dst = nullptr;
        return set_error(ST_ENOMEM, "st_library_begin: hipMalloc(%zu B) failed: %s", n * sizeof(T),
                         hipGetErrorString(e));
    }
    *bytes += (int64_t)(n * sizeof(T));
    return ST_OK;
}

template <typename T>
int dev_put(T *dst, const T *src, size_t n)
{
    if (n == 0) return ST_OK;
    ST_HIP_TRY(hipMemcpy(dst, src, n * sizeof(T), hipMemcpyHostToDevice));
    return ST_OK;
}

template <typename T>
int dev_upload(T **dst, const T *src, size_t n, int64_t *bytes)
{
    *dst = nullptr;
    if (n == 0) return ST_OK;
    hipError_t e = dev_malloc((void **)dst, n * sizeof(T));
    if (e != hipSuccess) {
        *dst = nullptr;
        return set_error(ST_ENOMEM, "st_library_create: hipMalloc(%zu B) failed: %s", n * sizeof(T),
                         hipGetErrorString(e));
    }
    *bytes += (int64_t)(n * sizeof(T));
    ST_HIP_TRY(hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice));
    return ST_OK;
}


// Launch shape of the fused kernel.
struct Launch {
    int slots = 1536, blocks_per_cu = 1, ppt = 1, grid = 0;
    bool lds_hist = true;     // histograms privatised in LDS (else device atomics)
    int hist_copies = 1;      // LDS copies of the privatised histograms
    size_t lds = 0;
};

}  // namespace
}  // namespace st

using namespace st;

struct st_library {
    int64_t n_tracks = 0, n_points = 0, n_blocks = 0;
    int precision = ST_F64;
    std::vector<st_geometry> geoms;
    std::vector<TrackDesc> tracks;
    std::vector<int32_t> fast_order;   // fast tracks, largest first
    std::vector<int32_t> exact;        // tracks for the exact path
    int64_t device_bytes = 0;
    double *x64 = nullptr, *y64 = nullptr, *z64 = nullptr, *e64 = nullptr;
    float4 *p32 = nullptr;
    double *zfirst = nullptr;
    BlockF32 *blocks = nullptr;
    TrackDesc *tracks_dev = nullptr;
    GeomDev *geoms_dev = nullptr;
    int32_t *order_dev = nullptr;      // fast_order on the device
    // streamed construction: begin sizes the device arrays, appends fill them
    // in track order, finish publishes the track table
    int64_t tracks_done = 0, points_done = 0, blocks_cap = 0;
    bool finished = false;
};

struct st_batch {
    st_library *lib = nullptr;
    st_batch_opts opts{};
    std::vector<st_hist *> hists;
    hipStream_t stream = nullptr;
    Launch ln;
    int hist_words = 0;
    BatchCtl *ctl = nullptr;   // per-pass control block (device)
    RecArgs rec{};                // recording buffers (cap 0: off)
    RecArgs *rec_dev = nullptr;   // their descriptor in device memory
    bool graph_ready = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // timing of the last run: fused kernel (graph) and exact path
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    // exact path
    std::vector<st_vcs *> methods;    // one per geometry
    st_track scratch{};
    st_pairs pairs{};
};

namespace {

BatchArgs make_args(const st_library *L, const st_batch_opts &o, st_hist *const *hists, int n_hists,
                    const Launch &ln, BatchCtl *ctl)
{
    BatchArgs a{};
    a.tracks = L->tracks_dev;
    a.x64 = L->x64;
    a.y64 = L->y64;
    a.z64 = L->z64;
    a.e64 = L->e64;
    a.p32 = L->p32;
    a.zfirst = L->zfirst;
    a.blocks = L->blocks;
    a.geoms = L->geoms_dev;
    a.order = L->order_dev;
    a.n_seg = (int64_t)L->fast_order.size() * o.n_oversamples;
    a.n_os = std::max(o.n_oversamples, 1);
    a.shift_xy = o.shift_xy;
    a.shift_z = o.shift_z;
    a.track_index_base = o.track_index_base;
    a.seed = o.seed;
    a.ctl = ctl;
    a.n_hists = n_hists;
    int off = 0;
    for (int h = 0; h < n_hists; ++h) {
        hists[h]->bins_zero = false;   // the pass counts into them
        a.hist[h].edges = hists[h]->edges_dev;
        a.hist[h].bins = (unsigned long long *)hists[h]->bins_dev;
        a.hist[h].guess = hists[h]->guess;
        a.hist[h].lds_offset = off;
        off += hists[h]->nbins;
    }
    a.hist_lds_words = ln.lds_hist ? off : 0;
    a.hist_copies = ln.lds_hist ? ln.hist_copies : 1;
    return a;
}

// Queues one pass: the control-block reset, then the fused kernel.
int launch_pass(int precision, const BatchArgs &args, const Launch &ln, hipStream_t stream)
{
    batch_reset_kernel<<<1, 64, 0, stream>>>(args.ctl);
    ST_HIP_TRY(hipGetLastError());
    if (args.n_seg == 0) return ST_OK;
    const void *k = pick_kernel(precision, ln.slots, ln.ppt);
    BatchArgs a = args;
    void *kargs[] = {&a};
    ST_HIP_TRY(hipLaunchKernel(k, dim3(ln.grid), dim3(kBT), kargs, ln.lds, stream));
    return ST_OK;
}

int launch_fast(st_batch *b)
{
    BatchArgs a = make_args(b->lib, b->opts, b->hists.data(), (int)b->hists.size(), b->ln, b->ctl);
    a.rec = b->rec.cap > 0 ? b->rec_dev : nullptr;
    return launch_pass(b->lib->precision, a, b->ln, b->stream);
}

// Launch shape for (table slots, workgroups per CU, points per thread, LDS
// copies of the histograms), capped by what the kernel's registers and LDS
// allow.
int configure_launch(int precision, int hist_words, int slots, int bpc, int ppt, int copies, Launch *out)
{
    Launch ln;
    ln.ppt = ppt == 2 ? 2 : 1;
    if (copies != 1 && copies != 2 && copies != 4)
        return set_error(ST_EINVAL, "batched run: hist_copies must be 1, 2 or 4");
    // privatise the histograms in LDS when they fit next to the table at two
    // workgroups per CU; otherwise count straight into the device bins (one copy)
    ln.lds_hist = lds_bytes(precision, slots, ln.ppt, hist_words * copies) <= 80 * 1024;
    if (!ln.lds_hist && copies > 1)
        return set_error(ST_ERANGE, "batched run: %d LDS copies of %d histogram bins do not fit", copies, hist_words);
    ln.hist_copies = copies;
    const size_t lds = lds_bytes(precision, slots, ln.ppt, ln.lds_hist ? hist_words * copies : 0);
    if (lds > 160 * 1024)
        return set_error(ST_ERANGE, "batched kernel needs %zu B of LDS (limit 160 KiB)", lds);
    const void *k = pick_kernel(precision, slots, ln.ppt);
    if (!k) return set_error(ST_EINVAL, "batched kernel: %d table slots take one point per thread only", slots);
    ST_HIP_TRY(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int per_cu = 0;
    ST_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, kBT, lds));
    if (per_cu < 1) return set_error(ST_EHIP, "batched kernel cannot be resident with %zu B LDS", lds);
    ln.slots = slots;
    ln.blocks_per_cu = std::max(1, std::min(bpc, per_cu));
    ln.grid = ln.blocks_per_cu * device_cu_count();
    ln.lds = lds;
    *out = ln;
    return ST_OK;
}

int configure(st_batch *b, int slots, int bpc, int ppt, int copies)
{
    return configure_launch(b->lib->precision, b->hist_words, slots, bpc, ppt, copies, &b->ln);
}

// The launch shape used unless the options name one or ask for tuning: every
// table size holds the largest survivor set (kMaxWindow) or a quarter table,
// plus the points of the iterations between occupancy checks, so every
// fast-path track fits.  Defaults: the fastest shapes measured on MI355X for
// 2 nm and 10 nm spheres.
int default_launch(int precision, int hist_words, const st_batch_opts &o, Launch *out)
{
    const bool f32 = precision == ST_F32;
    const int slots = o.table_slots ? o.table_slots : 1536;
    if (slots != 1280 && slots != 1472 && slots != 1536 && slots != 2048 && slots != 4096)
        return set_error(ST_EINVAL, "batched run: table_slots must be 1280, 1472, 1536, 2048 or 4096");
    const int bpc = o.blocks_per_cu ? o.blocks_per_cu : (f32 ? 5 : 4);
    const int ppt = o.points_per_thread ? o.points_per_thread : (f32 ? 1 : 2);
    if (ppt != 1 && ppt != 2) return set_error(ST_EINVAL, "batched run: points_per_thread must be 1 or 2");
    const int copies = o.hist_copies ? o.hist_copies : 1;
    // FP32, shape not given: 1280-slot tables at seven workgroups per CU
    // where they fit, else 1472 slots at six (cfg 4 on one box: 273.8 ms,
    // 278.0 ms, against 288 ms at 1536 x 5 x 1)
    if (f32 && !o.table_slots && !o.blocks_per_cu && !o.points_per_thread) {
        const int shapes[2][2] = {{1280, 7}, {1472, 6}};
        for (const auto &sh : shapes) {
            Launch occ;
            if (configure_launch(precision, hist_words, sh[0], sh[1], 1, copies, &occ) == ST_OK &&
                occ.blocks_per_cu == sh[1]) {
                *out = occ;
                return ST_OK;
            }
            clear_error();
        }
    }
    int rc = configure_launch(precision, hist_words, slots, bpc, ppt, copies, out);
    if (rc) return rc;
    // FP32 with fewer than five workgroups per CU (large privatised
    // histograms): a 1280-slot table if it fits more workgroups; else two
    // points per thread, at four workgroups per CU if the one-point table
    // fits four (the full cfg 5 run on one box: 1536 x 4 x 2 in 1290 ms,
    // 1536 x 4 x 1 in 1312 ms; earlier 1536 x 3 x 2, with a kernel that fit
    // three, 1564 ms against 1406 ms at 1536 x 4 x 1)
    if (f32 && !o.points_per_thread && out->blocks_per_cu < 5) {
        const int fit = out->blocks_per_cu;
        if (!o.table_slots) {
            Launch alt;
            if (configure_launch(precision, hist_words, 1280, bpc, 1, copies, &alt) == ST_OK &&
                alt.blocks_per_cu > fit) {
                *out = alt;
                return ST_OK;
            }
            clear_error();
        }
        if (fit < 4) return configure_launch(precision, hist_words, slots, bpc, 2, copies, out);
        Launch two;
        if (configure_launch(precision, hist_words, slots, bpc, 2, copies, &two) == ST_OK &&
            two.blocks_per_cu >= fit)
            *out = two;
        clear_error();
        return ST_OK;
    }
    return ST_OK;
}

// Exact per-segment path for tracks the fused kernel does not take.
int run_exact(st_batch *b)
{
    st_library *L = b->lib;
    for (int32_t t : L->exact) {
        const TrackDesc &td = L->tracks[t];
        const int64_t n = td.n;
        if (n > b->scratch.capacity) {
            st_track_free(&b->scratch);
            int rc = st_track_alloc(&b->scratch, n);
            if (rc) return rc;
        }
        if (n > b->pairs.capacity) {
            st_pairs_free(&b->pairs);
            int rc = st_pairs_alloc(&b->pairs, n);
            if (rc) return rc;
        }
        const size_t bytes = (size_t)n * sizeof(double);
        ST_HIP_TRY(hipMemcpyAsync(b->scratch.x, L->x64 + td.offset, bytes, hipMemcpyDeviceToDevice, b->stream));
        ST_HIP_TRY(hipMemcpyAsync(b->scratch.y, L->y64 + td.offset, bytes, hipMemcpyDeviceToDevice, b->stream));
        ST_HIP_TRY(hipMemcpyAsync(b->scratch.z, L->z64 + td.offset, bytes, hipMemcpyDeviceToDevice, b->stream));
        ST_HIP_TRY(hipMemcpyAsync(b->scratch.edep, L->e64 + td.offset, bytes, hipMemcpyDeviceToDevice, b->stream));
        b->scratch.n = n;
        st_pairs view = b->pairs;
        view.capacity = std::max<int64_t>(n, 1);
        st_vcs *m = b->methods[td.geom];
        int rc = st_vcs_begin_track(m, &b->scratch, b->opts.seed, b->opts.track_index_base + (uint32_t)t,
                                    b->opts.n_oversamples);
        if (rc) return rc;
        for (int os = 0; os < b->opts.n_oversamples; ++os) {
            rc = st_vcs_process_oversample(m, &b->scratch, &view);
            if (rc) return rc;
            for (st_hist *h : b->hists) {
                rc = hist_add_pairs(h, &view, b->stream);
                if (rc) return rc;
            }
        }
        m->in_track = false;
    }
    return ST_OK;
}

// Times one full pass of every launch candidate and keeps the fastest.  The
// histograms' counters are saved before and restored after, so tuning adds
// no counts; the library's streams are non-blocking, so every stream that may
// still write those counters is drained first.
int autotune(st_batch *b)
{
    struct Events {
        hipEvent_t e[2] = {nullptr, nullptr};
        ~Events()
        {
            for (hipEvent_t x : e)
                if (x) (void)hipEventDestroy(x);
        }
    } ev;
    ST_HIP_TRY(hipEventCreate(&ev.e[0]));
    ST_HIP_TRY(hipEventCreate(&ev.e[1]));
    ST_HIP_TRY(hipStreamSynchronize(b->stream));
    std::vector<std::vector<uint64_t>> saved;
    for (st_hist *h : b->hists) {
        ST_HIP_TRY(hipStreamSynchronize(h->stream));
        saved.emplace_back(h->nbins);
        ST_HIP_TRY(hipMemcpyAsync(saved.back().data(), h->bins_dev, h->nbins * sizeof(uint64_t),
                                  hipMemcpyDeviceToHost, b->stream));
    }
    ST_HIP_TRY(hipStreamSynchronize(b->stream));
    const int cand_slots[] = {1280, 1472, 1536, 2048, 4096};
    const int cand_bpc[] = {2, 3, 4, 5, 6, 7, 8};
    const int cand_copies[] = {1, 2, 4};
    float best = 1e30f;
    int best_s = b->ln.slots, best_b = b->ln.blocks_per_cu, best_p = b->ln.ppt, best_c = b->ln.hist_copies;
    for (int s : cand_slots)
        for (int pp = 1; pp <= 2; ++pp)
            for (int bp : cand_bpc)
                for (int cp : cand_copies) {
                if (configure(b, s, bp, pp, cp) != ST_OK) continue;
                if (b->ln.blocks_per_cu != bp) continue;
                ST_HIP_TRY(hipEventRecord(ev.e[0], b->stream));
                int rc = launch_fast(b);
                if (rc) return rc;
                ST_HIP_TRY(hipEventRecord(ev.e[1], b->stream));
                ST_HIP_TRY(hipEventSynchronize(ev.e[1]));
                float ms = 0;
                ST_HIP_TRY(hipEventElapsedTime(&ms, ev.e[0], ev.e[1]));
                if (ms < best) {
                    best = ms;
                    best_s = s;
                    best_b = bp;
                    best_p = pp;
                    best_c = cp;
                }
            }
    for (size_t i = 0; i < b->hists.size(); ++i)
        ST_HIP_TRY(hipMemcpyAsync(b->hists[i]->bins_dev, saved[i].data(), saved[i].size() * sizeof(uint64_t),
                                  hipMemcpyHostToDevice, b->stream));
    ST_HIP_TRY(hipStreamSynchronize(b->stream));
    clear_error();   // candidates that did not fit set an error string
    return configure(b, best_s, best_b, best_p, best_c);
}

// ---------------------------------------------------------------------------
// allocation-free pass (st_batch_run_static)
// ---------------------------------------------------------------------------
// Its control block lives in this file-scope device array: one copy per
// device, kept until the process ends.  The host-side name is a stand-in, not
// a device address: each device's address comes from hipGetSymbolAddress,
// taken once per device with that device current, and only that address is
// handed to kernels and copies.
constexpr size_t kStaticScratchBytes = 256;
static_assert(sizeof(BatchCtl) <= kStaticScratchBytes, "the control block fits the static scratch");
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];
constexpr int kMaxDevices = 64;

int static_scratch(BatchCtl **out)
{
    static std::mutex mu;
    static void *addr[kMaxDevices] = {};
    int dev = 0;
    ST_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        return set_error(ST_ERANGE, "st_batch_run_static: device %d beyond the %d supported", dev, kMaxDevices);
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(ST_EHIP, "st_batch_run_static: hipGetSymbolAddress(scratch) failed on device %d: %s",
                             dev, hipGetErrorString(e));
        addr[dev] = p;
    }
    *out = reinterpret_cast<BatchCtl *>(addr[dev]);
    return ST_OK;
}

// Waits for the work queued on `stream` and copies the control block back.
int read_ctl(const BatchCtl *ctl, hipStream_t stream, BatchCtl *out)
{
    ST_HIP_TRY(hipMemcpyAsync(out, ctl, sizeof *out, hipMemcpyDeviceToHost, stream));
    ST_HIP_TRY(hipStreamSynchronize(stream));
    return ST_OK;
}

// The kernel's error flag and the conservation totals of a finished pass.
int check_ctl(const BatchCtl &c)
{
    if (c.error)
        return set_error(ST_ESTATE,
                         "batched kernel reported an error (code %d: 1 full sphere table, 2 entry list overrun, "
                         "4 key outside its segment's box)",
                         c.error);
    if (c.stats[3] != c.stats[4] || c.stats[5] != c.stats[6])
        return set_error(ST_ESTATE,
                         "batched kernel lost sphere-table entries: energy words in %016llx / out %016llx, "
                         "entries created %llu / retired %llu",
                         c.stats[3], c.stats[4], c.stats[5], c.stats[6]);
    return ST_OK;
}

}  // namespace

extern "C" {

int st_library_begin(int64_t n_tracks, int64_t n_points, const st_geometry *geoms, int32_t n_geoms,
                     int32_t precision, st_library **out)
{
    if (!out) return set_error(ST_EINVAL, "st_library_begin: null output");
    *out = nullptr;
    if (n_tracks < 0 || n_points < 0 || !geoms || n_geoms < 1 || n_geoms > kMaxGeoms)
        return set_error(ST_EINVAL, "st_library_begin: bad arguments (1 <= n_geoms <= %d)", kMaxGeoms);
    if (precision != ST_F64 && precision != ST_F32)
        return set_error(ST_EINVAL, "st_library_begin: precision must be ST_F64 or ST_F32");
    for (int i = 0; i < n_geoms; ++i)
        if (!(geoms[i].diameter > 0) || geoms[i].n_linear < 1)
            return set_error(ST_EINVAL, "st_library_begin: invalid geometry %d", i);
    st_library *L = new (std::nothrow) st_library;
    if (!L) return set_error(ST_ENOMEM, "st_library_begin: out of host memory");
    L->n_tracks = n_tracks;
    L->n_points = n_points;
    L->precision = precision;
    L->geoms.assign(geoms, geoms + n_geoms);
    L->tracks.resize(n_tracks);
    // ceil(n / kBlk) <= n / kBlk + 1 per track
    L->blocks_cap = n_points / kBlk + n_tracks + 1;
    int rc = ST_OK;
    int64_t bytes = 0;
    if (!rc) rc = dev_alloc(&L->x64, n_points, &bytes);
    if (!rc) rc = dev_alloc(&L->y64, n_points, &bytes);
    if (!rc) rc = dev_alloc(&L->z64, n_points, &bytes);
    if (!rc) rc = dev_alloc(&L->e64, n_points, &bytes);
    if (!rc) rc = dev_alloc(&L->zfirst, L->blocks_cap, &bytes);
    if (!rc && precision == ST_F32) {
        rc = dev_alloc(&L->p32, n_points, &bytes);
        if (!rc) rc = dev_alloc(&L->blocks, L->blocks_cap, &bytes);
    }
    L->device_bytes = bytes;
    if (rc) {
        st_library_destroy(L);
        return rc;
    }
    *out = L;
    return ST_OK;
}

int st_library_append(st_library *L, const double *x, const double *y, const double *z, const double *edep,
                      const int64_t *offsets, int64_t n_tracks, const int32_t *geom_of_track)
{
    if (!L || !offsets || n_tracks < 0) return set_error(ST_EINVAL, "st_library_append: bad arguments");
    if (L->finished) return set_error(ST_EINVAL, "st_library_append: library already finished");
    const int n_geoms = (int)L->geoms.size();
    const int64_t n_points = offsets[n_tracks] - offsets[0];
    if (n_points > 0 && (!x || !y || !z || !edep)) return set_error(ST_EINVAL, "st_library_append: null arrays");
    for (int64_t t = 0; t < n_tracks; ++t) {
        const int64_t n = offsets[t + 1] - offsets[t];
        if (n < 0 || n > 0x7FFFFFFF)
            return set_error(ST_EINVAL, "st_library_append: bad offsets at track %lld", (long long)t);
        if (geom_of_track && (geom_of_track[t] < 0 || geom_of_track[t] >= n_geoms))
            return set_error(ST_EINVAL, "st_library_append: geometry index out of range at track %lld",
                             (long long)t);
    }
    if (n_tracks > L->n_tracks - L->tracks_done || n_points > L->n_points - L->points_done)
        return set_error(ST_EINVAL,
                         "st_library_append: %lld tracks / %lld points exceed what st_library_begin sized "
                         "(%lld / %lld left)",
                         (long long)n_tracks, (long long)n_points, (long long)(L->n_tracks - L->tracks_done),
                         (long long)(L->n_points - L->points_done));
    const int precision = L->precision;
    const int64_t pbase = L->points_done;   // this chunk's first point in the library
    int64_t bbase = 0;                       // this chunk's first block in the library
    if (L->tracks_done > 0) {
        const TrackDesc &last = L->tracks[L->tracks_done - 1];
        bbase = last.block0 + last.nblocks;
    }
    TrackDesc *tds = L->tracks.data() + L->tracks_done;
    const int64_t nblocks = plan_chunk(offsets, n_tracks, geom_of_track, pbase, bbase, tds);
    if (bbase + nblocks > L->blocks_cap)
        return set_error(ST_EINVAL, "st_library_append: block table overflow");
    ChunkLayout lay;
    fill_chunk(x, y, z, edep, offsets, n_tracks, L->geoms.data(), precision, pbase, bbase, nblocks, tds, lay);
    std::vector<double> &sx = lay.x, &sy = lay.y, &sz = lay.z, &se = lay.e, &zfirst = lay.zfirst;
    std::vector<PointF32> &p32 = lay.p32;
    std::vector<BlockF32> &blocks = lay.blocks;
    int rc = ST_OK;
    if (!rc) rc = dev_put(L->x64 + pbase, sx.data(), n_points);
    if (!rc) rc = dev_put(L->y64 + pbase, sy.data(), n_points);
    if (!rc) rc = dev_put(L->z64 + pbase, sz.data(), n_points);
    if (!rc) rc = dev_put(L->e64 + pbase, se.data(), n_points);
    if (!rc) rc = dev_put(L->zfirst + bbase, zfirst.data(), nblocks);
    if (!rc && precision == ST_F32) {
        rc = dev_put(reinterpret_cast<PointF32 *>(L->p32) + pbase, p32.data(), n_points);
        if (!rc) rc = dev_put(L->blocks + bbase, blocks.data(), nblocks);
    }
    if (rc) return rc;
    L->tracks_done += n_tracks;
    L->points_done += n_points;
    return ST_OK;
}

int st_library_finish(st_library *L)
{
    if (!L) return set_error(ST_EINVAL, "st_library_finish: null library");
    if (L->finished) return ST_OK;
    if (L->tracks_done != L->n_tracks || L->points_done != L->n_points)
        return set_error(ST_EINVAL, "st_library_finish: %lld of %lld tracks and %lld of %lld points appended",
                         (long long)L->tracks_done, (long long)L->n_tracks, (long long)L->points_done,
                         (long long)L->n_points);
    L->n_blocks = L->n_tracks ? L->tracks.back().block0 + L->tracks.back().nblocks : 0;
    L->exact.clear();
    L->fast_order.clear();
    for (int64_t t = 0; t < L->n_tracks; ++t) {
        if (L->tracks[t].flags & kFlagExact) L->exact.push_back((int32_t)t);
        else L->fast_order.push_back((int32_t)t);
    }
    std::stable_sort(L->fast_order.begin(), L->fast_order.end(),
                     [&](int32_t a2, int32_t b2) { return L->tracks[a2].n > L->tracks[b2].n; });

    const int n_geoms = (int)L->geoms.size();
    std::vector<GeomDev> gd(n_geoms);
    for (int i = 0; i < n_geoms; ++i) {
        const st_geometry &g = L->geoms[i];
        gd[i].H = g.half_length;
        gd[i].L = g.length;
        gd[i].D = g.diameter;
        gd[i].invD = 1.0 / g.diameter;
        gd[i].r2 = g.radius * g.radius;
        gd[i].gso = g.gso;
        gd[i].denom = (2. / 3.) * g.diameter;
        gd[i].denom_inv = 1.0 / gd[i].denom;
        gd[i].yscale = gd[i].denom_inv * kTwo32;
        gd[i].qc = g.gso / g.diameter;
        gd[i].r2q = (g.radius * g.radius) / (g.diameter * g.diameter);
        {
            const double dl = kTwoM44 * (7.0 * g.half_length + std::fabs(g.gso) + g.diameter + 16.0);
            const double dlq = dl / g.diameter;
            gd[i].fm = 0.5 - (dlq + 1e-12);
            gd[i].mrgq = 3.0 * dlq * (1.0 + 2.0 * dlq) + 1e-14 * gd[i].r2q;
        }
        gd[i].N = (unsigned long long)g.n_linear;
        gd[i].N2 = gd[i].N * gd[i].N;
    }
    int rc = ST_OK;
    int64_t bytes = 0;
    if (!rc) rc = dev_upload(&L->tracks_dev, L->tracks.data(), L->tracks.size(), &bytes);
    if (!rc) rc = dev_upload(&L->geoms_dev, gd.data(), gd.size(), &bytes);
    if (!rc) rc = dev_upload(&L->order_dev, L->fast_order.data(), L->fast_order.size(), &bytes);
    L->device_bytes += bytes;
    if (rc) return rc;
    L->finished = true;
    return ST_OK;
}

int st_library_create(const double *x, const double *y, const double *z, const double *edep,
                      const int64_t *offsets, int64_t n_tracks, const st_geometry *geoms, int32_t n_geoms,
                      const int32_t *geom_of_track, int32_t precision, st_library **out)
{
    if (!out) return set_error(ST_EINVAL, "st_library_create: null output");
    *out = nullptr;
    if (n_tracks < 0 || !offsets) return set_error(ST_EINVAL, "st_library_create: bad arguments");
    const int64_t n_points = offsets[n_tracks] - offsets[0];
    if (n_points < 0) return set_error(ST_EINVAL, "st_library_create: bad offsets");
    st_library *L = nullptr;
    int rc = st_library_begin(n_tracks, n_points, geoms, n_geoms, precision, &L);
    if (rc) return rc;
    // append runs of tracks of at most kStagePoints points (or one larger
    // track), so host staging stays bounded
    constexpr int64_t kStagePoints = int64_t(1) << 25;
    for (int64_t t0 = 0; t0 < n_tracks && !rc;) {
        int64_t t1 = t0 + 1;
        while (t1 < n_tracks && offsets[t1 + 1] - offsets[t0] <= kStagePoints) ++t1;
        rc = st_library_append(L, x, y, z, edep, offsets + t0, t1 - t0,
                               geom_of_track ? geom_of_track + t0 : nullptr);
        t0 = t1;
    }
    if (!rc) rc = st_library_finish(L);
    if (rc) {
        st_library_destroy(L);
        return rc;
    }
    *out = L;
    return ST_OK;
}

int st_library_info_get(const st_library *L, st_library_info *info)
{
    if (!L || !info) return set_error(ST_EINVAL, "st_library_info_get: null argument");
    info->n_tracks = L->n_tracks;
    info->n_points = L->n_points;
    info->device_bytes = L->device_bytes;
    info->fast_tracks = (int64_t)L->fast_order.size();
    info->exact_tracks = (int64_t)L->exact.size();
    info->precision = L->precision;
    info->n_geometries = (int32_t)L->geoms.size();
    return ST_OK;
}

void st_library_destroy(st_library *L)
{
    if (!L) return;
    (void)hipDeviceSynchronize();
    void *ptrs[] = {L->x64, L->y64, L->z64, L->e64, L->p32,
                    L->zfirst, L->blocks, L->tracks_dev, L->geoms_dev, L->order_dev};
    for (void *p : ptrs) (void)hipFree(p);
    delete L;
}

int st_batch_create(st_library *lib, const st_batch_opts *opts, st_hist *const *hists, int32_t n_hists,
                    st_batch **out)
{
    if (!out) return set_error(ST_EINVAL, "st_batch_create: null output");
    *out = nullptr;
    if (!lib || !opts || (n_hists > 0 && !hists) || n_hists < 1 || n_hists > kMaxHists)
        return set_error(ST_EINVAL, "st_batch_create: need a library and 1..%d histograms", kMaxHists);
    if (opts->n_oversamples < 0) return set_error(ST_EINVAL, "st_batch_create: n_oversamples < 0");
    if (!lib->finished)
        return set_error(ST_EINVAL, "st_batch_create: library not finished (st_library_finish)");
    for (int i = 0; i < n_hists; ++i)
        if (!hists[i]) return set_error(ST_EINVAL, "st_batch_create: null histogram %d", i);
    st_batch *b = new (std::nothrow) st_batch;
    if (!b) return set_error(ST_ENOMEM, "st_batch_create: out of host memory");
    b->lib = lib;
    b->opts = *opts;
    b->hists.assign(hists, hists + n_hists);
    b->stream = default_stream();
    for (st_hist *h : b->hists) b->hist_words += h->nbins;
    int rc = ST_OK;
    auto fail = [&](int code) {
        st_batch_destroy(b);
        return code;
    };
    hipError_t e = dev_malloc((void **)&b->ctl, sizeof(BatchCtl));
    if (e == hipSuccess) e = hipMemset(b->ctl, 0, sizeof(BatchCtl));
    if (e != hipSuccess) return fail(set_error(hip_status(e), "st_batch_create: %s", hipGetErrorString(e)));
    for (const st_geometry &g : lib->geoms) {
        st_vcs *m = nullptr;
        rc = st_vcs_create(&g, opts->shift_xy, opts->shift_z, 0, 0, &m);
        if (rc) return fail(rc);
        st_vcs_set_stream(m, b->stream);
        b->methods.push_back(m);
    }
    rc = default_launch(lib->precision, b->hist_words, *opts, &b->ln);
    if (rc) return fail(rc);
    if (opts->autotune && !lib->fast_order.empty() && opts->n_oversamples > 0) {
        rc = autotune(b);
        if (rc) return fail(rc);
    }
    *out = b;
    return ST_OK;
}

int st_batch_run_static(st_library *lib, const st_batch_opts *opts, st_hist *const *hists, int32_t n_hists,
                        void *hip_stream, int64_t *stats7)
{
    if (!lib || !opts || !hists || n_hists < 1 || n_hists > kMaxHists)
        return set_error(ST_EINVAL, "st_batch_run_static: need a library, options and 1..%d histograms", kMaxHists);
    if (!lib->finished) return set_error(ST_EINVAL, "st_batch_run_static: library not finished (st_library_finish)");
    if (opts->n_oversamples < 0) return set_error(ST_EINVAL, "st_batch_run_static: n_oversamples < 0");
    if (opts->autotune)
        return set_error(ST_EINVAL, "st_batch_run_static: autotune times candidate passes on buffers it "
                                    "allocates; tune with st_batch_create and pass the chosen shape");
    if (!lib->exact.empty())
        return set_error(ST_ERANGE,
                         "st_batch_run_static: %zu tracks of the library need the exact per-segment path, "
                         "whose per-track buffers are allocated (limit: 0 such tracks); use st_batch_run",
                         lib->exact.size());
    int hist_words = 0;
    for (int i = 0; i < n_hists; ++i) {
        if (!hists[i]) return set_error(ST_EINVAL, "st_batch_run_static: null histogram %d", i);
        hist_words += hists[i]->nbins;
    }
    if (stats7)
        for (int i = 0; i < kStats; ++i) stats7[i] = 0;
    if (opts->n_oversamples == 0) return ST_OK;
    Launch ln;
    int rc = default_launch(lib->precision, hist_words, *opts, &ln);
    if (rc) return rc;
    BatchCtl *ctl = nullptr;
    rc = static_scratch(&ctl);
    if (rc) return rc;
    hipStream_t stream = hip_stream ? (hipStream_t)hip_stream : default_stream();
    if (!stream) return set_error(ST_EHIP, "st_batch_run_static: no stream");
    const BatchArgs a = make_args(lib, *opts, hists, n_hists, ln, ctl);
    rc = launch_pass(lib->precision, a, ln, stream);
    if (rc || !stats7) return rc;
    BatchCtl c;
    rc = read_ctl(ctl, stream, &c);
    if (rc) return rc;
    for (int i = 0; i < kStats; ++i) stats7[i] = (int64_t)c.stats[i];
    return check_ctl(c);
}

int st_batch_set_stream(st_batch *b, void *hip_stream)
{
    if (!b) return set_error(ST_EINVAL, "st_batch_set_stream: null batch");
    b->stream = hip_stream ? (hipStream_t)hip_stream : default_stream();
    for (st_vcs *m : b->methods) st_vcs_set_stream(m, b->stream);
    if (b->exec) {
        (void)hipGraphExecDestroy(b->exec);
        b->exec = nullptr;
    }
    if (b->graph) {
        (void)hipGraphDestroy(b->graph);
        b->graph = nullptr;
    }
    b->graph_ready = false;
    return ST_OK;
}

int st_batch_run(st_batch *b)
{
    if (!b) return set_error(ST_EINVAL, "st_batch_run: null batch");
    if (b->opts.n_oversamples == 0) return ST_OK;
    if (!b->ev[0])
        for (hipEvent_t &e : b->ev) ST_HIP_TRY(hipEventCreate(&e));
    ST_HIP_TRY(hipEventRecord(b->ev[0], b->stream));
    if (b->opts.use_graph && b->rec.cap == 0) {
        if (!b->graph_ready) {
            ST_HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            int rc = launch_fast(b);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(b->stream, &g);
            if (rc) return rc;
            if (e != hipSuccess) return set_error(ST_EHIP, "stream capture failed: %s", hipGetErrorString(e));
            b->graph = g;
            ST_HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
            b->graph_ready = true;
        }
        ST_HIP_TRY(hipGraphLaunch(b->exec, b->stream));
    } else {
        int rc = launch_fast(b);
        if (rc) return rc;
    }
    ST_HIP_TRY(hipEventRecord(b->ev[1], b->stream));
    int rc = run_exact(b);
    if (rc) return rc;
    ST_HIP_TRY(hipEventRecord(b->ev[2], b->stream));
    return ST_OK;
}

int st_batch_times(st_batch *b, float *ms2)
{
    if (!b || !ms2) return set_error(ST_EINVAL, "st_batch_times: null argument");
    if (!b->ev[0]) return set_error(ST_ESTATE, "st_batch_times before st_batch_run");
    ST_HIP_TRY(hipEventSynchronize(b->ev[2]));
    ST_HIP_TRY(hipEventElapsedTime(&ms2[0], b->ev[0], b->ev[1]));
    ST_HIP_TRY(hipEventElapsedTime(&ms2[1], b->ev[1], b->ev[2]));
    return ST_OK;
}

int st_batch_sync(st_batch *b)
{
    if (!b) return set_error(ST_EINVAL, "st_batch_sync: null batch");
    BatchCtl c;
    int rc = read_ctl(b->ctl, b->stream, &c);
    if (rc) return rc;
    return check_ctl(c);
}

int st_batch_stats(st_batch *b, int64_t *stats3)
{
    if (!b || !stats3) return set_error(ST_EINVAL, "st_batch_stats: null argument");
    BatchCtl c;
    int rc = read_ctl(b->ctl, b->stream, &c);
    if (rc) return rc;
    for (int i = 0; i < 3; ++i) stats3[i] = (int64_t)c.stats[i];
    return ST_OK;
}

int st_batch_balance(st_batch *b, uint64_t *balance4)
{
    if (!b || !balance4) return set_error(ST_EINVAL, "st_batch_balance: null argument");
    BatchCtl c;
    int rc = read_ctl(b->ctl, b->stream, &c);
    if (rc) return rc;
    for (int i = 0; i < 4; ++i) balance4[i] = c.stats[3 + i];
    return ST_OK;
}

int st_batch_params(const st_batch *b, int32_t *p)
{
    if (!b || !p) return set_error(ST_EINVAL, "st_batch_params: null argument");
    p[0] = kBT;
    p[1] = b->ln.slots;
    p[2] = b->ln.blocks_per_cu;
    p[3] = b->ln.ppt;
    p[4] = b->ln.lds_hist ? b->ln.hist_copies : 0;
    return ST_OK;
}

int st_batch_record(st_batch *b, int64_t capacity)
{
    if (!b || capacity < 0) return set_error(ST_EINVAL, "st_batch_record: bad arguments");
    (void)hipFree(b->rec.seg);
    (void)hipFree(b->rec.key);
    (void)hipFree(b->rec.box);
    (void)hipFree(b->rec.sum);
    (void)hipFree(b->rec_dev);
    b->rec = RecArgs{};
    b->rec_dev = nullptr;
    if (capacity == 0) return ST_OK;
    ST_ALLOC_TRY(dev_malloc((void **)&b->rec.seg, capacity * sizeof(int64_t)));
    ST_ALLOC_TRY(dev_malloc((void **)&b->rec.key, capacity * sizeof(unsigned long long)));
    ST_ALLOC_TRY(dev_malloc((void **)&b->rec.box, 3 * capacity * sizeof(unsigned long long)));
    ST_ALLOC_TRY(dev_malloc((void **)&b->rec.sum, capacity * sizeof(double)));
    ST_ALLOC_TRY(dev_malloc((void **)&b->rec_dev, sizeof(RecArgs)));
    b->rec.cap = capacity;
    ST_HIP_TRY(hipMemcpy(b->rec_dev, &b->rec, sizeof(RecArgs), hipMemcpyHostToDevice));
    return ST_OK;
}

int st_batch_record_fetch(st_batch *b, int64_t *segments, uint64_t *keys, double *sums, int64_t *count)
{
    if (!b || !count) return set_error(ST_EINVAL, "st_batch_record_fetch: null argument");
    BatchCtl ctl;
    int rc = read_ctl(b->ctl, b->stream, &ctl);
    if (rc) return rc;
    const unsigned long long c = ctl.rec_count;
    *count = (int64_t)c;
    if ((int64_t)c > b->rec.cap) return set_error(ST_ERANGE, "st_batch_record_fetch: %llu records, capacity %lld",
                                                 c, (long long)b->rec.cap);
    // segment id (track * n_os + oversample) from the ticket: the ticket
    // counts segments in the order of the fast tracks
    const int64_t n_os = std::max(b->opts.n_oversamples, 1);
    std::vector<int64_t> seg(c);
    if (c) ST_HIP_TRY(hipMemcpy(seg.data(), b->rec.seg, c * sizeof(int64_t), hipMemcpyDeviceToHost));
    for (unsigned long long i = 0; i < c; ++i)
        seg[i] = (int64_t)b->lib->fast_order[seg[i] / n_os] * n_os + seg[i] % n_os;
    if (c && segments) std::memcpy(segments, seg.data(), c * sizeof(int64_t));
    if (c && keys) {
        // the reference's key from the local key and its slab's box
        std::vector<unsigned long long> loc(c), bx(3 * c);
        ST_HIP_TRY(hipMemcpy(loc.data(), b->rec.key, c * sizeof(uint64_t), hipMemcpyDeviceToHost));
        ST_HIP_TRY(hipMemcpy(bx.data(), b->rec.box, 3 * c * sizeof(uint64_t), hipMemcpyDeviceToHost));
        for (unsigned long long i = 0; i < c; ++i) {
            const TrackDesc &td = b->lib->tracks[seg[i] / n_os];
            const unsigned long long N = (unsigned long long)b->lib->geoms[td.geom].n_linear;
            const unsigned long long kx0 = bx[3 * i] & 0xFFFFFFFFull, ky0 = bx[3 * i] >> 32;
            const unsigned long long kz0 = bx[3 * i + 1] & 0xFFFFFFFFull, nx = bx[3 * i + 1] >> 32;
            const unsigned long long ny = bx[3 * i + 2];
            const unsigned long long lx = loc[i] % nx, ly = (loc[i] / nx) % ny, lz = loc[i] / (nx * ny);
            keys[i] = (kx0 + lx) + (ky0 + ly) * N + (kz0 + lz) * N * N;
        }
    }
    if (c && sums) ST_HIP_TRY(hipMemcpy(sums, b->rec.sum, c * sizeof(double), hipMemcpyDeviceToHost));
    return ST_OK;
}

void st_batch_destroy(st_batch *b)
{
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (st_vcs *m : b->methods) st_vcs_destroy(m);
    for (hipEvent_t e : b->ev)
        if (e) (void)hipEventDestroy(e);
    st_track_free(&b->scratch);
    st_pairs_free(&b->pairs);
    void *ptrs[] = {b->ctl, b->rec.seg, b->rec.key, b->rec.box, b->rec.sum, b->rec_dev};
    for (void *p : ptrs) (void)hipFree(p);
    delete b;
}

}  // extern "C"
