// layout_check.cpp -- CPU checks of the batched library's host-side layout
// (supertrack_amd/csrc/layout.{hpp,cpp}), built with g++ alone (no HIP) by
// tests/test_layout.py and, with AddressSanitizer and UBSan, by
// tests/sanitizers.py.  Prints one line per check and "ALL OK" at the end;
// exits non-zero on the first failure.
#include <algorithm>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../supertrack_amd/csrc/layout.hpp"

using namespace st;

static void fail(const char *what, long long a = 0, long long b = 0)
{
    std::printf("FAIL %s (%lld, %lld)\n", what, a, b);
    std::exit(1);
}

static st_geometry geometry(double half_length, double diameter)
{
    // SphericalGeometry (SphericalGeometry.cu:34-58), as st_geometry_make
    st_geometry g{};
    const double r = diameter / 2;
    int N = (int)std::floor((half_length * 2) / diameter);
    double L = N * diameter, H = L / 2, gso = -H + r;
    if (r > H) {
        H = r;
        N = 1;
        L = H * 2;
        gso = -H + r;
    }
    g.half_length = H;
    g.length = L;
    g.diameter = diameter;
    g.radius = r;
    g.gso = gso;
    g.n_linear = N;
    return g;
}

struct Lib {
    std::vector<double> x, y, z, e;
    std::vector<int64_t> off{0};
    std::vector<int32_t> geom;
};

// a track: primary along z with clusters, like the synthetic library, plus
// optional special points
static void add_track(Lib &L, std::mt19937_64 &rng, int64_t n, double spread, int geom)
{
    std::normal_distribution<double> nd(0.0, spread);
    std::uniform_real_distribution<double> ud(-2500.0, 2500.0);
    const double x0 = ud(rng) * 0.5, y0 = ud(rng) * 0.5;
    for (int64_t i = 0; i < n; ++i) {
        L.x.push_back(x0 + nd(rng));
        L.y.push_back(y0 + nd(rng));
        L.z.push_back(-5000.0 + 5000.0 * std::uniform_real_distribution<double>(0, 1)(rng));
        L.e.push_back(1.0 + 100.0 * std::uniform_real_distribution<double>(0, 1)(rng));
    }
    L.off.push_back((int64_t)L.x.size());
    L.geom.push_back(geom);
}

// The window bound restated by brute force: for every i, the points j >= i
// (z order) with z[j] - z[i] <= w; min(count, x-cells * y-cells * layers).
static int64_t window_bound_brute(const double *x, const double *y, const double *z, int64_t n, double D, double w,
                                  int64_t layers)
{
    int64_t best = 0;
    for (int64_t i = 0; i < n; ++i) {
        double xl = x[i], xh = x[i], yl = y[i], yh = y[i];
        int64_t cnt = 0;
        for (int64_t j = i; j < n && z[j] - z[i] <= w; ++j) {
            xl = std::min(xl, x[j]);
            xh = std::max(xh, x[j]);
            yl = std::min(yl, y[j]);
            yh = std::max(yh, y[j]);
            ++cnt;
        }
        const double cells = (std::floor((xh - xl) / D) + 2.0) * (std::floor((yh - yl) / D) + 2.0) * (double)layers;
        best = std::max(best, cells < (double)cnt ? (int64_t)cells : cnt);
    }
    return best;
}

static void check_layout(const Lib &L, const std::vector<st_geometry> &geoms, int precision, int64_t pbase,
                         int64_t bbase, const char *name)
{
    const int64_t nt = (int64_t)L.off.size() - 1;
    std::vector<TrackDesc> tds(nt);
    const int64_t nblocks = plan_chunk(L.off.data(), nt, L.geom.data(), pbase, bbase, tds.data());
    ChunkLayout out;
    fill_chunk(L.x.data(), L.y.data(), L.z.data(), L.e.data(), L.off.data(), nt, geoms.data(), precision, pbase, bbase,
               nblocks, tds.data(), out);
    int64_t blocks_seen = 0;
    for (int64_t t = 0; t < nt; ++t) {
        const TrackDesc &td = tds[t];
        const int64_t n = L.off[t + 1] - L.off[t];
        if (td.offset != pbase + L.off[t] - L.off[0]) fail("track offset", t, td.offset);
        if (td.n != n) fail("track size", t, td.n);
        if (td.block0 != bbase + blocks_seen || td.nblocks != (n + kBlk - 1) / kBlk) fail("block range", t, td.block0);
        blocks_seen += td.nblocks;
        const int64_t o = td.offset - pbase;
        // a permutation of the track's points, z ascending, NaN last
        std::vector<double> a(L.z.begin() + L.off[t], L.z.begin() + L.off[t + 1]);
        std::vector<double> b(out.z.begin() + o, out.z.begin() + o + n);
        for (int64_t k = 1; k < n; ++k)
            if (!std::isnan(b[k]) && (std::isnan(b[k - 1]) || b[k] < b[k - 1])) fail("z order", t, k);
        auto key = [](double v) { return std::isnan(v) ? HUGE_VAL : v; };
        std::vector<double> ka, kb;
        for (double v : a) ka.push_back(key(v));
        for (double v : b) kb.push_back(key(v));
        std::sort(ka.begin(), ka.end());
        std::sort(kb.begin(), kb.end());
        if (ka != kb) fail("z permutation", t);
        double xs = 0, xo = 0;
        for (int64_t k = 0; k < n; ++k) {
            xs += L.x[L.off[t] + k];
            xo += out.x[o + k];
        }
        if (std::fabs(xs - xo) > 1e-6 * (std::fabs(xs) + 1.0) && std::isfinite(xs)) fail("x carried with z", t);
        for (int b2 = 0; b2 < td.nblocks; ++b2)
            if (!(out.zfirst[td.block0 - bbase + b2] == out.z[o + (int64_t)b2 * kBlk]) &&
                !std::isnan(out.zfirst[td.block0 - bbase + b2]))
                fail("block start", t, b2);
        if (precision == ST_F32) {
            for (int b2 = 0; b2 < td.nblocks; ++b2) {
                const BlockF32 &bk = out.blocks[td.block0 - bbase + b2];
                if (!(bk.dl > 0)) fail("block error bound", t, b2);
                const st_geometry &g = geoms[td.geom];
                const double ox = bk.oqx * g.diameter + g.gso;   // block origin, within rounding
                for (int64_t p = o + (int64_t)b2 * kBlk; p < std::min<int64_t>(o + (int64_t)(b2 + 1) * kBlk, o + n);
                     ++p) {
                    const double d = out.x[p] - ox;
                    if (std::isfinite(d) && std::fabs((double)out.p32[p].x - d) > bk.dl + 1e-9 * std::fabs(ox))
                        fail("fp32 offset within the block bound", t, p);
                }
            }
        }
    }
    if (blocks_seen != nblocks) fail("block total", blocks_seen, nblocks);
    std::printf("ok %s: %lld tracks, %lld points, %lld blocks, pbase %lld\n", name, (long long)nt,
                (long long)L.x.size(), (long long)nblocks, (long long)pbase);
}

int main()
{
    std::mt19937_64 rng(7);
    const std::vector<st_geometry> geoms = {geometry(2500.0, 2.0), geometry(2500.0, 10.0), geometry(1000.0, 1.0)};
    // 1. ordinary tracks of every size class around the block edges
    Lib L;
    for (int64_t n : {0, 1, 255, 256, 257, 511, 512, 513, 3000, 70000}) add_track(L, rng, n, 4.0, (int)(n % 3));
    for (int precision : {ST_F64, ST_F32}) {
        check_layout(L, geoms, precision, 0, 0, precision == ST_F32 ? "ragged F32" : "ragged F64");
        // offsets beyond 2^32 points and 2^24 blocks (a library of > 4e9 points)
        check_layout(L, geoms, precision, (int64_t)5000000000LL, (int64_t)20000000, "far offsets");
    }
    // 2. eligibility: a NaN point, an infinite energy, a dense plane, a
    //    track above kMaxFastPoints
    {
        Lib M;
        add_track(M, rng, 1000, 4.0, 1);
        M.x[10] = NAN;
        add_track(M, rng, 1000, 4.0, 1);
        M.e.back() = INFINITY;
        // a plane of 100 x 100 points 1 nm apart at one z: 10^4 spheres of 1 nm
        for (int i = 0; i < 100; ++i)
            for (int j = 0; j < 100; ++j) {
                M.x.push_back(i * 1.0);
                M.y.push_back(j * 1.0);
                M.z.push_back(-2500.0);
                M.e.push_back(10.0);
            }
        M.off.push_back((int64_t)M.x.size());
        M.geom.push_back(2);
        add_track(M, rng, kMaxFastPoints + 1, 4.0, 1);
        add_track(M, rng, 5000, 4.0, 1);   // ordinary: fast
        const int64_t nt = (int64_t)M.off.size() - 1;
        std::vector<TrackDesc> tds(nt);
        const int64_t nb = plan_chunk(M.off.data(), nt, M.geom.data(), 0, 0, tds.data());
        ChunkLayout out;
        fill_chunk(M.x.data(), M.y.data(), M.z.data(), M.e.data(), M.off.data(), nt, geoms.data(), ST_F32, 0, 0, nb,
                   tds.data(), out);
        const uint32_t want[] = {kFlagExact, kFlagExact, kFlagExact, kFlagExact, 0u};
        for (int64_t t = 0; t < nt; ++t)
            if (tds[t].flags != want[t]) fail("eligibility", t, tds[t].flags);
        std::printf("ok eligibility: NaN, infinite energy, dense plane, %lld points -> exact path; ordinary -> fast\n",
                    (long long)(kMaxFastPoints + 1));
    }
    // 2b. eligibility: one layer of the track's x/y extent must hold at most
    //     2^32 - 2 spheres (32-bit sphere keys); only possible for N > 65535
    {
        const std::vector<st_geometry> gw = {geometry(50000.0, 1.0)};   // N = 100000
        Lib M;
        auto two_points = [&](double span) {
            M.x.push_back(-span / 2);
            M.y.push_back(-span / 2);
            M.z.push_back(0.0);
            M.e.push_back(10.0);
            M.x.push_back(span / 2);
            M.y.push_back(span / 2);
            M.z.push_back(1.0);
            M.e.push_back(10.0);
            M.off.push_back((int64_t)M.x.size());
            M.geom.push_back(0);
        };
        two_points(65536.0);   // 65542^2 > 2^32 - 2: exact path
        two_points(65520.0);   // 65526^2 < 2^32 - 2: fast
        const int64_t nt = (int64_t)M.off.size() - 1;
        std::vector<TrackDesc> tds(nt);
        const int64_t nb = plan_chunk(M.off.data(), nt, M.geom.data(), 0, 0, tds.data());
        ChunkLayout out;
        fill_chunk(M.x.data(), M.y.data(), M.z.data(), M.e.data(), M.off.data(), nt, gw.data(), ST_F32, 0, 0, nb,
                   tds.data(), out);
        if (tds[0].flags != kFlagExact || tds[1].flags != 0u) fail("key-box eligibility", tds[0].flags, tds[1].flags);
        std::printf("ok eligibility: a layer of 65542^2 spheres -> exact path; 65526^2 -> fast\n");
    }
    // 3. the sliding-window bound against brute force
    for (int rep = 0; rep < 40; ++rep) {
        const int64_t n = 50 + (int64_t)(rng() % 600);
        std::vector<double> x(n), y(n), z(n);
        std::normal_distribution<double> nd(0.0, 1.0 + (double)(rep % 7) * 3.0);
        for (int64_t i = 0; i < n; ++i) {
            x[i] = nd(rng);
            y[i] = nd(rng);
            z[i] = (double)(rng() % 4000) * 0.05;
        }
        std::sort(z.begin(), z.end());
        const double D = 0.5 + (double)(rep % 5);
        const double w = 2.0 * D;
        const int64_t got = window_bound(x.data(), y.data(), z.data(), n, D, w, 3);
        const int64_t want = window_bound_brute(x.data(), y.data(), z.data(), n, D, w, 3);
        // the fast version stops once the bound exceeds kMaxWindow
        if (!(got == want || (want > kMaxWindow && got > kMaxWindow))) fail("window bound", got, want);
    }
    std::printf("ok window bound: 40 random tracks against brute force\n");
    std::printf("ALL OK\n");
    return 0;
}
