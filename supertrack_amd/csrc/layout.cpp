// layout.cpp -- host-side layout of a batched track library (layout.hpp).
#include "layout.hpp"

#include <algorithm>
#include <cmath>
#include <deque>
#include <numeric>

#include "parallel.hpp"

namespace st {

int64_t window_bound(const double *x, const double *y, const double *z, int64_t n, double D, double w,
                     int64_t layers)
{
    if (n == 0) return 0;
    std::deque<int64_t> xmax, xmin, ymax, ymin;
    int64_t best = 0;
    int64_t j = 0;
    for (int64_t i = 0; i < n; ++i) {
        while (j < n && z[j] - z[i] <= w) {
            while (!xmax.empty() && x[xmax.back()] <= x[j]) xmax.pop_back();
            xmax.push_back(j);
            while (!xmin.empty() && x[xmin.back()] >= x[j]) xmin.pop_back();
            xmin.push_back(j);
            while (!ymax.empty() && y[ymax.back()] <= y[j]) ymax.pop_back();
            ymax.push_back(j);
            while (!ymin.empty() && y[ymin.back()] >= y[j]) ymin.pop_back();
            ymin.push_back(j);
            ++j;
        }
        const int64_t cnt = j - i;
        const double cx = std::floor((x[xmax.front()] - x[xmin.front()]) / D) + 2.0;
        const double cy = std::floor((y[ymax.front()] - y[ymin.front()]) / D) + 2.0;
        const double cells = cx * cy * (double)layers;
        const int64_t b = cells < (double)cnt ? (int64_t)cells : cnt;
        best = std::max(best, b);
        if (best >= kMaxWindow + 1) return best;   // already ineligible
        for (auto *dq : {&xmax, &xmin, &ymax, &ymin})
            while (!dq->empty() && dq->front() <= i) dq->pop_front();
    }
    return best;
}

int64_t plan_chunk(const int64_t *offsets, int64_t n_tracks, const int32_t *geom_of_track, int64_t pbase,
                   int64_t bbase, TrackDesc *tds)
{
    int64_t nblocks = 0;
    for (int64_t t = 0; t < n_tracks; ++t) {
        TrackDesc &td = tds[t];
        td.offset = pbase + (offsets[t] - offsets[0]);
        td.n = (int32_t)(offsets[t + 1] - offsets[t]);
        td.nblocks = (td.n + kBlk - 1) / kBlk;
        td.block0 = bbase + nblocks;
        td.geom = geom_of_track ? geom_of_track[t] : 0;
        td.flags = 0;
        nblocks += td.nblocks;
    }
    return nblocks;
}

void fill_chunk(const double *x, const double *y, const double *z, const double *edep, const int64_t *offsets,
                int64_t n_tracks, const st_geometry *geoms, int precision, int64_t pbase, int64_t bbase,
                int64_t nblocks, TrackDesc *tds, ChunkLayout &out)
{
    const int64_t n_points = offsets[n_tracks] - offsets[0];
    out.x.assign(n_points, 0.0);
    out.y.assign(n_points, 0.0);
    out.z.assign(n_points, 0.0);
    out.e.assign(n_points, 0.0);
    out.zfirst.assign(std::max<int64_t>(nblocks, 1), 0.0);
    out.p32.clear();
    out.blocks.clear();
    if (precision == ST_F32) {
        out.p32.resize(n_points);
        out.blocks.resize(std::max<int64_t>(nblocks, 1));
    }
    const double *X = x, *Y = y, *Z = z, *E = edep;
    const int64_t base = offsets[0];
    parallel_for(n_tracks, [&](int64_t t) {
        TrackDesc &td = tds[t];
        const int64_t o = td.offset - pbase, n = td.n;
        const int64_t b0 = td.block0 - bbase;
        std::vector<int64_t> idx(n);
        std::iota(idx.begin(), idx.end(), 0);
        const double *zt = Z + base + o;
        auto zkey = [&](int64_t i) { return std::isnan(zt[i]) ? HUGE_VAL : zt[i]; };
        std::stable_sort(idx.begin(), idx.end(), [&](int64_t a2, int64_t b2) { return zkey(a2) < zkey(b2); });
        bool finite = true;
        double xmin = HUGE_VAL, xmax = -HUGE_VAL, ymin = HUGE_VAL, ymax = -HUGE_VAL, esum = 0.0, emax = 0.0;
        for (int64_t k = 0; k < n; ++k) {
            const int64_t s = base + o + idx[k];
            out.x[o + k] = X[s];
            out.y[o + k] = Y[s];
            out.z[o + k] = Z[s];
            out.e[o + k] = E[s];
            finite = finite && std::isfinite(X[s]) && std::isfinite(Y[s]) && std::isfinite(Z[s]) &&
                     std::isfinite(E[s]);
            xmin = std::min(xmin, X[s]);
            xmax = std::max(xmax, X[s]);
            ymin = std::min(ymin, Y[s]);
            ymax = std::max(ymax, Y[s]);
            esum += std::fabs(E[s]);
            emax = std::max(emax, std::fabs(E[s]));
        }
        td.xmin = xmin;
        td.xmax = xmax;
        td.ymin = ymin;
        td.ymax = ymax;
        td.zmax = n ? out.z[o + n - 1] : 0.0;
        for (int b = 0; b < td.nblocks; ++b) out.zfirst[b0 + b] = out.z[o + (int64_t)b * kBlk];
        const st_geometry &g = geoms[td.geom];
        const double denom = (2. / 3.) * g.diameter;
        // fixed-point headroom: sum of |y| over the track
        const double ysum = esum / denom;
        const double ylim = (precision == ST_F64) ? 2.7e11 : 1.0e9;
        const double ypoint = (precision == ST_F64) ? kMaxPointY64 : kMaxPointY32;
        bool fast = finite && n <= kMaxFastPoints && ysum < ylim && emax / denom < ypoint;
        if (fast && n > 0) {
            // the fast path reads sphere indices from the low words of the
            // rounded lattice coordinates: |q| < 2^30 for every shift |s| <= H'
            const double cmax = std::max({std::fabs(xmin), std::fabs(xmax), std::fabs(ymin), std::fabs(ymax),
                                          std::fabs(out.z[o] + g.half_length),
                                          std::fabs(out.z[o + n - 1] + g.half_length)});
            fast = (cmax + g.half_length + std::fabs(g.gso)) / g.diameter < 1073741824.0;
            // one layer of the segment's key box (x-cells times y-cells, for
            // any shift) must fit the 32-bit keys of the sphere table
            const double nl = (double)g.n_linear;
            const double cx = std::min(nl, std::floor((xmax - xmin) / g.diameter) + 6.0);
            const double cy = std::min(nl, std::floor((ymax - ymin) / g.diameter) + 6.0);
            fast = fast && cx * cy <= 4294967294.0;
        }
        if (fast) {
            // entries kept by a flush come from the next point's layer (or, for
            // keys with an index equal to N, the layer below): a z-window two
            // diameters wide, touching at most three layers
            const double w = 2.0 * g.diameter * (1.0 + 1e-12) + 1e-9;
            const int64_t W = window_bound(out.x.data() + o, out.y.data() + o, out.z.data() + o, n, g.diameter, w, 3);
            fast = W <= kMaxWindow;
        }
        td.flags = fast ? 0u : kFlagExact;
        if (precision == ST_F32) {
            for (int b = 0; b < td.nblocks; ++b) {
                const int64_t p0 = o + (int64_t)b * kBlk, p1 = std::min<int64_t>(p0 + kBlk, o + n);
                double lo3[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, hi3[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
                const double *arr[3] = {out.x.data(), out.y.data(), out.z.data()};
                for (int64_t p = p0; p < p1; ++p)
                    for (int a3 = 0; a3 < 3; ++a3) {
                        lo3[a3] = std::min(lo3[a3], arr[a3][p]);
                        hi3[a3] = std::max(hi3[a3], arr[a3][p]);
                    }
                double org[3];
                for (int a3 = 0; a3 < 3; ++a3) {
                    org[a3] = (std::isfinite(lo3[a3]) && std::isfinite(hi3[a3])) ? 0.5 * (lo3[a3] + hi3[a3]) : 0.0;
                }
                double err = 0.0;
                for (int64_t p = p0; p < p1; ++p) {
                    for (int a3 = 0; a3 < 3; ++a3) {
                        const double d = arr[a3][p] - org[a3];
                        const float f = (float)d;
                        (a3 == 0 ? out.p32[p].x : a3 == 1 ? out.p32[p].y : out.p32[p].z) = f;
                        const double e1 = std::fabs((double)f - d) + std::fabs(d) * 2.3e-16;
                        if (std::isfinite(e1)) err = std::max(err, e1);
                        else err = HUGE_VAL;
                    }
                    out.p32[p].w = (float)out.e[p];
                }
                // dl: the FP32 rounding of the offsets plus, with a wide margin,
                // the FP64 rounding of each add and product on the way to the
                // sphere test (|shift| <= H' on each axis)
                BlockF32 &bk = out.blocks[b0 + b];
                const double invD = 1.0 / g.diameter;
                bk.oqx = (org[0] - g.gso) * invD;
                bk.oqy = (org[1] - g.gso) * invD;
                bk.oqz = ((org[2] + g.half_length) - g.gso) * invD;
                double dl = err + kTwoM44 * (std::fabs(org[0]) + std::fabs(org[1]) + std::fabs(org[2]) +
                                             7.0 * g.half_length + std::fabs(g.gso) + g.diameter + 16.0);
                if (!std::isfinite(dl)) dl = HUGE_VAL;
                const double dlq = dl * invD;
                const double r2q = (g.radius * g.radius) / (g.diameter * g.diameter);
                bk.dl = dl;
                bk.fm = 0.5 - (dlq + 1e-12);
                bk.mrgq = 3.0 * dlq * (1.0 + 2.0 * dlq) + 1e-14 * r2q;
            }
        }
    });

}

}  // namespace st
