// synth.cpp -- deterministic synthetic track library (host code).
//
// Each track is a straight primary along +z through (x0, y0) ~ U(-h, h)^2,
// entering at z = -2 * entry_half_length (the voxel's -z face once the
// method adds H').  Events follow at exponential spacing; each event deposits
// one point and, with probability p, starts a delta-ray cluster of
// Geometric(mean) points placed by a 3-D Gaussian random walk.  Energies are
// lognormal, clamped.  All randomness is Philox4x32-10 keyed by
// (0x53545254, stream_id) with counter (track, draw, 0, 0), turned into
// numbers by the inverse transforms below, so a track is a pure function of
// (spec, track index) and tracks can be generated in any order or in parallel.
#include <atomic>
#include <cmath>
#include <thread>
#include <vector>

#include "st_common.hpp"
#include "st_internal.hpp"

namespace st {

namespace {

class Stream {
public:
    Stream(uint32_t stream_id, uint64_t track) : k1_(stream_id), track_((uint32_t)track) {}

    uint32_t next_u32()
    {
        if (pos_ == 4) {
            u32x4 c = {{track_, (uint32_t)draw_, (uint32_t)(draw_ >> 32), 0u}};
            block_ = philox4x32_10(c, 0x53545254u, k1_);
            ++draw_;
            pos_ = 0;
        }
        return block_.v[pos_++];
    }

    // uniform double in (0, 1) with 53 random bits
    double uniform()
    {
        const uint64_t hi = next_u32() >> 5;   // 27 bits
        const uint64_t lo = next_u32() >> 6;   // 26 bits
        return ((double)((hi << 26) | lo) + 0.5) * (1.0 / 9007199254740992.0);
    }

    double normal()
    {
        if (have_spare_) {
            have_spare_ = false;
            return spare_;
        }
        const double u1 = uniform();
        const double u2 = uniform();
        const double r = std::sqrt(-2.0 * std::log(u1));
        const double th = 6.283185307179586 * u2;
        spare_ = r * std::sin(th);
        have_spare_ = true;
        return r * std::cos(th);
    }

private:
    uint32_t k1_;
    uint32_t track_;
    uint64_t draw_ = 0;
    u32x4 block_{};
    int pos_ = 4;
    bool have_spare_ = false;
    double spare_ = 0.0;
};

struct TrackParams {
    int64_t n_target;   // 0: length-limited
    double p;
    double length;
    double spacing;
    double x0, y0;
};

TrackParams track_params(const st_synth_spec &s, Stream &rng)
{
    TrackParams t{};
    const double un = rng.uniform();
    const double up = rng.uniform();
    const double ul = rng.uniform();
    const double ux = rng.uniform();
    const double uy = rng.uniform();
    t.p = (s.p_cluster_max > s.p_cluster) ? s.p_cluster + up * (s.p_cluster_max - s.p_cluster) : s.p_cluster;
    t.length = s.length_min_nm + ul * (s.length_max_nm - s.length_min_nm);
    t.x0 = -s.voxel_half_length + 2.0 * s.voxel_half_length * ux;
    t.y0 = -s.voxel_half_length + 2.0 * s.voxel_half_length * uy;
    if (s.n_min > 0) {
        const double lo = std::log((double)s.n_min), hi = std::log((double)std::max(s.n_max, s.n_min));
        int64_t n = (int64_t)std::llround(std::exp(lo + un * (hi - lo)));
        if (n < s.n_min) n = s.n_min;
        if (n > std::max(s.n_max, s.n_min)) n = std::max(s.n_max, s.n_min);
        t.n_target = n;
        const double per_event = 1.0 + t.p * s.cluster_mean;
        double sp = t.length * per_event / (double)n;
        if (sp < s.spacing_min_nm) sp = s.spacing_min_nm;
        if (sp > s.spacing_max_nm) sp = s.spacing_max_nm;
        t.spacing = sp;
    } else {
        t.n_target = 0;
        t.spacing = s.spacing_nm;
    }
    return t;
}

// Generates one track; with out == nullptr only counts.
int64_t generate(const st_synth_spec &s, int64_t track, double *x, double *y, double *z, double *e)
{
    Stream rng(s.stream_id, (uint64_t)track);
    const TrackParams t = track_params(s, rng);
    const double z_entry = -2.0 * s.entry_half_length;
    const double z_end = z_entry + t.length;
    const double q = 1.0 / s.cluster_mean;
    const double log1mq = std::log1p(-q);
    const double mu_ln = std::log(s.edep_median_ev);
    int64_t m = 0;
    auto edep = [&]() {
        double v = std::exp(mu_ln + s.edep_sigma_ln * rng.normal());
        if (v < s.edep_min_ev) v = s.edep_min_ev;
        if (v > s.edep_max_ev) v = s.edep_max_ev;
        return v;
    };
    auto emit = [&](double px, double py, double pz) {
        const double ev = edep();
        if (x) {
            x[m] = px;
            y[m] = py;
            z[m] = pz;
            e[m] = ev;
        }
        ++m;
    };
    double zc = z_entry;
    for (;;) {
        zc += -t.spacing * std::log(rng.uniform());
        if (t.n_target == 0 && zc > z_end) break;
        emit(t.x0, t.y0, zc);
        if (t.n_target && m >= t.n_target) break;
        if (rng.uniform() < t.p) {
            const int64_t k = 1 + (int64_t)std::floor(std::log(rng.uniform()) / log1mq);
            double px = t.x0, py = t.y0, pz = zc;
            bool full = false;
            for (int64_t j = 0; j < k; ++j) {
                px += s.walk_sigma_nm * rng.normal();
                py += s.walk_sigma_nm * rng.normal();
                pz += s.walk_sigma_nm * rng.normal();
                emit(px, py, pz);
                if (t.n_target && m >= t.n_target) {
                    full = true;
                    break;
                }
            }
            if (full) break;
        }
    }
    return m;
}

int check_spec(const st_synth_spec *s)
{
    if (!s) return set_error(ST_EINVAL, "synth: null spec");
    if (!(s->cluster_mean >= 1.0) || !(s->walk_sigma_nm >= 0) || !(s->edep_median_ev > 0) ||
        !(s->length_max_nm >= s->length_min_nm))
        return set_error(ST_EINVAL, "synth: invalid spec");
    if (s->n_min <= 0 && !(s->spacing_nm > 0))
        return set_error(ST_EINVAL, "synth: length-limited mode needs spacing_nm > 0");
    if (s->n_min > 0 && !(s->spacing_min_nm > 0 && s->spacing_max_nm >= s->spacing_min_nm))
        return set_error(ST_EINVAL, "synth: count mode needs 0 < spacing_min <= spacing_max");
    return ST_OK;
}

}  // namespace
}  // namespace st

using namespace st;

extern "C" {

int st_synth_counts(const st_synth_spec *spec, int64_t t0, int64_t t1, int64_t *counts)
{
    int rc = check_spec(spec);
    if (rc) return rc;
    if (!counts || t1 < t0) return set_error(ST_EINVAL, "st_synth_counts: bad range");
    const st_synth_spec s = *spec;
    parallel_for(t1 - t0, [&](int64_t i) {
        if (s.n_min > 0) {
            Stream rng(s.stream_id, (uint64_t)(t0 + i));
            counts[i] = track_params(s, rng).n_target;
        } else {
            counts[i] = generate(s, t0 + i, nullptr, nullptr, nullptr, nullptr);
        }
    });
    return ST_OK;
}

int st_synth_fill(const st_synth_spec *spec, int64_t t0, int64_t t1, const int64_t *offsets,
                  double *x, double *y, double *z, double *edep)
{
    int rc = check_spec(spec);
    if (rc) return rc;
    if (!offsets || !x || !y || !z || !edep || t1 < t0)
        return set_error(ST_EINVAL, "st_synth_fill: bad arguments");
    const st_synth_spec s = *spec;
    std::atomic<int> bad{0};
    parallel_for(t1 - t0, [&](int64_t i) {
        const int64_t o = offsets[i] - offsets[0];
        const int64_t m = generate(s, t0 + i, x + o, y + o, z + o, edep + o);
        if (m != offsets[i + 1] - offsets[i]) bad.store(1);
    });
    if (bad.load()) return set_error(ST_EINVAL, "st_synth_fill: offsets do not match st_synth_counts");
    return ST_OK;
}

}  // extern "C"
