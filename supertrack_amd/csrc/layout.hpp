// layout.hpp -- host-side layout of a batched track library: per-track
// records, each track's points sorted by z, per-block search keys and, for
// FP32 storage, block-relative FP32 points with per-block error bounds, and
// each track's eligibility for the fused kernel.  Plain C++ (no HIP): used by
// batch.hip to build the device library and by the CPU tests on its own.
#pragma once

#include <cstdint>
#include <vector>

#include "../../include/supertrack_amd.h"

namespace st {

constexpr int kBlk = 256;                              // points per library block
constexpr int64_t kMaxWindow = 768;                    // fast-path bound on spheres in a z-window two diameters wide
constexpr int64_t kMaxFastPoints = (int64_t)1 << 22;   // fast-path bound on points per track
constexpr uint32_t kFlagExact = 1u;                    // TrackDesc::flags: the exact per-segment path
constexpr double kTwoM44 = 1.0 / 17592186044416.0;
// Per-point lineal energy limits of the fused kernel (its fixed-point
// conversions are exact only under 2^51): ST_F32 y * 2^32, ST_F64 y * 2^24.
constexpr double kMaxPointY32 = 262144.0;      // 2^18
constexpr double kMaxPointY64 = 67108864.0;    // 2^26

struct TrackDesc {
    int64_t offset;     // first point in the sorted arrays
    int64_t block0;     // first block in the block arrays
    int32_t n;
    int32_t nblocks;
    int32_t geom;
    uint32_t flags;
    double xmin, xmax, ymin, ymax, zmax;
};

// ST_F32 block record (48 B, read through the scalar cache).  The fast path
// works in lattice units: q = (coordinate - gso) / D, so rint(q) is the
// sphere index and q - rint(q) the offset from the sphere centre; the voxel
// faces are the half-way planes q = -1/2 and q = N - 1/2, so a point that is
// not within dl of a half-way plane is inside the voxel iff every rint(q)
// lies in [0, N - 1].
//   oq*     (origin - gso) / D of the block's FP32 offsets (z with H' added);
//   dl      bound on |approximate - reference| of any shifted, offset
//           coordinate of the block in nm (FP32 rounding of the offsets plus,
//           with a wide margin, FP64 rounding of every step, any |shift| <= H');
//   fm      |q - rint(q)| >= fm: within dl of a half-way plane;
//   mrgq    |sum (rint(q) - q)^2 - r^2 / D^2| <= mrgq: within the error of
//           the squared distance.
struct BlockF32 {
    double oqx, oqy, oqz, dl, fm, mrgq;
};

// One FP32 point: offsets from its block's origin (nm) and the energy (eV);
// the same 16-byte layout as HIP's float4, which the kernel loads.
struct PointF32 {
    float x, y, z, w;
};

// Host staging of one appended run of tracks, indexed from the run's first
// point and block.
struct ChunkLayout {
    std::vector<double> x, y, z, e;   // z-sorted FP64 points
    std::vector<double> zfirst;       // z of each block's first point
    std::vector<PointF32> p32;        // ST_F32 only
    std::vector<BlockF32> blocks;     // ST_F32 only
};

// Fills the records of n_tracks tracks (offsets[0..n_tracks] index the
// caller's arrays) placed from library point pbase and block bbase: offset,
// n, block range, geometry (geom_of_track may be NULL: 0).  Returns the
// number of blocks the tracks take.
int64_t plan_chunk(const int64_t *offsets, int64_t n_tracks, const int32_t *geom_of_track, int64_t pbase,
                   int64_t bbase, TrackDesc *tds);

// Sorts each planned track's points by z (stable; NaN last) into out, and
// completes its record: x/y bounds, last z, and the fast-path eligibility
// (finite data, at most kMaxFastPoints points, fixed-point headroom, indices
// below 2^30 for any shift, at most 2^32 - 2 spheres in one layer of the
// track's x/y extent, at most kMaxWindow spheres in any z-window two
// diameters wide).  ST_F32 also gets the FP32 points and block records.
void fill_chunk(const double *x, const double *y, const double *z, const double *edep, const int64_t *offsets,
                int64_t n_tracks, const st_geometry *geoms, int precision, int64_t pbase, int64_t bbase,
                int64_t nblocks, TrackDesc *tds, ChunkLayout &out);

// Largest number of distinct spheres that the points of any z-window of
// width w can touch, for any shift: min(points, x-cells * y-cells * layers).
int64_t window_bound(const double *x, const double *y, const double *z, int64_t n, double D, double w,
                     int64_t layers);

}  // namespace st
