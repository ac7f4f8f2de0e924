// TrackFile.hh -- ROOT-free track library file (*.sttrk).
//
// The reference stores tracks in ROOT files: a "Tracks" tree with branches
// "x [nm]", "y [nm]", "z [nm]", "edep [eV]" and a "Track index" tree of
// cumulative end entries (Track.cu:25-48, ThreadAllocator.cc:135-157).  This
// file holds the same content:
//   char     magic[8] = "STTRACK1"
//   uint64   n_tracks, n_points
//   uint64   index[n_tracks]      cumulative end entry of each track
//   double   x[n_points], y[n_points], z[n_points], edep[n_points]
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace TrackFile {

struct Header {
    uint64_t n_tracks = 0, n_points = 0;
    std::vector<uint64_t> index;
};

bool ReadHeader(const std::string &path, Header &h);
// Reads entries [entry, exit) of the four columns.
bool ReadRange(const std::string &path, const Header &h, int64_t entry, int64_t exit, double *x, double *y,
               double *z, double *edep);
bool Write(const std::string &path, const std::vector<uint64_t> &index, const double *x, const double *y,
           const double *z, const double *edep, uint64_t n_points);

}  // namespace TrackFile
