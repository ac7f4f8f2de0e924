// Track.hh -- device SoA of one track (Track.cuh:7-23).
#pragma once

#include "supertrack_amd.h"

class ThreadTask;

class Track {
public:
    Track() = default;
    void Free();
    // Reads the task's entries from its track file and copies them to the GPU.
    void AllocateAndLoadTrack(const ThreadTask &task);
    void AllocateEmptyTrack(int nVals);
    // Reads the task's entries into this track's buffers (capacity must
    // hold them) by copies on the given hipStream_t; returns when done.
    void LoadTrack(const ThreadTask &task, void *stream);
    // Queues copies of n entries from host arrays (pinned for an asynchronous
    // copy) on the given hipStream_t; the arrays must stay valid until the
    // stream has run them.
    void LoadFromHost(const double *x, const double *y, const double *z, const double *edep, int64_t n,
                      void *stream);

    double *x = nullptr;
    double *y = nullptr;
    double *z = nullptr;
    double *edep = nullptr;

    st_track c{};   // the C ABI view of the same buffers
};
