// VolumeEdepPair.hh -- device (sphere key, lineal energy) pairs
// (VolumeEdepPair.cuh:5-16).
#pragma once

#include <cstdint>

#include "supertrack_amd.h"

class VolumeEdepPair {
public:
    VolumeEdepPair() = default;
    void Allocate(uint64_t numInputElements);
    void Free();

    uint64_t *volume = nullptr;
    double *edep = nullptr;
    int *numElements = nullptr;   // pair count: pinned host memory, read after the track's sync

    st_pairs c{};
};
