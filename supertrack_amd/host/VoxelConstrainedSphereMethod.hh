// VoxelConstrainedSphereMethod.hh -- spheres on a cubic lattice filling a
// voxel, tracks shifted at random (VoxelConstrainedSphereMethod.hh:13-49).
// Keys: [VoxelConstrainedSphere] ScoringRegionHalfLength, ScoringSphereDiameter,
// ShiftTrack (default true), SuggestedCudaBlocks, SuggestedCudaThreads, and
// the additive ShiftTrackZ (default false: z is never shifted, as in the
// reference) and QueueOversamples (default false: each ProcessTrack call
// launches its work; true: a track's oversamples run as one launch when the
// track ends, st_vcs_set_queued).
#pragma once

#include "SimulationMethod.hh"
#include "SphericalGeometry.hh"
#include "supertrack_amd.h"

class VoxelConstrainedSphereMethod : public SimulationMethod {
public:
    explicit VoxelConstrainedSphereMethod(const INIReader &macroReader);
    ~VoxelConstrainedSphereMethod() override;

    void ParseInput() override;
    void AllocateTrackProcess(Track track, ThreadTask task) override;
    void ProcessTrack(Track track, VolumeEdepPair &edepsInTarget) override;
    void FreeTrackProcess() override;
    void Free() override;
    void SetStream(void *stream) override;

    static SimulationMethod *Construct(const INIReader &macroReader);

    const SphericalGeometry &Geometry() const { return _sphericalGeometry; }
    bool ShiftTrack() const { return _randomShifts; }
    bool ShiftTrackZ() const { return _randomShiftsZ; }

private:
    int _suggestedCudaThreads = 0;
    int _suggestedCudaBlocks = 0;
    bool _randomShifts = true;
    bool _randomShiftsZ = false;
    bool _queueOversamples = false;
    bool _referenceSeeds = true;
    SphericalGeometry _sphericalGeometry;
    st_vcs *_vcs = nullptr;
};
