// SphericalGeometry.hh -- the voxel and its sphere lattice
// (SphericalGeometry.cuh:8-22, SphericalGeometry.cu:34-58).
#pragma once

#include "INIReader.hh"
#include "supertrack_amd.h"

struct SphericalGeometry {
    SphericalGeometry(double scoring_region_half_length, double sphere_diameter);
    // [VoxelConstrainedSphere] ScoringRegionHalfLength, ScoringSphereDiameter
    explicit SphericalGeometry(const INIReader &reader);

    double scoringRegionLength;
    double scoringRegionHalfLength;
    double greatestSphereOffset;
    int numSpheresLinear;
    double sphereDiameter;
    double sphereRadius;

    st_geometry c() const;
};
