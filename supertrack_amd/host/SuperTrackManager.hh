// SuperTrackManager.hh -- run driver: tasks from the macro's folder, the
// per-track loop of SuperTrackManager.cu:58-83 and the output file
// (SuperTrackManager.hh:14-37).  One process, one GPU: the reference's forked
// workers become host threads, each with its own HIP stream, whose
// histograms are summed (Histogram::ReduceVector).  With [VoxelConstrainedSphere] Batched = true the
// whole task list is scored by one batched pass instead.
#pragma once

#include <string>
#include <vector>

#include "Histogram.hh"
#include "INIReader.hh"
#include "SimulationMethod.hh"
#include "ThreadAllocation.hh"

class SuperTrackManager {
public:
    static SuperTrackManager &GetInstance()
    {
        static SuperTrackManager singleton;
        return singleton;
    }
    ~SuperTrackManager() { delete _inputFileReader; }
    SuperTrackManager(const SuperTrackManager &) = delete;
    void operator=(const SuperTrackManager &) = delete;

    void AddSimulationMethod(const std::string &name, SimulationMethod *(*constructorPointer)(const INIReader &));
    void Initialize(INIReader *reader);
    void Run();
    Hist1D &GetOutput() { return _output; }
    // Clears the run state so that another macro can be run in this process.
    void Reset();

private:
    SuperTrackManager() = default;
    void RunPerTrack();
    void RunBatched();
    void EndOfRun(Hist1D &output);

    INIReader *_inputFileReader = nullptr;
    Hist1D _output;
    std::vector<ThreadAllocation> _threadAllocations;
    bool _bInitialized = false;
};
