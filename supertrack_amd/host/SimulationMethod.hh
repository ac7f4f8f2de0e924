// SimulationMethod.hh -- the plugin interface of a superposition method and
// its name -> constructor registry (SimulationMethod.hh:16-30,
// SimulationMethodFactory.hh:10-35).
#pragma once

#include <string>
#include <utility>
#include <vector>

#include "INIReader.hh"
#include "ThreadAllocation.hh"

class Track;
class VolumeEdepPair;

class SimulationMethod {
public:
    explicit SimulationMethod(const INIReader &macroReader) : _macroReader(macroReader) {}
    virtual ~SimulationMethod() = default;

    virtual void ParseInput() = 0;
    virtual void AllocateTrackProcess(Track track, ThreadTask task) = 0;
    virtual void ProcessTrack(Track track, VolumeEdepPair &edepsInTarget) = 0;
    virtual void FreeTrackProcess() = 0;
    virtual void Free() = 0;
    // Runs the method's device work on a given hipStream_t, so that workers
    // can run concurrently (the reference's workers are separate processes).
    virtual void SetStream(void *stream) { (void)stream; }

protected:
    INIReader _macroReader;
};

class SimulationMethodFactory {
public:
    static SimulationMethodFactory &GetInstance()
    {
        static SimulationMethodFactory singleton;
        return singleton;
    }
    SimulationMethodFactory(const SimulationMethodFactory &) = delete;
    void operator=(const SimulationMethodFactory &) = delete;

    void AddSimulationMethod(const std::string &name, SimulationMethod *(*constructorPointer)(const INIReader &));
    // Constructs the method named by [Simulation] Method; prints and aborts
    // when no registered name matches.
    SimulationMethod *Construct(const INIReader &reader);

private:
    SimulationMethodFactory() = default;
    std::vector<std::pair<std::string, SimulationMethod *(*)(const INIReader &)>> _nameConstructorPairs;
};
