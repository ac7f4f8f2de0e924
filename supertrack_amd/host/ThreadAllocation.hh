// ThreadAllocation.hh -- tasks (one track each) and their split over workers
// (ThreadTask.hh:9-16, ThreadAllocation.cc:16-32, ThreadAllocator.cc:29-157).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "INIReader.hh"

class ThreadAllocation;

class ThreadTask {
public:
    ThreadTask(const std::string &fileName, int64_t entryPoint, int64_t exitPoint)
        : _fileName(fileName), _entryPoint(entryPoint), _exitPoint(exitPoint) {}
    std::string GetFilename() const { return _fileName; }
    int64_t GetEntryPoint() const { return _entryPoint; }
    int64_t GetExitPoint() const { return _exitPoint; }
    int64_t GetRandomSeed() const { return _randomSeed; }
    int GetThreadID() const { return _threadID; }
    int GetNOversamples() const { return _nOversamples; }
    // Position of the task in the run (all files, all tracks), used as the
    // track index of the shift stream in run-seeded mode.
    int64_t GetTrackIndex() const { return _trackIndex; }
    // The run's [Run] RandomSeed (after the -1 -> time substitution).
    int64_t GetRunSeed() const { return _runSeed; }

private:
    friend class ThreadAllocation;
    friend class ThreadAllocator;
    std::string _fileName;
    int64_t _entryPoint, _exitPoint;
    int _threadID = 0;
    int64_t _randomSeed = 0;
    int _nOversamples = 1;
    int64_t _trackIndex = 0;
    int64_t _runSeed = 0;
};

class ThreadAllocation {
public:
    ThreadAllocation(int threadID, int64_t randomSeed, int nOversamples)
        : _threadID(threadID), _randomSeed(randomSeed), _nOversamples(nOversamples) {}
    // The task's seed is the decimal concatenation of the run seed, the
    // worker id and the task's position in the worker (ThreadAllocation.cc:29-31).
    void AddTask(ThreadTask &task);
    std::vector<ThreadTask> GetTasks() const { return _tasks; }
    int64_t GetRandomSeed() const { return _randomSeed; }
    int GetNOversamples() const { return _nOversamples; }
    int GetThreadID() const { return _threadID; }

private:
    int _threadID;
    int64_t _randomSeed;
    int _nOversamples;
    std::vector<ThreadTask> _tasks;
};

class ThreadAllocator {
public:
    // [Run] Type = Folder, Path, FirstFile, LastFile, Threads, Oversamples,
    // RandomSeed (-1: the current time).  Track files are *.sttrk (see
    // TrackFile.hh) instead of ROOT files.
    explicit ThreadAllocator(const INIReader &reader);
    void ReturnThreadAllocations(std::vector<ThreadAllocation> &out);
    int64_t GetRandomSeed() const { return _randomSeed; }

private:
    void ParseFileLimits(int folderSize);
    std::string _folderPath;
    long _lowerFileLimit, _upperFileLimit, _numThreads, _nOversamples;
    int64_t _randomSeed;
};
