// host.cc -- the C++ classes of the reference's interface, implemented on
// the C ABI of libsupertrack_amd.  Configuration errors print a message and
// abort(), as the reference does; everything else reports through the C ABI.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <filesystem>
#include <fstream>
#include <future>
#include <iostream>
#include <sstream>
#include <thread>

#include <hip/hip_runtime_api.h>

#include "Histogram.hh"
#include "SimulationMethod.hh"
#include "SphericalGeometry.hh"
#include "SuperTrackManager.hh"
#include "ThreadAllocation.hh"
#include "Track.hh"
#include "TrackFile.hh"
#include "VolumeEdepPair.hh"
#include "VoxelConstrainedSphereMethod.hh"

namespace {

[[noreturn]] void fatal(const std::string &msg)
{
    std::cout << msg << std::endl;
    std::abort();
}

void check(int rc, const char *what)
{
    if (rc != ST_OK) fatal(std::string(what) + " failed: " + st_last_error());
}

}  // namespace

// ---------------------------------------------------------------- geometry --
SphericalGeometry::SphericalGeometry(double half_length, double diameter)
{
    st_geometry g;
    check(st_geometry_make(half_length, diameter, &g), "SphericalGeometry");
    scoringRegionLength = g.length;
    scoringRegionHalfLength = g.half_length;
    greatestSphereOffset = g.gso;
    numSpheresLinear = g.n_linear;
    sphereDiameter = g.diameter;
    sphereRadius = g.radius;
}

SphericalGeometry::SphericalGeometry(const INIReader &r)
    : SphericalGeometry(r.GetReal("VoxelConstrainedSphere", "ScoringRegionHalfLength", 0),
                        r.GetReal("VoxelConstrainedSphere", "ScoringSphereDiameter", 0))
{
}

st_geometry SphericalGeometry::c() const
{
    st_geometry g;
    g.half_length = scoringRegionHalfLength;
    g.length = scoringRegionLength;
    g.diameter = sphereDiameter;
    g.radius = sphereRadius;
    g.gso = greatestSphereOffset;
    g.n_linear = numSpheresLinear;
    g.pad = 0;
    return g;
}

// --------------------------------------------------------------- track file --
namespace TrackFile {

bool ReadHeader(const std::string &path, Header &h)
{
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    char magic[8];
    bool ok = std::fread(magic, 1, 8, f) == 8 && std::equal(magic, magic + 8, "STTRACK1") &&
              std::fread(&h.n_tracks, 8, 1, f) == 1 && std::fread(&h.n_points, 8, 1, f) == 1;
    if (ok) {
        h.index.resize(h.n_tracks);
        ok = h.n_tracks == 0 || std::fread(h.index.data(), 8, h.n_tracks, f) == h.n_tracks;
    }
    std::fclose(f);
    return ok;
}

bool ReadRange(const std::string &path, const Header &h, int64_t entry, int64_t exit, double *x, double *y,
               double *z, double *edep)
{
    if (entry < 0 || exit < entry || (uint64_t)exit > h.n_points) return false;
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    const int64_t n = exit - entry;
    const int64_t base = 8 + 16 + 8 * (int64_t)h.n_tracks;
    double *cols[4] = {x, y, z, edep};
    bool ok = true;
    for (int c = 0; c < 4 && ok; ++c) {
        const int64_t off = base + 8 * ((int64_t)c * (int64_t)h.n_points + entry);
        ok = std::fseek(f, off, SEEK_SET) == 0 && (n == 0 || std::fread(cols[c], 8, n, f) == (size_t)n);
    }
    std::fclose(f);
    return ok;
}

bool Write(const std::string &path, const std::vector<uint64_t> &index, const double *x, const double *y,
           const double *z, const double *edep, uint64_t n_points)
{
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const uint64_t nt = index.size();
    bool ok = std::fwrite("STTRACK1", 1, 8, f) == 8 && std::fwrite(&nt, 8, 1, f) == 1 &&
              std::fwrite(&n_points, 8, 1, f) == 1 && (nt == 0 || std::fwrite(index.data(), 8, nt, f) == nt);
    const double *cols[4] = {x, y, z, edep};
    for (int c = 0; c < 4 && ok; ++c) ok = n_points == 0 || std::fwrite(cols[c], 8, n_points, f) == n_points;
    return std::fclose(f) == 0 && ok;
}

}  // namespace TrackFile

// ----------------------------------------------------------- thread tasks --
void ThreadAllocation::AddTask(ThreadTask &task)
{
    const size_t ntasks = _tasks.size();
    _tasks.push_back(task);
    _tasks[ntasks]._threadID = _threadID;
    _tasks[ntasks]._nOversamples = _nOversamples;
    _tasks[ntasks]._runSeed = _randomSeed;
    std::stringstream concat;
    concat << _randomSeed << _threadID << ntasks;
    _tasks[ntasks]._randomSeed = std::stol(concat.str());
}

ThreadAllocator::ThreadAllocator(const INIReader &reader)
{
    if (reader.Get("Run", "Type", "") != "Folder")
        fatal("Macro file error: ([Run], Type) does not = folder. Only analysis runs over folders are supported.");
    _folderPath = reader.Get("Run", "Path", "");
    if (_folderPath.empty()) fatal("Macro file error: ([Run], Folder) not defined.");
    _lowerFileLimit = reader.GetInteger("Run", "FirstFile", -1);
    _upperFileLimit = reader.GetInteger("Run", "LastFile", -1);
    _numThreads = reader.GetInteger("Run", "Threads", 1);
    _nOversamples = reader.GetInteger("Run", "Oversamples", 1);
    _randomSeed = reader.GetInteger("Run", "RandomSeed", -1);
    if (_randomSeed == -1) _randomSeed = (int64_t)time(nullptr);
}

void ThreadAllocator::ParseFileLimits(int folderSize)
{
    if (_lowerFileLimit == -1 && _upperFileLimit == -1) {
        _lowerFileLimit = 1;
        _upperFileLimit = folderSize;
    } else if (_upperFileLimit < _lowerFileLimit) {
        fatal("Lower file limit greater than upper file limit. Aborting.");
    } else if (_upperFileLimit > folderSize) {
        fatal("Upper file limit greater than number of files. Aborting.");
    } else if (_lowerFileLimit < 1) {
        fatal("Lower file limit less than 1. Aborting.");
    }
}

void ThreadAllocator::ReturnThreadAllocations(std::vector<ThreadAllocation> &out)
{
    std::vector<std::string> files;
    std::error_code ec;
    for (const auto &entry : std::filesystem::directory_iterator(_folderPath, ec))
        if (entry.path().extension() == ".sttrk") files.push_back(entry.path().string());
    if (ec) fatal("Macro file error: cannot read folder " + _folderPath);
    std::sort(files.begin(), files.end());
    ParseFileLimits((int)files.size());
    std::vector<ThreadTask> tasks;
    for (long i = _lowerFileLimit; i <= _upperFileLimit; ++i) {
        TrackFile::Header h;
        if (!TrackFile::ReadHeader(files[i - 1], h)) fatal("Cannot read track file " + files[i - 1]);
        uint64_t start = 0;
        for (uint64_t t = 0; t < h.n_tracks; ++t) {
            tasks.emplace_back(files[i - 1], (int64_t)start, (int64_t)h.index[t]);
            start = h.index[t];
        }
    }
    for (size_t k = 0; k < tasks.size(); ++k) tasks[k]._trackIndex = (int64_t)k;
    if ((long)tasks.size() < _numThreads) {
        std::cout << "There are fewer tracks to analyze than the number of requested threads." << std::endl;
        std::cout << "Lowering number of active threads to: " << tasks.size() << " to accomodate." << std::endl;
        _numThreads = (long)tasks.size();
    }
    if (_numThreads < 1) return;
    const long per = (long)tasks.size() / _numThreads;
    const long rem = (long)tasks.size() % _numThreads;
    size_t cur = 0;
    for (long i = 0; i < _numThreads; ++i) {
        out.push_back(ThreadAllocation((int)i, _randomSeed, (int)_nOversamples));
        const long n = per + (i < rem ? 1 : 0);
        for (long j = 0; j < n; ++j) out.back().AddTask(tasks[cur++]);
    }
}

// -------------------------------------------------------- Track / pairs --
void Track::AllocateEmptyTrack(int nVals)
{
    check(st_track_alloc(&c, nVals), "Track::AllocateEmptyTrack");
    x = c.x;
    y = c.y;
    z = c.z;
    edep = c.edep;
}

void Track::AllocateAndLoadTrack(const ThreadTask &task)
{
    TrackFile::Header h;
    if (!TrackFile::ReadHeader(task.GetFilename(), h)) fatal("Cannot read track file " + task.GetFilename());
    const int64_t n = task.GetExitPoint() - task.GetEntryPoint();
    std::vector<double> hx(n), hy(n), hz(n), he(n);
    if (!TrackFile::ReadRange(task.GetFilename(), h, task.GetEntryPoint(), task.GetExitPoint(), hx.data(), hy.data(),
                              hz.data(), he.data()))
        fatal("Cannot read track entries from " + task.GetFilename());
    check(st_track_upload(&c, hx.data(), hy.data(), hz.data(), he.data(), n), "Track::AllocateAndLoadTrack");
    x = c.x;
    y = c.y;
    z = c.z;
    edep = c.edep;
}

void Track::LoadFromHost(const double *hx, const double *hy, const double *hz, const double *he, int64_t n,
                         void *stream)
{
    if (n > c.capacity) fatal("Track::LoadFromHost: track larger than the allocated capacity");
    const hipStream_t s = (hipStream_t)stream;
    const size_t bytes = (size_t)n * sizeof(double);
    bool ok = true;
    if (n > 0) {
        ok = ok && hipMemcpyAsync(c.x, hx, bytes, hipMemcpyHostToDevice, s) == hipSuccess;
        ok = ok && hipMemcpyAsync(c.y, hy, bytes, hipMemcpyHostToDevice, s) == hipSuccess;
        ok = ok && hipMemcpyAsync(c.z, hz, bytes, hipMemcpyHostToDevice, s) == hipSuccess;
        ok = ok && hipMemcpyAsync(c.edep, he, bytes, hipMemcpyHostToDevice, s) == hipSuccess;
    }
    if (!ok) fatal("Track::LoadFromHost: copy to the device failed");
    c.n = n;
}

void Track::LoadTrack(const ThreadTask &task, void *stream)
{
    TrackFile::Header h;
    if (!TrackFile::ReadHeader(task.GetFilename(), h)) fatal("Cannot read track file " + task.GetFilename());
    const int64_t n = task.GetExitPoint() - task.GetEntryPoint();
    if (n > c.capacity) fatal("Track::LoadTrack: track larger than the allocated capacity");
    std::vector<double> hx(n), hy(n), hz(n), he(n);
    if (!TrackFile::ReadRange(task.GetFilename(), h, task.GetEntryPoint(), task.GetExitPoint(), hx.data(), hy.data(),
                              hz.data(), he.data()))
        fatal("Cannot read track entries from " + task.GetFilename());
    const hipStream_t s = (hipStream_t)stream;
    const size_t bytes = (size_t)n * sizeof(double);
    bool ok = true;
    if (n > 0) {
        ok = ok && hipMemcpyAsync(c.x, hx.data(), bytes, hipMemcpyHostToDevice, s) == hipSuccess;
        ok = ok && hipMemcpyAsync(c.y, hy.data(), bytes, hipMemcpyHostToDevice, s) == hipSuccess;
        ok = ok && hipMemcpyAsync(c.z, hz.data(), bytes, hipMemcpyHostToDevice, s) == hipSuccess;
        ok = ok && hipMemcpyAsync(c.edep, he.data(), bytes, hipMemcpyHostToDevice, s) == hipSuccess;
    }
    ok = ok && hipStreamSynchronize(s) == hipSuccess;
    if (!ok) fatal("Track::LoadTrack: copy to the device failed");
    c.n = n;
}

void Track::Free()
{
    st_track_free(&c);
    x = y = z = edep = nullptr;
}

void VolumeEdepPair::Allocate(uint64_t n)
{
    check(st_pairs_alloc(&c, (int64_t)n), "VolumeEdepPair::Allocate");
    volume = c.volume;
    edep = c.edep;
    numElements = c.num_elements;
}

void VolumeEdepPair::Free()
{
    st_pairs_free(&c);
    volume = nullptr;
    edep = nullptr;
    numElements = nullptr;
}

// ------------------------------------------------------- method + factory --
void SimulationMethodFactory::AddSimulationMethod(const std::string &name,
                                                  SimulationMethod *(*ctor)(const INIReader &))
{
    _nameConstructorPairs.emplace_back(name, ctor);
}

SimulationMethod *SimulationMethodFactory::Construct(const INIReader &reader)
{
    const std::string name = reader.Get("Simulation", "Method", "");
    for (const auto &p : _nameConstructorPairs)
        if (p.first == name) return p.second(reader);
    fatal("No simulation method matching name: " + name + " found. Program aborting.");
}

VoxelConstrainedSphereMethod::VoxelConstrainedSphereMethod(const INIReader &reader)
    : SimulationMethod(reader), _sphericalGeometry(reader)
{
    ParseInput();
    const st_geometry g = _sphericalGeometry.c();
    check(st_vcs_create(&g, _randomShifts, _randomShiftsZ, _suggestedCudaBlocks, _suggestedCudaThreads, &_vcs),
          "VoxelConstrainedSphereMethod");
    if (_queueOversamples) check(st_vcs_set_queued(_vcs, 1), "VoxelConstrainedSphereMethod");
}

VoxelConstrainedSphereMethod::~VoxelConstrainedSphereMethod() { st_vcs_destroy(_vcs); }

void VoxelConstrainedSphereMethod::ParseInput()
{
    _suggestedCudaBlocks = (int)_macroReader.GetReal("VoxelConstrainedSphere", "SuggestedCudaBlocks", 0);
    _suggestedCudaThreads = (int)_macroReader.GetReal("VoxelConstrainedSphere", "SuggestedCudaThreads", 0);
    _randomShifts = _macroReader.GetBoolean("VoxelConstrainedSphere", "ShiftTrack", true);
    _randomShiftsZ = _macroReader.GetBoolean("VoxelConstrainedSphere", "ShiftTrackZ", false);
    _queueOversamples = _macroReader.GetBoolean("VoxelConstrainedSphere", "QueueOversamples", false);
    _referenceSeeds = _macroReader.GetBoolean("Run", "ReferenceSeeds", true);
}

SimulationMethod *VoxelConstrainedSphereMethod::Construct(const INIReader &reader)
{
    return new VoxelConstrainedSphereMethod(reader);
}

void VoxelConstrainedSphereMethod::AllocateTrackProcess(Track track, ThreadTask task)
{
    // per-task seed (ThreadAllocation.cc:29-31) with counter 0, or the run
    // seed with the task's run position as counter (what the batched form uses)
    const uint64_t seed = _referenceSeeds ? (uint64_t)task.GetRandomSeed() : (uint64_t)task.GetRunSeed();
    const uint32_t index = _referenceSeeds ? 0u : (uint32_t)task.GetTrackIndex();
    track.c.n = task.GetExitPoint() - task.GetEntryPoint();
    check(st_vcs_begin_track(_vcs, &track.c, seed, index, task.GetNOversamples()),
          "VoxelConstrainedSphereMethod::AllocateTrackProcess");
}

void VoxelConstrainedSphereMethod::ProcessTrack(Track track, VolumeEdepPair &edepsInTarget)
{
    check(st_vcs_process_oversample(_vcs, &track.c, &edepsInTarget.c), "VoxelConstrainedSphereMethod::ProcessTrack");
}

void VoxelConstrainedSphereMethod::FreeTrackProcess()
{
    check(st_vcs_end_track(_vcs), "VoxelConstrainedSphereMethod::FreeTrackProcess");
}

void VoxelConstrainedSphereMethod::Free() {}

void VoxelConstrainedSphereMethod::SetStream(void *stream)
{
    check(st_vcs_set_stream(_vcs, stream), "VoxelConstrainedSphereMethod::SetStream");
}

// -------------------------------------------------------------- Hist1D --
Hist1D::Hist1D(const std::string &name, const std::vector<double> &edges)
    : _name(name), _edges(edges), _contents(edges.size() + 1, 0.0)
{
}

double Hist1D::GetBinLowEdge(int i) const
{
    // TAxis::GetBinLowEdge: the edge array inside [1, nbins + 1]; outside it
    // the average bin width extrapolated from the lower edge
    const int nb = GetNbinsX();
    if (i >= 1 && i <= nb + 1) return _edges[i - 1];
    return _edges[0] + (double)(i - 1) * ((_edges[nb] - _edges[0]) / (double)nb);
}

namespace CPUHistogramUtils {

void PMF_to_PDF(Hist1D &h)
{
    const int length = h.GetNbinsX();
    float normalization = 0;
    for (int i = 0; i < length; ++i) {
        const double value = h.GetBinContent(i);
        const double low_edge = h.GetBinLowEdge(i);
        const double high_edge = h.GetBinLowEdge(i + 1);
        h.SetBinContent(i, value / (high_edge - low_edge));
        normalization += value;
    }
    for (int i = 0; i < length; ++i) h.SetBinContent(i, h.GetBinContent(i) / normalization);
}

void Prepare_for_Semilog(Hist1D &h)
{
    const int length = h.GetNbinsX();
    for (int i = 0; i < length; ++i) {
        const double value = h.GetBinContent(i);
        const double mid_value = (h.GetBinLowEdge(i + 1) + h.GetBinLowEdge(i)) / 2;
        h.SetBinContent(i, value * mid_value);
    }
}

bool ReadCSV(const std::string &path, Hist1D &h)
{
    std::ifstream in(path);
    if (!in) return false;
    std::string line, name = "spectrum";
    std::vector<double> lo, hi, cnt;
    while (std::getline(in, line)) {
        if (line.empty()) continue;
        if (line[0] == '#') {
            if (name == "spectrum" && line.size() > 2 && line.find(':') == std::string::npos) name = line.substr(2);
            continue;
        }
        if (line.rfind("low_edge", 0) == 0) continue;
        double a, b, c;
        if (std::sscanf(line.c_str(), "%lf,%lf,%lf", &a, &b, &c) != 3) return false;
        lo.push_back(a);
        hi.push_back(b);
        cnt.push_back(c);
    }
    if (lo.empty()) return false;
    std::vector<double> edges(lo);
    edges.push_back(hi.back());
    h = Hist1D(name, edges);
    for (size_t i = 0; i < cnt.size(); ++i) h.SetBinContent((int)i + 1, cnt[i]);
    return true;
}

}  // namespace CPUHistogramUtils

void Hist1D::Add(const Hist1D *other)
{
    if (other->_contents.size() != _contents.size()) fatal("Hist1D::Add: histograms with different binning");
    for (size_t i = 0; i < _contents.size(); ++i) _contents[i] += other->_contents[i];
}

double Hist1D::Integral() const
{
    double s = 0;
    for (int i = 1; i <= GetNbinsX(); ++i) s += _contents[i];
    return s;
}

bool Hist1D::WriteCSV(const std::string &path, const std::vector<std::pair<std::string, std::string>> &meta) const
{
    FILE *f = std::fopen(path.c_str(), "w");
    if (!f) return false;
    std::fprintf(f, "# %s\n", _name.c_str());
    for (const auto &kv : meta) std::fprintf(f, "# %s: %s\n", kv.first.c_str(), kv.second.c_str());
    std::fprintf(f, "low_edge,high_edge,count\n");
    for (int i = 1; i <= GetNbinsX(); ++i)
        std::fprintf(f, "%.17g,%.17g,%.17g\n", _edges[i - 1], _edges[i], _contents[i]);
    return std::fclose(f) == 0;
}

// ----------------------------------------------------------- Histogram --
Histogram::Histogram(int nbins, float lo, float hi, std::string type, int, int)
    : _type(std::move(type)), _nbins(nbins), _binLower(lo), _binUpper(hi)
{
    Construct();
}

Histogram::Histogram(const INIReader &r)
{
    _type = r.Get("Histogram", "Type", "");
    _binLower = r.GetFloat("Histogram", "BinLower", 0);
    _binUpper = r.GetFloat("Histogram", "BinUpper", 0);
    _nbins = (int)r.GetInteger("Histogram", "NBins", 0);
    Construct();
}

void Histogram::Construct()
{
    int type;
    if (_type == "log") type = ST_HIST_LOG;
    else if (_type == "lin") type = ST_HIST_LIN;
    else fatal("Unknown histogram type: " + _type + " available types are lin and log.");
    check(st_hist_create(type, _binLower, _binUpper, _nbins, &_hist), "Histogram");
    std::vector<double> edges(_nbins + 1);
    check(st_hist_edges(_hist, edges.data()), "Histogram edges");
    _CPU_histogram = Hist1D("Lineal energy histogram", edges);
}

Histogram::~Histogram() { st_hist_destroy(_hist); }

void Histogram::AllocateTrackProcess(VolumeEdepPair targetEdeps)
{
    check(st_hist_begin_track(_hist, &targetEdeps.c), "Histogram::AllocateTrackProcess");
}

void Histogram::SortReduceAndAddToHistogram(VolumeEdepPair targetEdeps)
{
    check(st_hist_add(_hist, &targetEdeps.c), "Histogram::SortReduceAndAddToHistogram");
}

void Histogram::SetStream(void *stream)
{
    check(st_hist_set_stream(_hist, stream), "Histogram::SetStream");
}

// AccumulateCPUHistogram (Histogram.cu:177-186) adds each track's counts to the
// CPU histogram; here they stay summed in the device counters until
// GetCPUHistogram / Print reads them, which gives the same integer totals
// without a device-to-host transfer and synchronisation per track.
void Histogram::FreeTrackProcess()
{
    check(st_hist_end_track(_hist, nullptr), "Histogram::FreeTrackProcess");
}

void Histogram::AccumulateBatched()
{
    std::vector<uint64_t> counts(_nbins);
    check(st_hist_fetch(_hist, counts.data(), 1), "Histogram::AccumulateBatched");
    for (int i = 0; i < _nbins; ++i) _CPU_histogram.AddBinContent(i + 1, (double)counts[i]);
}

void Histogram::Print()
{
    AccumulateBatched();
    double total = 0;
    for (int i = 1; i <= _nbins; ++i) {
        std::cout << "Bin: " << _CPU_histogram.GetBinLowEdge(i) << " Counts: " << _CPU_histogram.GetBinContent(i)
                  << std::endl;
        total += _CPU_histogram.GetBinContent(i);
    }
    std::cout << total << std::endl;
}

Hist1D Histogram::ReduceVector(std::vector<Hist1D> &toReduce)
{
    Hist1D out = toReduce.back();
    toReduce.pop_back();
    for (const Hist1D &h : toReduce) out.Add(&h);
    return out;
}

// ------------------------------------------------------------- manager --
void SuperTrackManager::AddSimulationMethod(const std::string &name, SimulationMethod *(*ctor)(const INIReader &))
{
    SimulationMethodFactory::GetInstance().AddSimulationMethod(name, ctor);
}

void SuperTrackManager::Initialize(INIReader *reader)
{
    if (reader->ParseError() != 0) fatal("Macro file error: cannot parse the macro file");
    delete _inputFileReader;
    _inputFileReader = reader;
    _threadAllocations.clear();
    ThreadAllocator alloc(*reader);
    alloc.ReturnThreadAllocations(_threadAllocations);
    _bInitialized = true;
}

void SuperTrackManager::Reset()
{
    _threadAllocations.clear();
    _output = Hist1D();
    _bInitialized = false;
}

void SuperTrackManager::Run()
{
    if (!_bInitialized) {
        std::cout << "Can not start run without initializing manager." << std::endl;
        return;
    }
    if (_inputFileReader->GetBoolean("VoxelConstrainedSphere", "Batched", false)) RunBatched();
    else RunPerTrack();
    EndOfRun(_output);
}

void SuperTrackManager::RunPerTrack()
{
    // One host thread per worker (the reference forks one process per
    // worker), each with its own stream, method, histogram and buffers sized
    // for its largest task; the worker histograms are summed at the end.
    const size_t nw = _threadAllocations.size();
    std::vector<Hist1D> outputs(nw);
    auto work = [&](size_t k) {
        const ThreadAllocation &alloc = _threadAllocations[k];
        const std::vector<ThreadTask> &tasks = alloc.GetTasks();
        hipStream_t stream = nullptr;
        if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess) fatal("hipStreamCreate failed");
        Histogram histogram(*_inputFileReader);
        histogram.SetStream(stream);
        SimulationMethod *method = SimulationMethodFactory::GetInstance().Construct(*_inputFileReader);
        method->SetStream(stream);
        int64_t cap = 1;
        for (const ThreadTask &task : tasks) cap = std::max<int64_t>(cap, task.GetExitPoint() - task.GetEntryPoint());
        Track track;
        track.AllocateEmptyTrack((int)cap);
        VolumeEdepPair edepsInTarget;
        edepsInTarget.Allocate((uint64_t)cap);
        // Ingest: the next task's entries are read from its file into one of
        // two pinned host buffers while the current task runs on the GPU.
        double *pinned[2] = {nullptr, nullptr};
        for (double *&p : pinned)
            if (hipHostMalloc((void **)&p, 4 * (size_t)cap * sizeof(double), hipHostMallocDefault) != hipSuccess)
                fatal("hipHostMalloc failed");
        auto read_task = [&](size_t j) {
            const ThreadTask &t = tasks[j];
            double *b = pinned[j & 1];
            TrackFile::Header h;
            if (!TrackFile::ReadHeader(t.GetFilename(), h) ||
                !TrackFile::ReadRange(t.GetFilename(), h, t.GetEntryPoint(), t.GetExitPoint(), b, b + cap,
                                      b + 2 * cap, b + 3 * cap))
                fatal("Cannot read track entries from " + t.GetFilename());
        };
        std::future<void> next;
        if (!tasks.empty()) next = std::async(std::launch::async, read_task, 0);
        for (size_t j = 0; j < tasks.size(); ++j) {
            const ThreadTask &task = tasks[j];
            next.get();
            // the buffer of task j + 1 was last read by the copy of task j - 1,
            // which the stream synchronisation at the end of that task completed
            if (j + 1 < tasks.size()) next = std::async(std::launch::async, read_task, j + 1);
            const int64_t n = task.GetExitPoint() - task.GetEntryPoint();
            const double *b = pinned[j & 1];
            track.LoadFromHost(b, b + cap, b + 2 * cap, b + 3 * cap, n, stream);
            edepsInTarget.c.capacity = std::max<int64_t>(n, 1);
            method->AllocateTrackProcess(track, task);
            histogram.AllocateTrackProcess(edepsInTarget);
            for (int os = 0; os < task.GetNOversamples(); ++os) {
                method->ProcessTrack(track, edepsInTarget);
                histogram.SortReduceAndAddToHistogram(edepsInTarget);
            }
            // cudaDeviceSynchronize of the reference's loop: FreeTrackProcess
            // synchronises the stream (all of this worker's work is on it)
            method->FreeTrackProcess();
            histogram.FreeTrackProcess();
        }
        edepsInTarget.c.capacity = cap;
        delete method;
        outputs[k] = histogram.GetCPUHistogram();
        track.Free();
        edepsInTarget.Free();
        for (double *p : pinned) (void)hipHostFree(p);
        (void)hipStreamDestroy(stream);
    };
    std::vector<std::thread> threads;
    for (size_t k = 0; k < nw; ++k) threads.emplace_back(work, k);
    for (std::thread &t : threads) t.join();
    if (outputs.empty()) {
        Histogram empty(*_inputFileReader);
        _output = empty.GetCPUHistogram();
    } else {
        _output = Histogram::ReduceVector(outputs);
    }
}

void SuperTrackManager::RunBatched()
{
    const INIReader &r = *_inputFileReader;
    if (r.Get("Simulation", "Method", "") != "VoxelConstrainedSphere")
        fatal("Batched runs are implemented for the VoxelConstrainedSphere method only.");
    std::vector<ThreadTask> tasks;
    for (const ThreadAllocation &a : _threadAllocations)
        for (const ThreadTask &t : a.GetTasks()) tasks.push_back(t);
    std::sort(tasks.begin(), tasks.end(),
              [](const ThreadTask &a, const ThreadTask &b) { return a.GetTrackIndex() < b.GetTrackIndex(); });
    std::vector<int64_t> offsets(1, 0);
    for (const ThreadTask &t : tasks) offsets.push_back(offsets.back() + (t.GetExitPoint() - t.GetEntryPoint()));
    const int64_t n = offsets.back();
    std::vector<double> x(n), y(n), z(n), e(n);
    for (size_t k = 0; k < tasks.size(); ++k) {
        TrackFile::Header h;
        const ThreadTask &t = tasks[k];
        if (!TrackFile::ReadHeader(t.GetFilename(), h) ||
            !TrackFile::ReadRange(t.GetFilename(), h, t.GetEntryPoint(), t.GetExitPoint(), x.data() + offsets[k],
                                  y.data() + offsets[k], z.data() + offsets[k], e.data() + offsets[k]))
            fatal("Cannot read track file " + t.GetFilename());
    }
    const SphericalGeometry geometry(r);
    const st_geometry g = geometry.c();
    const std::string prec = r.Get("VoxelConstrainedSphere", "Precision", "f64");
    if (prec != "f64" && prec != "f32") fatal("Unknown precision: " + prec + " available precisions are f64 and f32.");
    st_library *lib = nullptr;
    check(st_library_create(x.data(), y.data(), z.data(), e.data(), offsets.data(), (int64_t)tasks.size(), &g, 1,
                            nullptr, prec == "f32" ? ST_F32 : ST_F64, &lib),
          "st_library_create");
    Histogram histogram(r);
    st_batch_opts o{};
    o.n_oversamples = tasks.empty() ? 0 : tasks[0].GetNOversamples();
    o.shift_xy = r.GetBoolean("VoxelConstrainedSphere", "ShiftTrack", true);
    o.shift_z = r.GetBoolean("VoxelConstrainedSphere", "ShiftTrackZ", false);
    o.seed = tasks.empty() ? 0 : (uint64_t)tasks[0].GetRunSeed();
    o.track_index_base = 0;
    o.use_graph = 1;
    o.autotune = r.GetBoolean("VoxelConstrainedSphere", "Autotune", false);
    st_hist *hp = histogram.Handle();
    st_batch *b = nullptr;
    check(st_batch_create(lib, &o, &hp, 1, &b), "st_batch_create");
    check(st_batch_run(b), "st_batch_run");
    check(st_batch_sync(b), "st_batch_sync");
    histogram.AccumulateBatched();
    st_batch_destroy(b);
    st_library_destroy(lib);
    _output = histogram.GetCPUHistogram();
}

void SuperTrackManager::EndOfRun(Hist1D &output)
{
    const std::string folder = _inputFileReader->Get("Output", "Path", "");
    const std::string prefix = _inputFileReader->Get("Output", "NamePrefix", "");
    const int64_t seed = _threadAllocations.empty() ? 0 : _threadAllocations[0].GetRandomSeed();
    std::stringstream name;
    name << folder << prefix << seed << ".csv";
    double tracks = 0;
    for (const auto &a : _threadAllocations) tracks += (double)a.GetTasks().size();
    const long nos = _inputFileReader->GetInteger("Run", "Oversamples", 1);
    std::vector<std::pair<std::string, std::string>> meta = {
        {"Number of Tracks", std::to_string(tracks)},
        {"Number of Oversamples", std::to_string(nos)},
        {"Effective Number of Tracks", std::to_string(tracks * (double)nos)},
    };
    if (!folder.empty() || !prefix.empty()) {
        if (!output.WriteCSV(name.str(), meta)) std::cout << "Could not write output " << name.str() << std::endl;
    }
}
