// Histogram.hh -- lineal-energy histogram filled on the GPU
// (Histogram.cuh:19-78) and Hist1D, the ROOT-free stand-in for the TH1D it
// hands back (same bin numbering: 0 underflow, 1..nbins, nbins+1 overflow).
#pragma once

#include <string>
#include <vector>

#include "INIReader.hh"
#include "VolumeEdepPair.hh"
#include "supertrack_amd.h"

class Hist1D {
public:
    Hist1D() = default;
    Hist1D(const std::string &name, const std::vector<double> &edges);
    int GetNbinsX() const { return (int)_edges.size() - 1; }
    double GetBinContent(int i) const { return _contents.at(i); }
    double GetBinLowEdge(int i) const;
    void AddBinContent(int i, double w) { _contents.at(i) += w; }
    void SetBinContent(int i, double w) { _contents.at(i) = w; }
    void Add(const Hist1D *other);
    double Integral() const;
    const std::vector<double> &Edges() const { return _edges; }
    const std::vector<double> &Contents() const { return _contents; }
    // "low_edge,high_edge,count" per bin, preceded by "# key: value" metadata lines.
    bool WriteCSV(const std::string &path, const std::vector<std::pair<std::string, std::string>> &meta) const;

private:
    std::string _name;
    std::vector<double> _edges;
    std::vector<double> _contents;
};

// Spectrum post-processing of a finished histogram (utils.cc:80-117).  Both
// loops run over TH1 bins 0..nbins-1, as the reference's do: the underflow bin
// is included and the last bin is left as it is.
namespace CPUHistogramUtils {
// Counts per bin width, divided by the (single-precision) sum of the counts.
void PMF_to_PDF(Hist1D &h);
// Density times the bin's mid value: y f(y) for a semilog plot.
void Prepare_for_Semilog(Hist1D &h);
// Reads a histogram written by Hist1D::WriteCSV (under- and overflow zero).
bool ReadCSV(const std::string &path, Hist1D &h);
}  // namespace CPUHistogramUtils

class Histogram {
public:
    Histogram(int nbins, float binLower, float binUpper, std::string type, int suggestedCudaAccumulateBlocks = 4,
              int suggestedCudaAccumulatedThreads = 32);
    // [Histogram] Type (lin/log), NBins, BinLower, BinUpper (read as floats),
    // SuggestedCudaAccumulateBlocks/Threads (hints)
    explicit Histogram(const INIReader &reader);
    ~Histogram();
    Histogram(const Histogram &) = delete;
    void operator=(const Histogram &) = delete;

    void AllocateTrackProcess(VolumeEdepPair targetEdeps);
    void FreeTrackProcess();
    void SortReduceAndAddToHistogram(VolumeEdepPair targetEdeps);
    // device work of this histogram on a given hipStream_t (one per worker)
    void SetStream(void *stream);
    void Print();

    // The CPU histogram, after adding the counts still held on the device
    // (FreeTrackProcess leaves each track's counts there: one transfer at the
    // end instead of one per track; the sums are exact integers either way).
    Hist1D GetCPUHistogram()
    {
        AccumulateBatched();
        return _CPU_histogram;
    }
    static Hist1D ReduceVector(std::vector<Hist1D> &toReduce);

    st_hist *Handle() { return _hist; }
    // Adds the device counters (batched runs, tracks ended by FreeTrackProcess)
    // to the CPU copy and clears them.
    void AccumulateBatched();

private:
    void Construct();
    std::string _type;
    int _nbins = 0;
    float _binLower = 0, _binUpper = 0;
    st_hist *_hist = nullptr;
    Hist1D _CPU_histogram;
};
