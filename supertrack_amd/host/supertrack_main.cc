// supertrack_main.cc -- command line entry (SuperTrack.cu:9-50):
//   supertrack [macro.ini]   (default ../macros/default.ini)
//   supertrack --spectrum in.csv out.csv [semilog]
#include <chrono>
#include <ctime>
#include <iostream>
#include <string>

#include "SuperTrackManager.hh"
#include "VoxelConstrainedSphereMethod.hh"

// --ini-dump <file> <section> <name>: prints what each INIReader getter
// returns for one key (used to test the macro parsing rules).
static int ini_dump(const std::string &file, const std::string &section, const std::string &name)
{
    INIReader r(file);
    std::cout << "error=" << r.ParseError() << "\n";
    std::string v = r.Get(section, name, "<default>");
    std::string esc;
    for (char c : v) esc += (c == '\n') ? std::string("\\n") : std::string(1, c);
    std::cout << "get=[" << esc << "]\n";
    std::cout << "integer=" << r.GetInteger(section, name, -999) << "\n";
    std::cout.precision(17);
    std::cout << "real=" << r.GetReal(section, name, -999.0) << "\n";
    std::cout << "float=" << (double)r.GetFloat(section, name, -999.0f) << "\n";
    std::cout << "bool=" << (r.GetBoolean(section, name, true) ? 1 : 0) << (r.GetBoolean(section, name, false) ? 1 : 0)
              << "\n";
    return 0;
}

// --spectrum <in.csv> <out.csv> [semilog]: PMF -> PDF (and y f(y)) of a
// histogram written by a run (CPUHistogramUtils, utils.cc:80-117); no GPU.
static int spectrum(const std::string &in, const std::string &out, bool semilog)
{
    Hist1D h;
    if (!CPUHistogramUtils::ReadCSV(in, h)) {
        std::cerr << "Cannot read histogram " << in << std::endl;
        return 1;
    }
    CPUHistogramUtils::PMF_to_PDF(h);
    if (semilog) CPUHistogramUtils::Prepare_for_Semilog(h);
    if (!h.WriteCSV(out, {{"spectrum", semilog ? "y f(y)" : "f(y)"}})) {
        std::cerr << "Cannot write " << out << std::endl;
        return 1;
    }
    return 0;
}

int main(int argc, char **argv)
{
    if (argc == 5 && std::string(argv[1]) == "--ini-dump") return ini_dump(argv[2], argv[3], argv[4]);
    if ((argc == 4 || argc == 5) && std::string(argv[1]) == "--spectrum")
        return spectrum(argv[2], argv[3], argc == 5 && std::string(argv[4]) == "semilog");
    time_t start = time(nullptr);
    std::string path = "../macros/default.ini";
    if (argc == 2) path = argv[1];
    SuperTrackManager &manager = SuperTrackManager::GetInstance();
    manager.AddSimulationMethod("VoxelConstrainedSphere", &VoxelConstrainedSphereMethod::Construct);
    manager.Initialize(new INIReader(path));
    // the HIP runtime and the device are set up before the run clock starts:
    // "Run time" covers the run, not the process start
    int devices = 0;
    if (st_device_count(&devices) != ST_OK || devices < 1) {
        std::cerr << "No GPU: " << st_last_error() << std::endl;
        return 1;
    }
    const auto t0 = std::chrono::steady_clock::now();
    manager.Run();
    const auto t1 = std::chrono::steady_clock::now();
    const Hist1D &out = manager.GetOutput();
    std::cout << "Scored spheres: " << out.Integral() << std::endl;
    std::cout << "Run time: " << std::chrono::duration<double, std::milli>(t1 - t0).count() << " ms" << std::endl;
    std::cout << "SuperTrack concluding after: " << (time(nullptr) - start) << " seconds." << std::endl;
    // every device object of the run is gone: release the library's streams
    // before the HIP runtime is torn down at exit
    if (st_shutdown() != ST_OK) {
        std::cerr << "st_shutdown: " << st_last_error() << std::endl;
        return 1;
    }
    return 0;
}
