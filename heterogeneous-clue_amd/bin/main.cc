// clue_hip_main: runs the CLUE2D / CLUE3D HIP modules over the events of an input file, with the
// command line and final report of the reference's driver (src/serial/bin/main.cc:43-240).
//
// numberOfStreams EDM streams, each owning its own module instances, run as tasks on a pool of
// numberOfThreads threads (edm::StreamSchedule, edm::TaskPool); an external-work module's device
// work frees the thread for other streams until it completes. Streams pull events from a shared
// source until maxEvents are done, or until runForMinutes have passed at the end of a pass over the
// file (cycling over the file's events like Source2D/3D::produce, src/serial/bin/Source.cc:152-227).
// Modules are found by name in plugin libraries loaded with dlopen (src/serial/bin/PluginManager.cc,
// SharedLibrary.cc). --validation adds the CLUE2D validator (src/serial/bin/main.cc:80-85, 157-168).
#include <dlfcn.h>
#include <time.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <iterator>
#include <exception>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "DataFormats/CLUEValidatorTypes.h"
#include "DataFormats/CLUE_config.h"
#include "DataFormats/ClusterCollection.h"
#include "DataFormats/PointsCloud.h"
#include "Framework/ESPluginFactory.h"
#include "Framework/Event.h"
#include "Framework/EventSetup.h"
#include "Framework/PluginFactory.h"
#include "Framework/ProductRegistry.h"
#include "Framework/StreamSchedule.h"
#include "Framework/TaskPool.h"

namespace {

  // ------------------------------------------------------------ input (Source.cc:8-58, 88-144)
  // A file is a sequence of events: uint32 n, then n records of float32 fields
  // (2D: x, y, layer, weight; 3D: x, y, z, eta, phi, r_over_absz, radius, layer, energy, isSilicon).
  template <int NF>
  bool read_event(std::ifstream& in, std::vector<float>& rec) {
    uint32_t n = 0;
    in.read(reinterpret_cast<char*>(&n), sizeof(n));
    if (in.gcount() == 0) return false;
    if (in.gcount() != sizeof(n)) throw std::runtime_error("truncated event header");
    rec.resize(static_cast<size_t>(n) * NF);
    in.read(reinterpret_cast<char*>(rec.data()), static_cast<std::streamsize>(rec.size() * sizeof(float)));
    if (static_cast<size_t>(in.gcount()) != rec.size() * sizeof(float)) throw std::runtime_error("truncated event records");
    return true;
  }

  std::vector<PointsCloud> read2D(std::filesystem::path const& f) {
    std::ifstream in(f, std::ios::binary);
    if (!in) throw std::runtime_error("cannot open " + f.string());
    std::vector<PointsCloud> evs;
    std::vector<float> rec;
    while (read_event<4>(in, rec)) {
      PointsCloud pc;
      const size_t n = rec.size() / 4;
      for (size_t i = 0; i < n; ++i) {
        pc.x.push_back(rec[4 * i]);
        pc.y.push_back(rec[4 * i + 1]);
        pc.layer.push_back(static_cast<int>(rec[4 * i + 2]));
        pc.weight.push_back(rec[4 * i + 3]);
      }
      evs.push_back(std::move(pc));
    }
    return evs;
  }

  // A toyDetector file is one event with no header: records {x, y, layer, weight} (f32), repeated on
  // each of the 100 layers with layer + l (Source.cc:60-79); a trailing partial record is ignored.
  std::vector<PointsCloud> readToyDetector(std::filesystem::path const& f) {
    std::ifstream in(f, std::ios::binary);
    if (!in) throw std::runtime_error("cannot open " + f.string());
    const std::vector<char> bytes{std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>()};
    const size_t n = bytes.size() / (4 * sizeof(float));
    std::vector<float> rec(n * 4);
    std::memcpy(rec.data(), bytes.data(), rec.size() * sizeof(float));
    PointsCloud pc;
    constexpr int kLayers = 100;
    for (int l = 0; l < kLayers; ++l)
      for (size_t i = 0; i < n; ++i) {
        pc.x.push_back(rec[4 * i]);
        pc.y.push_back(rec[4 * i + 1]);
        pc.layer.push_back(static_cast<int>(rec[4 * i + 2] + static_cast<float>(l)));
        pc.weight.push_back(rec[4 * i + 3]);
      }
    return {std::move(pc)};
  }

  std::vector<ClusterCollection> read3D(std::filesystem::path const& f) {
    std::ifstream in(f, std::ios::binary);
    if (!in) throw std::runtime_error("cannot open " + f.string());
    std::vector<ClusterCollection> evs;
    std::vector<float> r;
    while (read_event<10>(in, r)) {
      ClusterCollection c;
      const size_t n = r.size() / 10;
      for (size_t i = 0; i < n; ++i) {
        const float* p = &r[10 * i];
        c.x.push_back(p[0]);
        c.y.push_back(p[1]);
        c.z.push_back(p[2]);
        c.eta.push_back(p[3]);
        c.phi.push_back(p[4]);
        c.r_over_absz.push_back(p[5]);
        c.radius.push_back(p[6]);
        c.layer.push_back(static_cast<int>(p[7]));
        c.energy.push_back(p[8]);
        c.isSilicon.push_back(static_cast<int>(p[9]));
      }
      evs.push_back(std::move(c));
    }
    return evs;
  }

  // ------------------------------------------------------------ event numbering (Source.cc:152-190)
  // Hands out 0-based event numbers: the first maxEvents, or, with runForMinutes >= 0, whole passes
  // over the file's events until that many minutes have passed when a pass ends.
  class EventCounter {
  public:
    EventCounter(int maxEvents, int runForMinutes, int eventsInFile)
        : maxEvents_(maxEvents), minutes_(runForMinutes), cycle_(eventsInFile), start_(std::chrono::steady_clock::now()) {}

    int take() {
      if (stop_.load()) return -1;
      const int k = next_.fetch_add(1);
      if (minutes_ < 0) {
        if (k < maxEvents_) return k;
        stop_.store(true);
        return -1;
      }
      if (k - lastCheck_.load() >= cycle_) {
        std::lock_guard<std::mutex> g(m_);
        if (k - lastCheck_.load() >= cycle_) {
          const auto elapsed = std::chrono::steady_clock::now() - start_;
          if (std::chrono::duration_cast<std::chrono::minutes>(elapsed).count() >= minutes_) stop_.store(true);
          lastCheck_.store((k / cycle_) * cycle_);
        }
      }
      return stop_.load() ? -1 : k;
    }

    void stop() { stop_.store(true); }

  private:
    const int maxEvents_, minutes_, cycle_;
    const std::chrono::steady_clock::time_point start_;
    std::atomic<int> next_{0}, lastCheck_{0};
    std::atomic<bool> stop_{false};
    std::mutex m_;
  };

  // ------------------------------------------------------------ running the streams
  // numberOfStreams EDM streams on a pool of numberOfThreads threads (src/serial/bin/main.cc:188-197,
  // EventProcessor.cc:26-45); the first exception stops every stream and is rethrown here.
  template <typename Input>
  int process(std::vector<Input> const& events, std::vector<std::string> const& modules, edm::EventSetup const& es,
              int numberOfThreads, int numberOfStreams, int maxEvents, int runForMinutes, double& seconds) {
    std::vector<std::unique_ptr<edm::StreamSchedule<Input>>> streams;
    for (int s = 0; s < numberOfStreams; ++s)
      streams.push_back(std::make_unique<edm::StreamSchedule<Input>>(s, modules));
    EventCounter counter(maxEvents, runForMinutes, static_cast<int>(events.size()));
    std::atomic<int> processed{0};
    std::exception_ptr err;
    std::mutex m;
    std::condition_variable cv;
    int running = numberOfStreams;
    const auto t0 = std::chrono::steady_clock::now();
    {
      edm::TaskPool pool(numberOfThreads);
      for (auto& s : streams)
        s->start(pool, events, es, [&counter] { return counter.take(); }, processed, [&](std::exception_ptr e) {
          std::lock_guard<std::mutex> g(m);
          if (e) {
            if (!err) err = e;
            counter.stop();  // the other streams stop at their next event
          }
          if (--running == 0) cv.notify_one();
        });
      std::unique_lock<std::mutex> lk(m);
      cv.wait(lk, [&] { return running == 0; });
    }
    seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (err) std::rethrow_exception(err);
    for (auto& s : streams) s->endJob();
    return processed.load();
  }

  double cpu_seconds() {
    timespec ts;
    clock_gettime(CLOCK_PROCESS_CPUTIME_ID, &ts);
    return static_cast<double>(ts.tv_sec) + 1e-9 * static_cast<double>(ts.tv_nsec);
  }

  void print_help(std::string const& name) {
    std::cout
        << name << ": CLUE / CLUE3D on the HIP backend\n\n"
        << "Usage: " << name
        << " [--dim 2|3] [--numberOfThreads NT] [--numberOfStreams NS] [--maxEvents ME | --runForMinutes MIN]\n"
        << "       [--inputFile PATH] [--configFile PATH] [--validation] [--strictValidation] [--outputDir DIR]\n"
        << "       [--dumpDir DIR] [--plugin LIB] [--empty]\n\n"
        << "  --dim               2 (CLUE, default) or 3 (CLUE3D)\n"
        << "  --numberOfThreads   threads running the EDM streams' modules (default 1; 0: one per core)\n"
        << "  --numberOfStreams   concurrent events (default: number of threads)\n"
        << "  --maxEvents         events to process (default: all events of the input file; 10 for a\n"
        << "                      toyDetector file)\n"
        << "  --runForMinutes     process passes over the input file until this many minutes have passed\n"
        << "  --inputFile         raw input (2D: uint32 n + n x {x,y,layer,weight} f32 per event;\n"
        << "                      3D: n x {x,y,z,eta,phi,r_over_absz,radius,layer,energy,isSilicon});\n"
        << "                      a 2D file named toyDetector* holds one layer's {x,y,layer,weight} records\n"
        << "                      with no header, repeated on all 100 layers (default: data/input/raw2D.bin\n"
        << "                      or raw3D.bin next to this program)\n"
        << "  --configFile        2D: CSV dc,rhoc,outlierDeltaFactor,produceOutput (default:\n"
        << "                      config/hgcal_config.csv next to this program); 3D: name,value lines\n"
        << "  --validation        check each event against <input dir>/../output/reference/ref_<N>k_20_25_2.csv\n"
        << "                      for a toyDetector_<N>k input (CLUE 2D only); after --inputFile it selects\n"
        << "                      config/toyDetector_config.csv next to this program\n"
        << "  --strictValidation  --validation, and every result row must print exactly as in the file\n"
        << "  --outputDir         where CLUEOutputProducer writes its CSV files (2D, produceOutput = 1)\n"
        << "  --dumpDir           write bit-exact binary outputs per event (Event_<id>.bin)\n"
        << "  --plugin            plugin library to load (default: libclue_plugins.so next to this program)\n"
        << "  --empty             run no algorithm modules (framework overhead only)\n";
  }

}  // namespace

int main(int argc, char** argv) {
  std::vector<std::string> args(argv, argv + argc);
  int dim = 2, numberOfThreads = 1, numberOfStreams = 0, maxEvents = -1, runForMinutes = -1;
  std::filesystem::path inputFile, configFile, outputDir, dumpDir;
  std::vector<std::string> plugins;
  bool empty = false, validation = false, strictValidation = false;
  const std::filesystem::path exeDir = std::filesystem::path(args[0]).parent_path();
  try {
    for (size_t i = 1; i < args.size(); ++i) {
      auto next = [&]() -> std::string const& {
        if (i + 1 >= args.size()) throw std::invalid_argument("missing value after " + args[i]);
        return args[++i];
      };
      if (args[i] == "-h" || args[i] == "--help") {
        print_help(args[0]);
        return EXIT_SUCCESS;
      } else if (args[i] == "--dim") {
        dim = std::stoi(next());
      } else if (args[i] == "--numberOfThreads") {
        numberOfThreads = std::stoi(next());
      } else if (args[i] == "--numberOfStreams") {
        numberOfStreams = std::stoi(next());
      } else if (args[i] == "--maxEvents") {
        maxEvents = std::stoi(next());
      } else if (args[i] == "--runForMinutes") {
        runForMinutes = std::stoi(next());
      } else if (args[i] == "--inputFile") {
        inputFile = next();
      } else if (args[i] == "--configFile") {
        configFile = next();
      } else if (args[i] == "--validation" || args[i] == "--strictValidation") {
        validation = true;
        strictValidation = strictValidation || args[i] == "--strictValidation";
        if (inputFile.string().find("toyDetector") != std::string::npos)
          configFile = exeDir / "config" / "toyDetector_config.csv";
      } else if (args[i] == "--outputDir") {
        outputDir = next();
      } else if (args[i] == "--dumpDir") {
        dumpDir = next();
      } else if (args[i] == "--plugin") {
        plugins.push_back(next());
      } else if (args[i] == "--empty") {
        empty = true;
      } else {
        std::cout << "Invalid parameter " << args[i] << std::endl << std::endl;
        print_help(args[0]);
        return EXIT_FAILURE;
      }
    }
  } catch (std::exception const& e) {
    std::cout << "Invalid command line: " << e.what() << std::endl;
    return EXIT_FAILURE;
  }
  if (dim != 2 && dim != 3) {
    std::cout << "The dimensionality of the algorithm is either 2 or 3!" << std::endl;
    return EXIT_FAILURE;
  }
  if (maxEvents >= 0 && runForMinutes >= 0) {
    std::cout << "Got both --maxEvents and --runForMinutes, please give only one of them" << std::endl;
    return EXIT_FAILURE;
  }
  if (numberOfThreads <= 0) numberOfThreads = static_cast<int>(std::max(1u, std::thread::hardware_concurrency()));
  if (numberOfStreams <= 0) numberOfStreams = numberOfThreads;
  if (inputFile.empty()) inputFile = exeDir / "data" / "input" / (dim == 2 ? "raw2D.bin" : "raw3D.bin");
  if (configFile.empty() && dim == 2) configFile = exeDir / "config" / "hgcal_config.csv";
  if (!std::filesystem::exists(inputFile)) {
    std::cout << "Input file '" << inputFile.string() << "' does not exist" << std::endl;
    return EXIT_FAILURE;
  }
  if (dim == 2 && !std::filesystem::exists(configFile)) {
    std::cout << "Config file '" << configFile.string() << "' does not exist" << std::endl;
    return EXIT_FAILURE;
  }
  if (plugins.empty())
    plugins.push_back((std::filesystem::canonical("/proc/self/exe").parent_path() / "libclue_plugins.so").string());
  for (auto const& p : plugins)
    if (!dlopen(p.c_str(), RTLD_NOW | RTLD_GLOBAL)) {
      std::cout << "Cannot load plugin library " << p << ": " << dlerror() << std::endl;
      return EXIT_FAILURE;
    }

  try {
    edm::EventSetup es;
    std::vector<std::string> modules;
    double seconds = 0;
    int processed = 0;
    const double cpu0 = cpu_seconds();
    auto announce = [&](int events) {
      if (runForMinutes < 0)
        std::cout << "Processing " << events << " events, of which " << numberOfStreams << " concurrently, with "
                  << numberOfThreads << " threads." << std::endl;
      else
        std::cout << "Processing for about " << runForMinutes << " minutes with " << numberOfStreams
                  << " concurrent events and " << numberOfThreads << " threads." << std::endl;
    };
    if (dim == 2) {
      edm::ESPluginFactory::create("CLUEHIPClusterizerESProducer", configFile)->produce(es);
      Parameters const& par = es.get<Parameters>();
      std::cerr << "Running CLUE 2D algorithm with the following parameters: \n"
                << "dc = " << par.dc << '\n'
                << "rhoc = " << par.rhoc << '\n'
                << "outlierDeltaFactor = " << par.outlierDeltaFactor << std::endl;
      if (par.produceOutput) std::cerr << "Producing output at the end" << std::endl;
      if (!empty) {
        modules.push_back("CLUEHIPClusterizer");
        if (par.produceOutput) {
          edm::ESPluginFactory::create("CLUEOutputESProducer", outputDir.empty() ? "output" : outputDir)->produce(es);
          modules.push_back("CLUEOutputProducer");
        }
        if (validation) {
          edm::ESPluginFactory::create("CLUEValidatorESProducer", inputFile)->produce(es);
          es.put(std::make_unique<ValidationMode>(ValidationMode{strictValidation}));
          modules.push_back("CLUEValidator");
        }
        if (!dumpDir.empty()) {
          edm::ESPluginFactory::create("CLUEDumpESProducer", dumpDir)->produce(es);
          modules.push_back("CLUEBinaryDump");
        }
      }
      const bool toy = inputFile.string().find("toyDetector") != std::string::npos;
      auto events = toy ? readToyDetector(inputFile) : read2D(inputFile);
      if (events.empty()) throw std::runtime_error("input file holds no events");
      if (maxEvents < 0 && runForMinutes < 0) maxEvents = toy ? 10 : static_cast<int>(events.size());
      announce(maxEvents);
      processed = process(events, modules, es, numberOfThreads, numberOfStreams, maxEvents, runForMinutes, seconds);
    } else {
      if (!configFile.empty()) {
        edm::ESPluginFactory::create("CLUEHIPTracksterizerESProducer", configFile)->produce(es);
        std::cerr << "Running CLUE 3D algorithm with the parameters of " << configFile.string() << "\n";
      } else {
        std::cerr << "Running CLUE 3D algorithm with default parameters\n";
      }
      if (!empty) {
        modules.push_back("CLUEHIPTracksterizer");
        if (validation) std::cerr << "Validation not available for CLUE 3D" << std::endl;
        if (!dumpDir.empty()) {
          edm::ESPluginFactory::create("CLUEDumpESProducer", dumpDir)->produce(es);
          modules.push_back("CLUE3DBinaryDump");
        }
      }
      auto events = read3D(inputFile);
      if (events.empty()) throw std::runtime_error("input file holds no events");
      if (maxEvents < 0 && runForMinutes < 0) maxEvents = static_cast<int>(events.size());
      announce(maxEvents);
      processed = process(events, modules, es, numberOfThreads, numberOfStreams, maxEvents, runForMinutes, seconds);
    }
    const double cpu = cpu_seconds() - cpu0;
    std::cout << "Processed " << processed << " events in " << std::scientific << seconds << " seconds, throughput "
              << std::defaultfloat << (processed / seconds) << " events/s, CPU usage per thread: " << std::fixed
              << std::setprecision(1) << (cpu / seconds / numberOfThreads * 100) << "%" << std::endl;
  } catch (std::runtime_error const& e) {
    std::cout << "\n----------\nCaught std::runtime_error" << std::endl;
    std::cout << e.what() << std::endl;
    return EXIT_FAILURE;
  } catch (std::exception const& e) {
    std::cout << "\n----------\nCaught std::exception" << std::endl;
    std::cout << e.what() << std::endl;
    return EXIT_FAILURE;
  }
  return EXIT_SUCCESS;
}
