// Test harness: runs the modules downstream of the CLUE2D clusterizer (CLUEValidator,
// CLUEOutputProducer) from the plugin library on a PointsCloudSerial read from a file, on the CPU.
// The EventSetup holds what the driver puts for a CLUE2D job: Parameters, the validator's
// OutputDirPath (made by CLUEValidatorESProducer from the input path), ValidationMode, and, with an
// output directory, the CSV writer's directory.
// usage: run_module <plugin lib> <product file> <input path> <dc> <rhoc> <odf> <strict 0|1>
//                   [output dir]
//   product file: int32 n; f32 x[n], y[n]; i32 layer[n]; f32 weight[n], rho[n], delta[n];
//                 i32 nearestHigher[n], isSeed[n], clusterIndex[n]
// Exit 0 when every module returned, 1 with "Caught std::runtime_error" and the message otherwise.
#include <dlfcn.h>

#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "DataFormats/CLUEValidatorTypes.h"
#include "DataFormats/CLUE_config.h"
#include "DataFormats/PointsCloud.h"
#include "Framework/ESPluginFactory.h"
#include "Framework/Event.h"
#include "Framework/EventSetup.h"
#include "Framework/PluginFactory.h"
#include "Framework/ProductRegistry.h"
#include "Framework/Worker.h"

template <class T>
static void read_into(FILE* f, std::vector<T>& v, int n) {
  v.resize(n);
  if (n && std::fread(v.data(), sizeof(T), n, f) != static_cast<size_t>(n)) throw std::runtime_error("short product file");
}

int main(int argc, char** argv) {
  if (argc < 8) {
    std::cerr << "usage: run_module <plugin lib> <product> <input path> <dc> <rhoc> <odf> <strict> [output dir]\n";
    return 2;
  }
  if (!dlopen(argv[1], RTLD_NOW | RTLD_GLOBAL)) {
    std::cerr << "cannot load " << argv[1] << ": " << dlerror() << "\n";
    return 2;
  }
  try {
    PointsCloudSerial pc;
    FILE* f = std::fopen(argv[2], "rb");
    if (!f) throw std::runtime_error("cannot open product file");
    int n = 0;
    if (std::fread(&n, 4, 1, f) != 1) throw std::runtime_error("short product file");
    read_into(f, pc.x, n);
    read_into(f, pc.y, n);
    read_into(f, pc.layer, n);
    read_into(f, pc.weight, n);
    read_into(f, pc.rho, n);
    read_into(f, pc.delta, n);
    read_into(f, pc.nearestHigher, n);
    read_into(f, pc.isSeed, n);
    read_into(f, pc.clusterIndex, n);
    std::fclose(f);

    edm::EventSetup es;
    Parameters par;
    par.dc = std::stof(argv[4]);
    par.rhoc = std::stof(argv[5]);
    par.outlierDeltaFactor = std::stof(argv[6]);
    par.produceOutput = argc > 8;
    es.put(std::make_unique<Parameters>(par));
    edm::ESPluginFactory::create("CLUEValidatorESProducer", argv[3])->produce(es);
    es.put(std::make_unique<ValidationMode>(ValidationMode{std::atoi(argv[7]) != 0}));
    std::vector<std::string> modules{"CLUEValidator"};
    if (argc > 8) {
      edm::ESPluginFactory::create("CLUEOutputESProducer", argv[8])->produce(es);
      modules.push_back("CLUEOutputProducer");
    }

    edm::ProductRegistry reg;
    auto token = reg.produces<PointsCloudSerial>();
    std::vector<std::unique_ptr<edm::Worker>> workers;
    int index = 1;
    for (auto const& m : modules) {
      reg.beginModuleConstruction(index++);
      workers.push_back(edm::PluginFactory::create(m, reg));
    }
    edm::Event ev(0, 1, reg);
    ev.emplace(token, std::move(pc));
    for (auto& w : workers) w->doWork(ev, es);
  } catch (std::exception const& e) {
    std::cout << "Caught std::runtime_error\n" << e.what() << std::endl;
    return 1;
  }
  return 0;
}
