// CLUEValidator: checks each event's CLUE2D output product against the reference result file of a
// toyDetector input, with the acceptance rule of src/serial/plugin-CLUEValidator/CLUEValidator.cc:
//  * the file is <dir of OutputDirPath>/reference/ref_<k>k_20_25_2.csv when the output path names
//    toyDetector_<k>k, k = 1..10 (:90-133); other inputs are reported as not validatable;
//  * inputs: x, y, weight within 1e-3 absolute, layer exact (:42-61, 172-180);
//  * results, only when dc = 20, rhoc = 25, outlierDeltaFactor = 2, the parameters the files were
//    made with (:182-198): rho within 1e-3, delta clamped at 999 within 1e-3, nearestHigher and
//    isSeed exact, cluster ids equal once each seed's id is mapped to the file's id at that seed
//    (-1 stays -1, an unmapped id reads as 0) (:63-88).
// Where the reference stops on a failed assert, a failed check here throws std::runtime_error that
// names the field and the first differing point; a row-count mismatch is a failure too.
// With ValidationMode{strict = true} in the EventSetup, every row the CSV writer would print for a
// point must also equal the file's row byte for byte (no tolerance, cluster ids unmapped).
#include <cmath>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include "DataFormats/CLUECsvRow.h"
#include "DataFormats/CLUEValidatorTypes.h"
#include "DataFormats/CLUE_config.h"
#include "DataFormats/PointsCloud.h"
#include "Framework/EDProducer.h"
#include "Framework/Event.h"
#include "Framework/EventSetup.h"
#include "Framework/PluginFactory.h"

namespace {

  std::string reference_file_for(std::string const& outputPath) {
    for (int k = 1; k <= 10; ++k)
      if (outputPath.find("toyDetector_" + std::to_string(k) + "k") != std::string::npos)
        return "ref_" + std::to_string(k) + "k_20_25_2.csv";
    return {};
  }

  // The file's columns, parsed field by field as the reference does (header skipped, first field of
  // a row ignored; a field that does not parse ends the table), and each row's text.
  struct ReferenceTable {
    PointsCloudSerial pc;
    std::vector<std::string> rows;
  };

  ReferenceTable read_reference(std::filesystem::path const& file, std::ostream& log) {
    std::ifstream in(file);
    if (!in) throw std::runtime_error("CLUEValidator: cannot open reference file " + file.string());
    ReferenceTable t;
    std::string line;
    std::getline(in, line);
    for (int n = 1; std::getline(in, line); ++n) {
      if (!line.empty() && line.back() == '\r') line.pop_back();
      if (line.empty()) continue;
      std::vector<std::string> f;
      std::istringstream ls(line);
      for (std::string v; std::getline(ls, v, ',');) f.push_back(v);
      f.resize(10);
      try {
        const float x = std::stof(f[1]), y = std::stof(f[2]);
        const int layer = std::stoi(f[3]);
        const float weight = std::stof(f[4]), rho = std::stof(f[5]), delta = std::stof(f[6]);
        const int nh = std::stoi(f[7]), seed = std::stoi(f[8]), cluster = std::stoi(f[9]);
        t.pc.x.push_back(x);
        t.pc.y.push_back(y);
        t.pc.layer.push_back(layer);
        t.pc.weight.push_back(weight);
        t.pc.rho.push_back(rho);
        t.pc.delta.push_back(delta);
        t.pc.nearestHigher.push_back(nh);
        t.pc.isSeed.push_back(seed);
        t.pc.clusterIndex.push_back(cluster);
        t.rows.push_back(line);
      } catch (std::exception const& e) {
        log << e.what() << "\nBad Input: '" << line << "' in line " << n << '\n';
        break;
      }
    }
    return t;
  }

  template <class T>
  void require_equal(char const* name, std::vector<T> const& mine, std::vector<T> const& ref, std::ostream& log) {
    if (mine.size() != ref.size())
      throw std::runtime_error(std::string("CLUEValidator: ") + name + ": " + std::to_string(mine.size()) +
                               " points, the reference file holds " + std::to_string(ref.size()));
    for (std::size_t i = 0; i < mine.size(); ++i) {
      bool same;
      if constexpr (std::is_same_v<T, int>)
        same = mine[i] == ref[i];
      else
        same = std::abs(mine[i] - ref[i]) <= 0.001f;
      if (!same) {
        std::ostringstream m;
        m << "CLUEValidator: " << name << " differs from the reference at i=" << i << ", " << mine[i]
          << " /= " << ref[i];
        throw std::runtime_error(m.str());
      }
    }
    log << "Output " << name << " -> Ok\n";
  }

  void require_same_clusters(PointsCloudSerial const& mine, PointsCloudSerial const& ref, std::ostream& log) {
    const std::size_t n = mine.clusterIndex.size();
    if (ref.clusterIndex.size() != n) throw std::runtime_error("CLUEValidator: cluster ids: point counts differ");
    std::unordered_map<int, int> toRef;
    for (std::size_t i = 0; i < n; ++i)
      if (mine.isSeed[i]) toRef[mine.clusterIndex[i]] = ref.clusterIndex[i];
    for (std::size_t i = 0; i < n; ++i) {
      const int id = mine.clusterIndex[i];
      const auto it = toRef.find(id);
      const int mapped = id == -1 ? -1 : (it == toRef.end() ? 0 : it->second);
      if (mapped != ref.clusterIndex[i]) {
        std::ostringstream m;
        m << "CLUEValidator: cluster ids differ from the reference at i=" << i << ", original=" << id
          << ", mapped=" << mapped << " /= " << ref.clusterIndex[i];
        throw std::runtime_error(m.str());
      }
    }
    log << "Output cluster IDs -> Ok\n";
  }

  void require_same_rows(PointsCloudSerial const& mine, ReferenceTable const& ref, std::ostream& log) {
    for (std::size_t i = 0; i < ref.rows.size(); ++i) {
      std::ostringstream row;
      write_clue2d_row(row, mine, i);
      if (row.str() != ref.rows[i])
        throw std::runtime_error("CLUEValidator (strict): row " + std::to_string(i) + " is '" + row.str() +
                                 "', the reference file has '" + ref.rows[i] + "'");
    }
    log << "Output rows -> identical to the reference file\n";
  }

}  // namespace

class CLUEValidator : public edm::EDProducer {
public:
  explicit CLUEValidator(edm::ProductRegistry& reg) : resultsToken_(reg.consumes<PointsCloudSerial>()) {}

private:
  void produce(edm::Event& event, edm::EventSetup const& eventSetup) override {
    auto const& out = eventSetup.get<OutputDirPath>();
    auto const& pc = event.get(resultsToken_);
    Parameters const& par = eventSetup.get<Parameters>();
    bool strict = false;
    try {
      strict = eventSetup.get<ValidationMode>().strict;
    } catch (std::runtime_error const&) {
      // optional: the reference's driver puts no ValidationMode
    }
    const std::string refName = reference_file_for(out.outFile.string());
    std::ostringstream log;  // one write per event, so concurrent streams do not interleave lines
    if (refName.empty()) {
      log << "\nThere is no reference output data for the input file selected.\n"
          << "Please select one of the toyDetectors input files to validate the plugin results\n\n";
      std::cout << log.str() << std::flush;
      return;
    }
    const auto refPath = out.outFile.parent_path() / "reference" / refName;
    log << "Validating output results from " << refPath << '\n';
    try {
      const ReferenceTable ref = read_reference(refPath, log);
      require_equal("x", pc.x, ref.pc.x, log);
      require_equal("y", pc.y, ref.pc.y, log);
      require_equal("layer", pc.layer, ref.pc.layer, log);
      require_equal("weight", pc.weight, ref.pc.weight, log);
      log << "Input variables are correct!\n";
      if (par.dc == 20 && par.rhoc == 25 && par.outlierDeltaFactor == 2) {
        log << "Using the same parameters as reference file, checking output results\n";
        require_equal("rho", pc.rho, ref.pc.rho, log);
        std::vector<float> clamped(pc.delta);
        for (float& d : clamped)
          if (d > 999) d = 999;
        require_equal("delta", clamped, ref.pc.delta, log);
        require_equal("nearestHigher", pc.nearestHigher, ref.pc.nearestHigher, log);
        require_equal("isSeed", pc.isSeed, ref.pc.isSeed, log);
        require_same_clusters(pc, ref.pc, log);
        if (strict) require_same_rows(pc, ref, log);
        log << "CLUE output is correct!\n";
      } else {
        log << "Parameters different from the ones used in the reference file, cannot check results\n";
      }
    } catch (...) {
      std::cout << log.str() << std::flush;
      throw;
    }
    std::cout << log.str() << std::flush;
  }

  edm::EDGetTokenT<PointsCloudSerial> resultsToken_;
};

DEFINE_FWK_MODULE(CLUEValidator);
