// Puts OutputDirPath for the validator: <input dir>/../output/<input name without its 4-character
// extension>_output.csv, the path the reference derives from the input file
// (src/serial/plugin-CLUEValidator/CLUEValidatorESProducer.cc:8-31). The driver passes the input
// file itself as the configuration path.
#include <filesystem>
#include <memory>
#include <stdexcept>
#include <string>

#include "DataFormats/CLUEValidatorTypes.h"
#include "Framework/ESPluginFactory.h"
#include "Framework/ESProducer.h"
#include "Framework/EventSetup.h"

class CLUEValidatorESProducer : public edm::ESProducer {
public:
  explicit CLUEValidatorESProducer(std::filesystem::path const& inputFile) : input_{inputFile} {
    std::string name = inputFile.filename().string();
    if (name.size() < 4) throw std::runtime_error("CLUEValidatorESProducer: input file name '" + name + "' is too short");
    name.erase(name.size() - 4);
    outName_ = name + "_output.csv";
  }

  void produce(edm::EventSetup& eventSetup) override {
    eventSetup.put(std::make_unique<OutputDirPath>(OutputDirPath{input_.parent_path().parent_path() / "output" / outName_}));
  }

private:
  std::filesystem::path input_;
  std::string outName_;
};

DEFINE_FWK_EVENTSETUP_MODULE(CLUEValidatorESProducer);
