// The module and conditions-producer registries live in this library, so that every plugin
// library loaded into a job enrols into the same two instances.
#include "Framework/ESPluginFactory.h"
#include "Framework/PluginFactory.h"

namespace edm {
  template class NamedRegistry<Worker, ProductRegistry&>;
  template class NamedRegistry<ESProducer, std::filesystem::path const&>;

  namespace PluginFactory {
    std::unique_ptr<Worker> create(std::string const& name, ProductRegistry& reg) {
      auto& r = ModuleRegistry::instance();
      r.describe("module");
      return r.make(name, reg);
    }
    std::vector<std::string> names() { return ModuleRegistry::instance().names(); }
  }  // namespace PluginFactory

  namespace ESPluginFactory {
    std::unique_ptr<ESProducer> create(std::string const& name, std::filesystem::path const& config) {
      auto& r = ProducerRegistry::instance();
      r.describe("conditions producer");
      return r.make(name, config);
    }
  }  // namespace ESPluginFactory
}  // namespace edm
