// Module registry. Interface of the reference (src/serial/Framework/PluginFactory.h:13-56): the
// DEFINE_FWK_MODULE(T) macro, edm::PluginFactory::create(name, registry), and the polymorphic class
// template edm::PluginFactory::impl::Maker<T>, whose typeinfo symbol is how the reference's build
// indexes which plugin library holds which module (src/serial/Makefile:77-78). The registry behind
// them is NamedRegistry.
#ifndef CLUE_FW_PluginFactory_h
#define CLUE_FW_PluginFactory_h

#include <memory>
#include <string>
#include <vector>

#include "Framework/NamedRegistry.h"
#include "Framework/Worker.h"

namespace edm {
  // one instance for every library of a job: defined in libclue_framework (PluginFactory.cc)
  extern template class NamedRegistry<Worker, ProductRegistry&>;

  namespace PluginFactory {
    using ModuleRegistry = NamedRegistry<Worker, ProductRegistry&>;

    namespace impl {
      // Builds the worker of module class T.
      template <typename T>
      class Maker final : public Factory<Worker, ProductRegistry&> {
      public:
        std::unique_ptr<Worker> make(ProductRegistry& reg) const override { return std::make_unique<WorkerT<T>>(reg); }
      };

      // Enrols Maker<T> under `name`; the value is only there to run this from a static initialiser.
      template <typename T>
      bool enrol(char const* name) {
        ModuleRegistry::instance().enrol(name, std::make_unique<Maker<T>>());
        return true;
      }
    }  // namespace impl

    // Throws std::runtime_error for an unknown module name.
    std::unique_ptr<Worker> create(std::string const& name, ProductRegistry& reg);
    // Names of the modules of every library loaded so far.
    std::vector<std::string> names();
  }  // namespace PluginFactory
}  // namespace edm

#define EDM_PLUGIN_SYM(x, y) EDM_PLUGIN_SYM2(x, y)
#define EDM_PLUGIN_SYM2(x, y) x##y
#define DEFINE_FWK_MODULE(type) \
  static bool const EDM_PLUGIN_SYM(edm_module_enrolled_, __LINE__) = edm::PluginFactory::impl::enrol<type>(#type);

#endif
