// Conditions-producer registry. Interface of the reference (src/serial/Framework/ESPluginFactory.h:
// 14-58): DEFINE_FWK_EVENTSETUP_MODULE(T) for a producer constructed from a configuration path,
// edm::ESPluginFactory::create(name, path), and the polymorphic class template
// edm::ESPluginFactory::impl::Maker<T> whose typeinfo indexes the plugin libraries. The registry
// behind them is NamedRegistry.
#ifndef CLUE_FW_ESPluginFactory_h
#define CLUE_FW_ESPluginFactory_h

#include <filesystem>
#include <memory>
#include <string>

#include "Framework/ESProducer.h"
#include "Framework/NamedRegistry.h"

namespace edm {
  // one instance for every library of a job: defined in libclue_framework (PluginFactory.cc)
  extern template class NamedRegistry<ESProducer, std::filesystem::path const&>;

  namespace ESPluginFactory {
    using ProducerRegistry = NamedRegistry<ESProducer, std::filesystem::path const&>;

    namespace impl {
      template <typename T>
      class Maker final : public Factory<ESProducer, std::filesystem::path const&> {
      public:
        std::unique_ptr<ESProducer> make(std::filesystem::path const& config) const override {
          return std::make_unique<T>(config);
        }
      };

      template <typename T>
      bool enrol(char const* name) {
        ProducerRegistry::instance().enrol(name, std::make_unique<Maker<T>>());
        return true;
      }
    }  // namespace impl

    std::unique_ptr<ESProducer> create(std::string const& name, std::filesystem::path const& config);
  }  // namespace ESPluginFactory
}  // namespace edm

#define EDM_ES_PLUGIN_SYM(x, y) EDM_ES_PLUGIN_SYM2(x, y)
#define EDM_ES_PLUGIN_SYM2(x, y) x##y
#define DEFINE_FWK_EVENTSETUP_MODULE(type)                                   \
  static bool const EDM_ES_PLUGIN_SYM(edm_es_module_enrolled_, __LINE__) = \
      edm::ESPluginFactory::impl::enrol<type>(#type);

#endif
