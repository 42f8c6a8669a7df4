// Module registry: DEFINE_FWK_MODULE(T) registers T under its class name when the library holding
// it is loaded; create(name, reg) builds a Worker for it. Same names and macro as the reference
// (src/serial/Framework/PluginFactory.h:13-56) so module sources build against either framework.
#ifndef CLUE_FW_PluginFactory_h
#define CLUE_FW_PluginFactory_h

#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "Framework/Worker.h"

namespace edm {
  namespace PluginFactory {
    namespace impl {
      class MakerBase {
      public:
        virtual ~MakerBase() = default;
        virtual std::unique_ptr<Worker> create(ProductRegistry& reg) const = 0;
      };

      template <typename T>
      class Maker : public MakerBase {
      public:
        std::unique_ptr<Worker> create(ProductRegistry& reg) const override {
          return std::make_unique<WorkerT<T>>(reg);
        }
      };

      class Registry {
      public:
        void add(std::string const& name, std::unique_ptr<MakerBase> maker);
        MakerBase const* get(std::string const& name) const;
        std::vector<std::string> names() const;

      private:
        std::unordered_map<std::string, std::unique_ptr<MakerBase>> makers_;
      };

      Registry& getGlobalRegistry();

      template <typename T>
      class Registrar {
      public:
        explicit Registrar(std::string const& name) { getGlobalRegistry().add(name, std::make_unique<Maker<T>>()); }
      };
    }  // namespace impl

    // Throws std::runtime_error for an unknown module name.
    std::unique_ptr<Worker> create(std::string const& name, ProductRegistry& reg);
  }  // namespace PluginFactory
}  // namespace edm

#define EDM_PLUGIN_SYM(x, y) EDM_PLUGIN_SYM2(x, y)
#define EDM_PLUGIN_SYM2(x, y) x##y
#define DEFINE_FWK_MODULE(type) \
  static edm::PluginFactory::impl::Registrar<type> EDM_PLUGIN_SYM(maker, __LINE__)(#type);

#endif
