#include <stdexcept>

#include "Framework/ESPluginFactory.h"
#include "Framework/PluginFactory.h"

namespace edm {
  namespace PluginFactory {
    namespace impl {
      void Registry::add(std::string const& name, std::unique_ptr<MakerBase> maker) {
        if (!makers_.emplace(name, std::move(maker)).second)
          throw std::logic_error("module " + name + " registered twice");
      }
      MakerBase const* Registry::get(std::string const& name) const {
        auto it = makers_.find(name);
        return it == makers_.end() ? nullptr : it->second.get();
      }
      std::vector<std::string> Registry::names() const {
        std::vector<std::string> out;
        for (auto const& kv : makers_) out.push_back(kv.first);
        return out;
      }
      Registry& getGlobalRegistry() {
        static Registry r;
        return r;
      }
    }  // namespace impl

    std::unique_ptr<Worker> create(std::string const& name, ProductRegistry& reg) {
      auto const* maker = impl::getGlobalRegistry().get(name);
      if (!maker) throw std::runtime_error("no module named " + name + " is loaded");
      return maker->create(reg);
    }
  }  // namespace PluginFactory

  namespace ESPluginFactory {
    namespace impl {
      void Registry::add(std::string const& name, std::unique_ptr<MakerBase> maker) {
        if (!makers_.emplace(name, std::move(maker)).second)
          throw std::logic_error("ES module " + name + " registered twice");
      }
      MakerBase const* Registry::get(std::string const& name) const {
        auto it = makers_.find(name);
        return it == makers_.end() ? nullptr : it->second.get();
      }
      Registry& getGlobalRegistry() {
        static Registry r;
        return r;
      }
    }  // namespace impl

    std::unique_ptr<ESProducer> create(std::string const& name, std::filesystem::path const& config) {
      auto const* maker = impl::getGlobalRegistry().get(name);
      if (!maker) throw std::runtime_error("no ES module named " + name + " is loaded");
      return maker->create(config);
    }
  }  // namespace ESPluginFactory
}  // namespace edm
