// A registry of factories by name, shared by the module and the conditions-producer registries.
// Factories enrol themselves from static initialisers of the libraries that hold them (one
// registry per product type, created on first use), and are looked up by name when the job is
// configured. Enrolment and lookup happen on one thread, before any event is processed.
#ifndef CLUE_FW_NamedRegistry_h
#define CLUE_FW_NamedRegistry_h

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace edm {
  // Builds one product from an argument; one concrete factory per registered class.
  template <typename Product, typename Arg>
  class Factory {
  public:
    virtual ~Factory() = default;
    virtual std::unique_ptr<Product> make(Arg arg) const = 0;
  };

  template <typename Product, typename Arg>
  class NamedRegistry {
  public:
    using FactoryType = Factory<Product, Arg>;

    static NamedRegistry& instance() {
      static NamedRegistry r;
      return r;
    }
    // A second factory under one name is a configuration error (two libraries define one class).
    void enrol(std::string const& name, std::unique_ptr<FactoryType> f) {
      auto [it, fresh] = byName_.try_emplace(name, std::move(f));
      if (!fresh) throw std::logic_error(kind_ + " " + name + " is registered twice");
    }
    std::unique_ptr<Product> make(std::string const& name, Arg arg) const {
      auto it = byName_.find(name);
      if (it == byName_.end()) {
        std::string known;
        for (auto const& kv : byName_) known += (known.empty() ? "" : ", ") + kv.first;
        throw std::runtime_error("no " + kind_ + " named " + name + " is loaded (loaded: " + known + ")");
      }
      return it->second->make(arg);
    }
    std::vector<std::string> names() const {
      std::vector<std::string> out;
      for (auto const& kv : byName_) out.push_back(kv.first);
      return out;
    }
    void describe(std::string kind) { kind_ = std::move(kind); }

  private:
    std::map<std::string, std::unique_ptr<FactoryType>> byName_;
    std::string kind_ = "plugin";
  };
}  // namespace edm

#endif
