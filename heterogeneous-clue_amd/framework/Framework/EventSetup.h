// Conditions products by type: put<T> (throws on a duplicate) and get<T> (throws if missing).
// API of the reference's edm::EventSetup (src/serial/Framework/EventSetup.h:29-53).
#ifndef CLUE_FW_EventSetup_h
#define CLUE_FW_EventSetup_h

#include <memory>
#include <stdexcept>
#include <string>
#include <typeindex>
#include <unordered_map>

namespace edm {
  class EventSetup {
  public:
    template <typename T>
    void put(std::unique_ptr<T> prod) {
      auto ok = products_.try_emplace(std::type_index(typeid(T)), std::make_unique<Holder<T>>(std::move(prod)));
      if (!ok.second) throw std::runtime_error(std::string("Product of type ") + typeid(T).name() + " already exists");
    }

    template <typename T>
    T const& get() const {
      auto it = products_.find(std::type_index(typeid(T)));
      if (it == products_.end())
        throw std::runtime_error(std::string("Product of type ") + typeid(T).name() + " is not produced");
      return *static_cast<Holder<T> const&>(*it->second).obj;
    }

  private:
    struct HolderBase {
      virtual ~HolderBase() = default;
    };
    template <typename T>
    struct Holder : HolderBase {
      explicit Holder(std::unique_ptr<T> p) : obj(std::move(p)) {}
      std::unique_ptr<T> obj;
    };
    std::unordered_map<std::type_index, std::unique_ptr<HolderBase>> products_;
  };
}  // namespace edm

#endif
