// One product per C++ type per event: produces<T>() registers it (a second registration throws),
// consumes<T>() looks it up (throws if nobody produces it). Same contract as the reference's
// edm::ProductRegistry (src/serial/Framework/ProductRegistry.h:15-70).
#ifndef CLUE_FW_ProductRegistry_h
#define CLUE_FW_ProductRegistry_h

#include <set>
#include <stdexcept>
#include <string>
#include <typeindex>
#include <unordered_map>

#include "Framework/EDGetToken.h"
#include "Framework/EDPutToken.h"

namespace edm {
  class ProductRegistry {
  public:
    static constexpr int kSourceIndex = 0;

    template <typename T>
    EDPutTokenT<T> produces() {
      const std::type_index ti{typeid(T)};
      const unsigned int index = static_cast<unsigned int>(products_.size());
      if (!products_.emplace(ti, Entry{module_, index}).second)
        throw std::runtime_error(std::string("Product of type ") + typeid(T).name() + " already exists");
      return EDPutTokenT<T>{index};
    }

    template <typename T>
    EDGetTokenT<T> consumes() {
      auto it = products_.find(std::type_index(typeid(T)));
      if (it == products_.end())
        throw std::runtime_error(std::string("Product of type ") + typeid(T).name() + " is not produced");
      consumed_.insert(it->second.module);
      return EDGetTokenT<T>{it->second.index};
    }

    unsigned int size() const { return static_cast<unsigned int>(products_.size()); }

    void beginModuleConstruction(int module) {
      module_ = module;
      consumed_.clear();
    }
    std::set<int> const& consumedModules() const { return consumed_; }

  private:
    struct Entry {
      int module;
      unsigned int index;
    };
    int module_ = kSourceIndex;
    std::set<int> consumed_;
    std::unordered_map<std::type_index, Entry> products_;
  };
}  // namespace edm

#endif
