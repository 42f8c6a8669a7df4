// An event: its id, the EDM stream processing it, and the products its modules put into it.
// Public API of the reference's edm::Event (src/serial/Framework/Event.h:31-53): construction from
// (stream, event, registry), streamID(), eventID(), get(token) and emplace(token, args...). The
// storage is this framework's own: one slot per registered product, each holding the product by an
// owning pointer with a deleter for its type and the product's std::type_index, so that a get()
// before the product exists or with another type fails with an exception instead of reading freed
// or foreign memory.
#ifndef CLUE_FW_Event_h
#define CLUE_FW_Event_h

#include <memory>
#include <stdexcept>
#include <typeindex>
#include <utility>
#include <vector>

#include "Framework/ProductRegistry.h"

namespace edm {
  using StreamID = int;

  class Event {
  public:
    Event(int streamId, int eventId, ProductRegistry const& reg)
        : streamId_(streamId), eventId_(eventId), slots_(reg.size()) {}

    StreamID streamID() const { return streamId_; }
    int eventID() const { return eventId_; }

    template <typename T>
    T const& get(EDGetTokenT<T> const& token) const {
      Slot const& s = slot(token.index());
      if (!s.object) throw std::runtime_error("product requested before it was produced");
      if (s.type != std::type_index(typeid(T))) throw std::logic_error("product requested as another type");
      return *static_cast<T const*>(s.object.get());
    }

    // Constructs the product in place from args (brace initialisation, as for aggregates); a
    // product put twice replaces the first.
    template <typename T, typename... Args>
    void emplace(EDPutTokenT<T> const& token, Args&&... args) {
      Slot& s = slot(token.index());
      s.object = Owned(new T{std::forward<Args>(args)...}, &destroy<T>);
      s.type = std::type_index(typeid(T));
    }

  private:
    using Owned = std::unique_ptr<void, void (*)(void*)>;
    struct Slot {
      Owned object{nullptr, &no_product};
      std::type_index type{typeid(void)};
    };
    template <typename T>
    static void destroy(void* p) {
      delete static_cast<T*>(p);
    }
    static void no_product(void*) {}

    Slot& slot(unsigned int i) {
      if (i >= slots_.size()) throw std::out_of_range("token of a product this event was not built for");
      return slots_[i];
    }
    Slot const& slot(unsigned int i) const {
      if (i >= slots_.size()) throw std::out_of_range("token of a product this event was not built for");
      return slots_[i];
    }

    StreamID streamId_;
    int eventId_;
    std::vector<Slot> slots_;
  };
}  // namespace edm

#endif
