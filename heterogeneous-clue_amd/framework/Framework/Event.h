// An event: its id, the EDM stream processing it, and its products (owned, type-erased).
// API of the reference's edm::Event (src/serial/Framework/Event.h:31-53).
#ifndef CLUE_FW_Event_h
#define CLUE_FW_Event_h

#include <memory>
#include <stdexcept>
#include <utility>
#include <vector>

#include "Framework/ProductRegistry.h"

namespace edm {
  using StreamID = int;

  class WrapperBase {
  public:
    virtual ~WrapperBase() = default;
  };

  template <typename T>
  class Wrapper : public WrapperBase {
  public:
    template <typename... Args>
    explicit Wrapper(Args&&... args) : obj_{std::forward<Args>(args)...} {}
    T const& product() const { return obj_; }

  private:
    T obj_;
  };

  class Event {
  public:
    Event(int streamId, int eventId, ProductRegistry const& reg)
        : streamId_(streamId), eventId_(eventId), products_(reg.size()) {}

    StreamID streamID() const { return streamId_; }
    int eventID() const { return eventId_; }

    template <typename T>
    T const& get(EDGetTokenT<T> const& token) const {
      auto const& p = products_.at(token.index());
      if (!p) throw std::runtime_error("product requested before it was produced");
      return static_cast<Wrapper<T> const&>(*p).product();
    }

    template <typename T, typename... Args>
    void emplace(EDPutTokenT<T> const& token, Args&&... args) {
      products_.at(token.index()) = std::make_unique<Wrapper<T>>(std::forward<Args>(args)...);
    }

  private:
    StreamID streamId_;
    int eventId_;
    std::vector<std::unique_ptr<WrapperBase>> products_;
  };
}  // namespace edm

#endif
