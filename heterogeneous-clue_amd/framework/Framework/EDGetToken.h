// Typed product tokens (API of src/serial/Framework/EDGetToken.h, EDPutToken.h in the reference).
#ifndef CLUE_FW_EDGetToken_h
#define CLUE_FW_EDGetToken_h

namespace edm {
  class ProductRegistry;

  template <typename T>
  class EDGetTokenT {
  public:
    EDGetTokenT() = default;
    unsigned int index() const { return index_; }
    bool isValid() const { return index_ != kInvalid; }

  private:
    friend class ProductRegistry;
    explicit EDGetTokenT(unsigned int i) : index_(i) {}
    static constexpr unsigned int kInvalid = ~0u;
    unsigned int index_ = kInvalid;
  };
}  // namespace edm

#endif
