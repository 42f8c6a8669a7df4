#ifndef CLUE_FW_EDPutToken_h
#define CLUE_FW_EDPutToken_h

namespace edm {
  class ProductRegistry;

  template <typename T>
  class EDPutTokenT {
  public:
    EDPutTokenT() = default;
    unsigned int index() const { return index_; }
    bool isValid() const { return index_ != kInvalid; }

  private:
    friend class ProductRegistry;
    explicit EDPutTokenT(unsigned int i) : index_(i) {}
    static constexpr unsigned int kInvalid = ~0u;
    unsigned int index_ = kInvalid;
  };
}  // namespace edm

#endif
