// Conditions producer (API of src/serial/Framework/ESProducer.h:7-13); constructed from the path
// of its configuration file by the ES plugin factory.
#ifndef CLUE_FW_ESProducer_h
#define CLUE_FW_ESProducer_h

namespace edm {
  class EventSetup;

  class ESProducer {
  public:
    ESProducer() = default;
    virtual ~ESProducer() = default;
    virtual void produce(EventSetup& eventSetup) = 0;
  };
}  // namespace edm

#endif
