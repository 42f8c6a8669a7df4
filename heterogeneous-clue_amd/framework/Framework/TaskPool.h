// Fixed pool of worker threads running queued tasks in FIFO order: this framework's stand-in for
// the TBB arena the reference's EventProcessor runs its streams in (src/serial/bin/main.cc:188-197,
// EventProcessor.cc, StreamSchedule.cc). Tasks never wait on each other: an external-work module's
// acquire() returns at once and its completion queues the produce() step, so a few threads serve
// more EDM streams than there are threads while the device works.
#ifndef CLUE_FW_TaskPool_h
#define CLUE_FW_TaskPool_h

#include <condition_variable>
#include <deque>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace edm {
  class TaskPool {
  public:
    explicit TaskPool(int threads);
    TaskPool(TaskPool const&) = delete;
    TaskPool& operator=(TaskPool const&) = delete;
    // Runs every task still queued, then joins the threads.
    ~TaskPool();

    // Queue a task; safe from any thread, including threads the pool does not own (device callbacks).
    // A task must not throw: the pool ends the program if one does.
    void push(std::function<void()> task);

  private:
    void loop();

    std::mutex m_;
    std::condition_variable cv_;
    std::deque<std::function<void()>> queue_;
    bool stop_ = false;
    std::vector<std::thread> threads_;
  };
}  // namespace edm

#endif
