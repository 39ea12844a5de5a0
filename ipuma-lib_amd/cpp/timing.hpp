// timing.hpp -- swatlib::TickTock, the wall-clock interval the reference's
// Job and SubmittedBatch carry (src/swatlib/timing.hpp, used at
// src/ipu_batch_affine.h:28-48 and ipusw/ipumultidriver.hpp:55-80): tick()
// starts an interval, tock() ends it and adds it to a running total.
#ifndef IPUMA_AMD_TIMING_HPP
#define IPUMA_AMD_TIMING_HPP

#include <chrono>
#include <cstdint>
#include <stdexcept>

namespace swatlib {

class TickTock {
 public:
  using clock = std::chrono::steady_clock;

  void tick() {
    begin_ = clock::now();
    running_ = true;
    closed_ = false;
  }
  void tock() {
    if (!running_) throw std::runtime_error("Not started.");
    end_ = clock::now();
    closed_ = true;
    total_ += std::chrono::duration_cast<std::chrono::microseconds>(end_ - begin_);
  }
  // sum of every tick..tock interval, in microseconds
  int64_t accumulate_microseconds() const { return total_.count(); }
  void add(const TickTock &other) { total_ += other.total_; }
  // an interval measured elsewhere (e.g. HIP events), added to the total
  void add_microseconds(int64_t us) { total_ += std::chrono::microseconds(us); }
  // the last tick..tock interval
  template <typename Unit = std::chrono::milliseconds>
  typename Unit::rep duration() const {
    if (!(running_ && closed_)) throw std::runtime_error("Not started and ended yet.");
    return std::chrono::duration_cast<Unit>(end_ - begin_).count();
  }

 private:
  clock::time_point begin_{}, end_{};
  std::chrono::microseconds total_{0};
  bool running_ = false, closed_ = false;
};

}  // namespace swatlib
#endif
