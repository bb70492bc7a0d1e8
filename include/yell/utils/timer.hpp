// yell/utils/timer.hpp -- AutoTimer with the interface of the reference's
// include/yell/utils/timer.hpp (used by example.cpp): it adds the elapsed
// milliseconds since construction or reset() to *ret on stop() and on
// destruction, and with a tag prints "<tag> <*ret>" when it adds.
#pragma once
#include <chrono>
#include <iostream>
#include <string>

class AutoTimer {
 public:
  using Time_t = std::chrono::nanoseconds;
  using Clock = std::chrono::high_resolution_clock;

  explicit AutoTimer(double *ret) : ret_(ret), start_(Clock::now()) {}
  AutoTimer(double *ret, std::string const &tag) : verbose_(true), tag_(tag), ret_(ret), start_(Clock::now()) {}
  ~AutoTimer() { stop(); }

  void reset() { start_ = Clock::now(); }
  void stop() {
    if (!ret_) return;
    *ret_ += std::chrono::duration<double, std::milli>(Clock::now() - start_).count();
    if (verbose_) std::cout << tag_ << " " << *ret_ << "\n";
  }

 protected:
  bool verbose_ = false;
  std::string tag_;
  double *ret_ = nullptr;
  Clock::time_point start_;
};
