// Minimal wall-clock timer with the interface the efanna sample programs use
// (boost::timer::auto_cpu_timer, elapsed().wall in nanoseconds).  Boost is
// not installed here; this header only lets the samples compile against the
// efanna facade in the compatibility test.  It measures wall time only.
#ifndef EFG_TEST_COMPAT_BOOST_TIMER_HPP
#define EFG_TEST_COMPAT_BOOST_TIMER_HPP
#include <chrono>
#include <cstdint>

namespace boost {
namespace timer {
struct cpu_times {
  int64_t wall, user, system;
};
class auto_cpu_timer {
 public:
  auto_cpu_timer() : t0_(std::chrono::steady_clock::now()) {}
  cpu_times elapsed() const {
    int64_t ns = std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t0_).count();
    return cpu_times{ns, 0, 0};
  }

 private:
  std::chrono::steady_clock::time_point t0_;
};
}  // namespace timer
}  // namespace boost
#endif
