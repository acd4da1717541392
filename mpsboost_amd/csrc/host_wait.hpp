// Bounded wait of a host thread for a sequence number that the device publishes into
// host-mapped memory (the per-level split candidates of a host-planned fit).
//
// A level's round trip is 25-45 us, so the wait first spins (with the CPU's pause hint) for
// about that long, then yields its core to other runnable threads a few hundred times, then
// sleeps in short steps. Several fits can wait at once, one thread each; past the spin
// phase a waiting thread no longer holds a core, whatever the number of waiting threads and the
// cores the process may use. The wait ends at a deadline: a sequence number that never arrives
// becomes an error that names what was waited for, not a silent hang.
//
// Header-only and free of HIP calls, so it builds and is tested on a CPU with plain threads.

#pragma once

#include <chrono>
#include <cstdint>
#include <thread>

#include <sched.h>

namespace mpsb {

struct WaitResult {
  bool published;   // the word reached the sequence number
  int stopped;      // non-zero: what poll() returned when it ended the wait
  uint64_t seen;    // last value read
  double seconds;   // time spent waiting
};

constexpr uint32_t kWaitSpinIterations = 2048;   // ~ 40-300 us of pause hints
constexpr uint32_t kWaitYieldIterations = 256;   // then sched_yield
constexpr uint32_t kWaitSleepMicros = 50;        // then sleep steps of this length

inline void cpu_relax() {
#if defined(__x86_64__) || defined(__i386__)
  __builtin_ia32_pause();
#endif
}

// Waits until *word >= seq (acquire load), the deadline passes, or poll() returns non-zero.
// poll() is called every 256 spins, then on every yield or sleep step; it is where the caller
// checks for a failed stream.
template <class Poll>
WaitResult wait_for_sequence(const uint64_t* word, uint64_t seq, double limit_seconds, Poll&& poll) {
  using clock = std::chrono::steady_clock;
  const auto start = clock::now();
  WaitResult r{false, 0, 0, 0.0};
  for (uint64_t i = 0;; ++i) {
    r.seen = __atomic_load_n(word, __ATOMIC_ACQUIRE);
    if (r.seen >= seq) {
      r.published = true;
      break;
    }
    const bool spinning = i < kWaitSpinIterations;
    if (!spinning || i % 256 == 255) {
      if ((r.stopped = poll()) != 0) break;
      r.seconds = std::chrono::duration<double>(clock::now() - start).count();
      if (r.seconds > limit_seconds) break;
    }
    if (spinning)
      cpu_relax();
    else if (i < kWaitSpinIterations + kWaitYieldIterations)
      sched_yield();
    else
      std::this_thread::sleep_for(std::chrono::microseconds(kWaitSleepMicros));
  }
  r.seconds = std::chrono::duration<double>(clock::now() - start).count();
  return r;
}

}  // namespace mpsb
