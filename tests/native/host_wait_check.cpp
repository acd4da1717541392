// CPU check of the bounded host wait (mpsboost_amd/csrc/host_wait.hpp) with plain threads
// standing in for the device: waiters block on host words that producer threads publish.
// Prints one JSON object per scenario; tests/test_host_wait.py asserts on them.
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <vector>

#include <sched.h>
#include <sys/resource.h>
#include <time.h>

#include "host_wait.hpp"

using mpsb::wait_for_sequence;

static double thread_cpu_seconds() {
  timespec t{};
  clock_gettime(CLOCK_THREAD_CPUTIME_ID, &t);
  return double(t.tv_sec) + 1e-9 * double(t.tv_nsec);
}

// n waiters, each publishing level by level like a fit: the producer for waiter k publishes
// sequence numbers 1..levels, each after a busy-work delay, while the waiter waits for each.
static void many_waiters(int waiters, int levels, int delay_us, int cores) {
  if (cores > 0) {
    cpu_set_t set;
    CPU_ZERO(&set);
    for (int c = 0; c < cores; ++c) CPU_SET(c, &set);
    sched_setaffinity(0, sizeof set, &set);
  }
  std::vector<uint64_t> words(size_t(waiters) * 8, 0);  // one cache line per word
  std::atomic<int> done{0}, failed{0};
  std::vector<double> cpu(waiters, 0.0);
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<std::thread> threads;
  for (int k = 0; k < waiters; ++k) {
    uint64_t* w = &words[size_t(k) * 8];
    threads.emplace_back([=, &failed] {
      for (int l = 1; l <= levels; ++l) {
        // Busy work standing in for the device's level: it needs a core to finish.
        const auto until = std::chrono::steady_clock::now() + std::chrono::microseconds(delay_us);
        volatile uint64_t x = 0;
        while (std::chrono::steady_clock::now() < until) x = x + 1;
        __atomic_store_n(w, uint64_t(l), __ATOMIC_RELEASE);
      }
    });
    threads.emplace_back([=, &done, &failed, &cpu] {
      const double c0 = thread_cpu_seconds();
      for (int l = 1; l <= levels; ++l) {
        auto r = wait_for_sequence(w, uint64_t(l), 60.0, [] { return 0; });
        if (!r.published) failed.fetch_add(1);
      }
      cpu[k] = thread_cpu_seconds() - c0;
      done.fetch_add(1);
    });
  }
  for (auto& t : threads) t.join();
  const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  double cpu_sum = 0;
  for (double c : cpu) cpu_sum += c;
  std::printf("{\"scenario\": \"many_waiters\", \"waiters\": %d, \"levels\": %d, \"delay_us\": %d, \"cores\": %d, "
              "\"done\": %d, \"failed\": %d, \"wall_s\": %.4f, \"waiter_cpu_s\": %.4f}\n",
              waiters, levels, delay_us, cores, done.load(), failed.load(), wall, cpu_sum);
}

static void deadline() {
  uint64_t word = 3;
  auto r = wait_for_sequence(&word, 4, 0.2, [] { return 0; });
  std::printf("{\"scenario\": \"deadline\", \"published\": %d, \"stopped\": %d, \"seen\": %llu, \"seconds\": %.4f}\n",
              int(r.published), r.stopped, (unsigned long long)r.seen, r.seconds);
}

static void poll_stops() {
  uint64_t word = 0;
  int calls = 0;
  auto r = wait_for_sequence(&word, 1, 60.0, [&] { return ++calls == 3 ? 7 : 0; });
  std::printf("{\"scenario\": \"poll_stops\", \"published\": %d, \"stopped\": %d, \"calls\": %d, \"seconds\": %.4f}\n",
              int(r.published), r.stopped, calls, r.seconds);
}

static void already_published() {
  uint64_t word = 9;
  int calls = 0;
  auto r = wait_for_sequence(&word, 9, 0.0, [&] { return ++calls, 0; });
  std::printf("{\"scenario\": \"already\", \"published\": %d, \"calls\": %d}\n", int(r.published), calls);
}

int main() {
  already_published();
  deadline();
  poll_stops();
  many_waiters(8, 40, 300, 0);
  // More waiting fits than cores, and the producers need those cores to publish.
  many_waiters(16, 20, 2000, 2);
  return 0;
}
