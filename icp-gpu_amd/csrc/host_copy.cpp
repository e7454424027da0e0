// Parallel host copy into pinned staging (host_copy.hpp).
#include "host_copy.hpp"

#include <condition_variable>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

namespace blsctx {
namespace {

// Below this size one thread copies alone (waking the workers costs more than it saves).
constexpr size_t kSplitMin = size_t(1) << 20;
// workers besides the calling thread
constexpr int kWorkers = 3;

// The workers wait on a job generation counter; a job is split into kWorkers + 1 equal
// parts, part 0 copied by the caller.  The pool is allocated once and never freed, so
// that workers still blocked on its condition variable at process exit never see it
// destroyed.
struct Pool {
  std::mutex m;
  std::condition_variable go, done;
  unsigned long gen = 0;
  int pending = 0;
  char* dst = nullptr;
  const char* src = nullptr;
  size_t n = 0, per = 0;

  Pool() {
    for (int w = 1; w <= kWorkers; w++) std::thread([this, w] { run(w); }).detach();
  }
  void part(int k) {
    const size_t a = (size_t)k * per;
    if (a >= n) return;
    const size_t b = a + per < n ? a + per : n;
    std::memcpy(dst + a, src + a, b - a);
  }
  void run(int w) {
    unsigned long seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(m);
        go.wait(lk, [&] { return gen != seen; });
        seen = gen;
      }
      part(w);
      std::lock_guard<std::mutex> lk(m);
      if (--pending == 0) done.notify_one();
    }
  }
  void copy(void* d, const void* s, size_t bytes) {
    {
      std::lock_guard<std::mutex> lk(m);
      dst = static_cast<char*>(d);
      src = static_cast<const char*>(s);
      n = bytes;
      // parts on 4 KiB boundaries
      per = ((bytes + kWorkers) / (kWorkers + 1) + 4095) & ~size_t(4095);
      pending = kWorkers;
      gen++;
    }
    go.notify_all();
    part(0);
    std::unique_lock<std::mutex> lk(m);
    done.wait(lk, [&] { return pending == 0; });
  }
};

}  // namespace

void host_copy(void* dst, const void* src, size_t n) {
  if (n < kSplitMin) {
    std::memcpy(dst, src, n);
    return;
  }
  static Pool* pool = new Pool();
  pool->copy(dst, src, n);
}

}  // namespace blsctx
