// Error plumbing shared by the library's translation units.  No exception
// crosses the C ABI: internal failures throw cdx::Error, and every extern "C"
// entry point converts it into a status code plus a message kept per context.
#pragma once
#include <hip/hip_runtime.h>

#include <condition_variable>
#include <mutex>
#include <stdexcept>
#include <string>

namespace cdx {

// Process-wide gate between HIP graph recording and every other HIP call the
// library makes.  While one thread records a stream, HIP rejects (and
// poisons the recording with) calls such as synchronous copies, pinned
// allocations or device synchronisation made by other threads.  Library
// code that calls HIP holds the gate shared (GateShared, re-entrant per
// thread); recording holds it exclusive (GateExclusive), which waits for
// the shared holders to leave and keeps new ones out (writers first, so a
// busy server cannot starve a recording).  A thread that holds it shared
// gives its hold up while it records.
class Gate {
 public:
  static Gate& get() {
    static Gate* g = new Gate();  // lives until exit
    return *g;
  }
  void lock_shared() {
    std::unique_lock<std::mutex> lk(mu_);
    cv_.wait(lk, [&] { return !writer_ && writers_waiting_ == 0; });
    ++readers_;
  }
  void unlock_shared() {
    std::lock_guard<std::mutex> lk(mu_);
    if (--readers_ == 0) cv_.notify_all();
  }
  void lock() {
    std::unique_lock<std::mutex> lk(mu_);
    ++writers_waiting_;
    cv_.wait(lk, [&] { return !writer_ && readers_ == 0; });
    --writers_waiting_;
    writer_ = true;
  }
  void unlock() {
    std::lock_guard<std::mutex> lk(mu_);
    writer_ = false;
    cv_.notify_all();
  }
  static int& depth() {
    static thread_local int d = 0;
    return d;
  }

 private:
  std::mutex mu_;
  std::condition_variable cv_;
  int readers_ = 0, writers_waiting_ = 0;
  bool writer_ = false;
};

struct GateShared {
  GateShared() {
    if (Gate::depth()++ == 0) Gate::get().lock_shared();
  }
  ~GateShared() {
    if (--Gate::depth() == 0) Gate::get().unlock_shared();
  }
  GateShared(const GateShared&) = delete;
  GateShared& operator=(const GateShared&) = delete;
};

struct GateExclusive {
  int saved;
  GateExclusive() : saved(Gate::depth()) {
    if (saved) Gate::get().unlock_shared();
    Gate::get().lock();
    Gate::depth() = saved + 1;  // shared holds nested inside are no-ops
  }
  ~GateExclusive() {
    Gate::depth() = saved;
    Gate::get().unlock();
    if (saved) Gate::get().lock_shared();
  }
  GateExclusive(const GateExclusive&) = delete;
  GateExclusive& operator=(const GateExclusive&) = delete;
};

enum Status : int {
  OK = 0,
  ERR_INVALID = 1,
  ERR_HIP = 2,
  ERR_UNSUPPORTED_SHAPE = 3,
  ERR_CAPACITY = 4,
  ERR_NO_DEVICE = 5,
  ERR_OUT_OF_MEMORY = 6,
  ERR_NOT_READY = 7,
};

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void hip_check(hipError_t e, const char* what, const char* file, int line) {
  if (e != hipSuccess) {
    int code = (e == hipErrorOutOfMemory) ? ERR_OUT_OF_MEMORY : (e == hipErrorNoDevice ? ERR_NO_DEVICE : ERR_HIP);
    throw Error(code, std::string(what) + ": " + hipGetErrorString(e) + " (" + file + ":" + std::to_string(line) + ")");
  }
}

}  // namespace cdx

#define CDX_HIP_CHECK(x) ::cdx::hip_check((x), #x, __FILE__, __LINE__)
#define CDX_HIP_CHECK_VOID(x) ::cdx::hip_check((x), #x, __FILE__, __LINE__)
#define CDX_REQUIRE(cond, code, msg)              \
  do {                                            \
    if (!(cond)) throw ::cdx::Error((code), (msg)); \
  } while (0)
