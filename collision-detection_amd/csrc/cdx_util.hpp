// Error plumbing shared by the library's translation units.  No exception
// crosses the C ABI: internal failures throw cdx::Error, and every extern "C"
// entry point converts it into a status code plus a message kept per context.
#pragma once
#include <hip/hip_runtime.h>
#include <stdexcept>
#include <string>

namespace cdx {

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
