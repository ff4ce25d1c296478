// Shared definitions of the MI355X tensor-block engine (host side).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "talsh.h"

namespace xt {

constexpr int kMaxRank = MAX_TENSOR_RANK;

// Size in bytes of one element of a TAL-SH data kind, 0 if invalid.
inline int kind_size(int kind)
{
  switch (kind) {
  case R4: return 4;
  case R8: return 8;
  case C4: return 8;
  case C8: return 16;
  default: return 0;
  }
}
inline bool kind_complex(int kind) { return kind == C4 || kind == C8; }
inline bool kind_double(int kind) { return kind == R8 || kind == C8; }

// Engine diagnostics go to stderr with a fixed prefix; errors are also
// returned as TAL-SH status codes, never swallowed.
void log_error(const char *fmt, ...);

#define XT_HIP_OK(expr)                                                                              \
  ([&]() -> bool {                                                                                   \
    hipError_t e_ = (expr);                                                                          \
    if (e_ != hipSuccess) {                                                                          \
      ::xt::log_error("%s:%d %s -> %s", __FILE__, __LINE__, #expr, hipGetErrorString(e_));           \
      return false;                                                                                  \
    }                                                                                                \
    return true;                                                                                     \
  }())

// Fast unsigned division by a runtime-invariant divisor (round-up method of
// Granlund & Montgomery, "Division by invariant integers using
// multiplication", PLDI 1994): q = (mulhi(n, m) + n) >> s for n < 2^31.
struct FastDiv {
  uint32_t d = 1, m = 0, s = 0;
  FastDiv() = default;
  explicit FastDiv(uint32_t div)
  {
    d = div;
    uint32_t l = 0;
    while ((1ull << l) < div) ++l;
    s = l;
    // m = floor(2^32 * (2^l - d) / d) + 1
    m = (uint32_t)(((1ull << 32) * ((1ull << l) - div)) / div + 1);
    if (div == 1) { m = 0; s = 0; }
  }
};

}  // namespace xt
