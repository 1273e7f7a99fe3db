// bk_common.h — error model of the C ABI and small numeric helpers shared by the
// host sources.
//
// Error model (crates/lgbm-capi/src/lib.rs:63-90,181): every exported function
// runs inside guard(), which turns any C++ exception into a -1 return and a
// thread-local message readable through bk_last_error().
#pragma once

#include <cmath>
#include <cstdint>
#include <exception>
#include <string>

namespace bk {

std::string& last_error();

template <typename F>
int guard(F&& f) noexcept {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    last_error() = e.what();
  } catch (...) {
    last_error() = "unknown error";
  }
  return -1;
}

// Rust `f64 as i32`: saturating, NaN -> 0.
inline int32_t sat_i32(double x) {
  if (std::isnan(x)) return 0;
  if (x >= 2147483647.0) return 2147483647;
  if (x <= -2147483648.0) return (int32_t)(-2147483647 - 1);
  return (int32_t)x;
}

}  // namespace bk
