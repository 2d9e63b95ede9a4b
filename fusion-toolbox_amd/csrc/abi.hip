// Host-side support of the C ABI: error strings, version, and the exact
// conversion of the reference's angle gate into a cosine threshold.
#include "stb_common.hpp"

#include <cmath>
#include <cstdarg>
#include <cstring>

namespace stb {

namespace {
thread_local char g_error[1024] = "";
}

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_error, sizeof(g_error), fmt, ap);
  va_end(ap);
}

int check_launch(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("%s: kernel launch failed: %s", what, hipGetErrorString(e));
    return STB_ERR_LAUNCH;
  }
  return STB_OK;
}

namespace {

// Order-preserving map of finite floats onto integers (and back), so that a
// bisection over the integers walks the float lattice.
int64_t ord_f32(float f) {
  uint32_t b;
  std::memcpy(&b, &f, 4);
  return (b & 0x80000000u) ? -int64_t(b & 0x7fffffffu) : int64_t(b);
}
float unord_f32(int64_t k) {
  const uint32_t b = k < 0 ? (uint32_t(-k) | 0x80000000u) : uint32_t(k);
  float f;
  std::memcpy(&f, &b, 4);
  return f;
}
// Same for doubles; the ordinal of a double fits in an int64 after folding
// the sign bit (|ordinal| < 2^63).
int64_t ord_f64(double d) {
  uint64_t b;
  std::memcpy(&b, &d, 8);
  return (b >> 63) ? -int64_t(b & 0x7fffffffffffffffull) : int64_t(b);
}
double unord_f64(int64_t k) {
  const uint64_t b =
      k < 0 ? (uint64_t(-k) | 0x8000000000000000ull) : uint64_t(k);
  double d;
  std::memcpy(&d, &b, 8);
  return d;
}

}  // namespace

// The reference gate rejects when fabs(acos(c)) >= thr in the scalar type
// (include/slamtb/math.hpp:14-18, src/correspondence.hpp:85-86). acos is
// non-increasing on [-1, 1], so the rejected c form an interval [-1, c*];
// c* is found by bisection with the same libm acos.
float angle_gate_cos_f32(double angle_thresh) {
  const float thr = float(angle_thresh);
  auto rejects = [thr](float c) { return std::fabs(std::acos(c)) >= thr; };
  if (!rejects(-1.0f)) return -2.0f;
  if (rejects(1.0f)) return 1.0f;
  int64_t lo = ord_f32(-1.0f), hi = ord_f32(1.0f);
  while (hi - lo > 1) {
    const int64_t mid = lo + (hi - lo) / 2;
    if (rejects(unord_f32(mid)))
      lo = mid;
    else
      hi = mid;
  }
  return unord_f32(lo);
}

double angle_gate_cos_f64(double angle_thresh) {
  const double thr = angle_thresh;
  auto rejects = [thr](double c) { return std::fabs(std::acos(c)) >= thr; };
  if (!rejects(-1.0)) return -2.0;
  if (rejects(1.0)) return 1.0;
  int64_t lo = ord_f64(-1.0), hi = ord_f64(1.0);
  while (hi - lo > 1) {
    const int64_t mid = lo + (hi - lo) / 2;
    if (rejects(unord_f64(mid)))
      lo = mid;
    else
      hi = mid;
  }
  return unord_f64(lo);
}

}  // namespace stb

extern "C" const char* stb_last_error(void) { return stb::g_error; }

extern "C" int stb_abi_version(void) { return 2; }

extern "C" size_t stb_icp_state_bytes(void) { return sizeof(stb_icp_state); }

// Exposed for tests: the cosine thresholds used by the correspondence gate.
extern "C" float stb_angle_gate_cos_f32(double angle_thresh) {
  return stb::angle_gate_cos_f32(angle_thresh);
}
extern "C" double stb_angle_gate_cos_f64(double angle_thresh) {
  return stb::angle_gate_cos_f64(angle_thresh);
}
