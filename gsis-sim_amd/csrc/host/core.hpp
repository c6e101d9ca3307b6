#pragma once
// Basic host types shared by the setup, the solver orchestration and the C ABI:
// a double-precision 3-vector, index types and the error type raised on bad input.
// Mirrors the roles of src/core/{types,vec3,error}.hpp in the reference (gsis::Error
// carries the source location; src/core/error.hpp:16-47).

#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

namespace gsk {

using Index = std::int32_t;   // node / face / cell numbers (.pmesh int32)
using Offset = std::int64_t;  // CSR offsets
inline constexpr Index kInvalid = -1;
inline constexpr double kPi = 3.14159265358979323846;
inline constexpr double kSqrtPi = 1.77245385090551602730;
inline constexpr double kBoltzmann = 1.380649e-23;

struct Vec3 {
  double x = 0.0, y = 0.0, z = 0.0;
  double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
  double& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
  Vec3& operator+=(const Vec3& b) { x += b.x; y += b.y; z += b.z; return *this; }
  Vec3& operator-=(const Vec3& b) { x -= b.x; y -= b.y; z -= b.z; return *this; }
  Vec3& operator*=(double s) { x *= s; y *= s; z *= s; return *this; }
  Vec3& operator/=(double s) { x /= s; y /= s; z /= s; return *this; }
};
inline Vec3 operator+(Vec3 a, const Vec3& b) { return a += b; }
inline Vec3 operator-(Vec3 a, const Vec3& b) { return a -= b; }
inline Vec3 operator-(const Vec3& a) { return {-a.x, -a.y, -a.z}; }
inline Vec3 operator*(double s, Vec3 a) { return a *= s; }
inline Vec3 operator*(Vec3 a, double s) { return a *= s; }
inline Vec3 operator/(Vec3 a, double s) { return a /= s; }
inline double dot(const Vec3& a, const Vec3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline Vec3 cross(const Vec3& a, const Vec3& b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double norm2(const Vec3& a) { return dot(a, a); }
inline double norm(const Vec3& a) { return std::sqrt(norm2(a)); }

class Error : public std::runtime_error {
 public:
  explicit Error(const std::string& m) : std::runtime_error(m) {}
};

// printf-style formatting (the toolchain's libstdc++ has no <format>)
inline std::string strf(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  va_list ap2;
  va_copy(ap2, ap);
  const int n = std::vsnprintf(nullptr, 0, fmt, ap);
  va_end(ap);
  std::string s(n > 0 ? static_cast<std::size_t>(n) : 0, '\0');
  if (n > 0) std::vsnprintf(s.data(), s.size() + 1, fmt, ap2);
  va_end(ap2);
  return s;
}

[[noreturn]] inline void fail_at(const char* file, int line, const std::string& msg) {
  const char* base = file;
  for (const char* p = file; *p; ++p)
    if (*p == '/') base = p + 1;
  throw Error(msg + strf(" [%s:%d]", base, line));
}

#define GSK_CHECK(cond, msg)                              \
  do {                                                    \
    if (!(cond)) ::gsk::fail_at(__FILE__, __LINE__, msg); \
  } while (0)

}  // namespace gsk
