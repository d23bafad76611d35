// ORACLE — test infrastructure only.  CPU restatement of the reference
// OSPRay path-tracer arithmetic; never linked into the product library.
//
// Small-vector algebra and the scalar transcendental functions used on the
// path.  The reference uses ISPC stdlib sin/cos/acos/exp/log and rkcommon
// rcp/rsqrt/frame(); neither library is available here (parity unpinned at
// that boundary), so this file states them explicitly:
//   * rcp(x)   = 1/x      (IEEE division, correctly rounded)
//   * rsqrt(x) = 1/sqrt(x) (two correctly rounded operations)
//   * sin/cos/acos/exp/log: Cody-Waite range reduction plus the Cephes
//     single-precision minimax polynomials (S. L. Moshier, Cephes Math
//     Library), evaluated with explicit fmaf so the result does not depend on
//     the compiler's contraction choices.  Accuracy is checked against libm
//     in tests/test_oracle_math.py.
// Everything here must be compiled with -ffp-contract=off.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>

namespace orc {

static constexpr float kInf = __builtin_inff();
static constexpr float kPi = 3.14159265358979323846f;
static constexpr float kTwoPi = 6.28318530717958647692f;
static constexpr float kFourPi = 12.5663706143591729538f;
static constexpr float kOneOverPi = 0.318309886183790671538f;
static constexpr float kOneOverTwoPi = 0.159154943091895335769f;
static constexpr float kPiOver2 = 1.57079632679489661923f;

inline float rcp(float x) { return 1.0f / x; }
inline float rsqrt(float x) { return 1.0f / std::sqrt(x); }
inline float sqr(float x) { return x * x; }
inline float clampf(float x, float lo = 0.f, float hi = 1.f)
{
  return x < lo ? lo : (x > hi ? hi : x);
}
inline float minf(float a, float b) { return a < b ? a : b; }
inline float maxf(float a, float b) { return a > b ? a : b; }
inline uint32_t fbits(float f)
{
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}
inline float bitsf(uint32_t u)
{
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

// u * 2^-32: the unit-interval float of a 32-bit integer (rkcommon
// to_float_unorm; restated, parity unpinned).
inline float to_float_unorm(uint32_t u) { return (float)u * 2.3283064365386963e-10f; }

inline uint32_t reverseBits(uint32_t x)
{
  x = ((x >> 1) & 0x55555555u) | ((x & 0x55555555u) << 1);
  x = ((x >> 2) & 0x33333333u) | ((x & 0x33333333u) << 2);
  x = ((x >> 4) & 0x0F0F0F0Fu) | ((x & 0x0F0F0F0Fu) << 4);
  x = ((x >> 8) & 0x00FF00FFu) | ((x & 0x00FF00FFu) << 8);
  return (x >> 16) | (x << 16);
}

// ---- transcendental functions -------------------------------------------

// returns r in [-pi/4, pi/4] and quadrant q with x = q*pi/2 + r
inline float reduce_pio2(float x, int &q)
{
  const float k = std::nearbyint(x * 0.636619772367581343076f);
  q = (int)k;
  float r = std::fma(-k, 1.57079637050628662109375f, x);
  r = std::fma(-k, -4.37113900018624283e-8f, r);
  r = std::fma(-k, -1.71512451438197e-15f, r);
  return r;
}

inline float sin_poly(float r)
{
  const float z = r * r;
  float p = std::fma(z, -1.9515295891e-4f, 8.3321608736e-3f);
  p = std::fma(z, p, -1.6666654611e-1f);
  return std::fma(r * z, p, r);
}

inline float cos_poly(float r)
{
  const float z = r * r;
  float p = std::fma(z, 2.443315711809948e-5f, -1.388731625493765e-3f);
  p = std::fma(z, p, 4.166664568298827e-2f);
  const float t = std::fma(z, -0.5f, 1.0f);
  return std::fma(z * z, p, t);
}

inline void fsincos(float x, float &s, float &c)
{
  int q;
  const float r = reduce_pio2(x, q);
  const float sr = sin_poly(r);
  const float cr = cos_poly(r);
  switch (q & 3) {
  case 0: s = sr; c = cr; break;
  case 1: s = cr; c = -sr; break;
  case 2: s = -sr; c = -cr; break;
  default: s = -cr; c = sr; break;
  }
}

inline float fasin_core(float x)
{
  // asin for |x| <= 0.5
  const float z = x * x;
  float p = std::fma(z, 4.2163199048e-2f, 2.4181311049e-2f);
  p = std::fma(z, p, 4.5470025998e-2f);
  p = std::fma(z, p, 7.4953002686e-2f);
  p = std::fma(z, p, 1.6666752422e-1f);
  return std::fma(x * z, p, x);
}

inline float facos(float x)
{
  if (x < -0.5f) {
    const float s = std::sqrt(0.5f * (1.0f + x));
    return kPi - 2.0f * fasin_core(s);
  }
  if (x > 0.5f) {
    const float s = std::sqrt(0.5f * (1.0f - x));
    return 2.0f * fasin_core(s);
  }
  return kPiOver2 - fasin_core(x);
}

inline float fldexp(float x, int e)
{
  // x * 2^e for x in [0.5, 4); two scaling steps cover the float range
  if (e > 127) {
    x *= bitsf(0x7f000000u); // 2^127
    e -= 127;
    if (e > 127)
      e = 127;
  }
  if (e < -126) {
    x *= bitsf(0x00800000u); // 2^-126
    e += 126;
    if (e < -126)
      e = -126;
  }
  return x * bitsf((uint32_t)(e + 127) << 23);
}

inline float fexp(float x)
{
  if (x != x)
    return x;
  if (x > 88.72283905206835f)
    return kInf;
  if (x < -103.972077083991796f)
    return 0.0f;
  const float n = std::floor(std::fma(x, 1.44269504088896341f, 0.5f));
  float r = std::fma(-n, 0.693359375f, x);
  r = std::fma(-n, -2.12194440e-4f, r);
  const float z = r * r;
  float p = std::fma(r, 1.9875691500e-4f, 1.3981999507e-3f);
  p = std::fma(r, p, 8.3334519073e-3f);
  p = std::fma(r, p, 4.1665795894e-2f);
  p = std::fma(r, p, 1.6666665459e-1f);
  p = std::fma(r, p, 5.0000001201e-1f);
  const float y = std::fma(p, z, r) + 1.0f;
  return fldexp(y, (int)n);
}

inline float flog(float x);

// x^y for x >= 0 as exp(y log x) (the power-cosine lobe of the OBJ specular
// BSDF, math/sampling.ih:70-87, Specular.ih:23-78)
inline float fpow(float x, float y)
{
  if (x == 0.0f)
    return y > 0.0f ? 0.0f : (y == 0.0f ? 1.0f : kInf);
  return fexp(y * flog(x));
}

inline float flog(float x)
{
  if (x != x || x < 0.0f)
    return __builtin_nanf("");
  if (x == 0.0f)
    return -kInf;
  if (x == kInf)
    return kInf;
  uint32_t u = fbits(x);
  int e = (int)((u >> 23) & 0xff);
  if (e == 0) { // denormal: scale up
    x *= 8388608.0f; // 2^23
    u = fbits(x);
    e = (int)((u >> 23) & 0xff) - 23;
  }
  e -= 126;
  float m = bitsf((u & 0x807fffffu) | 0x3f000000u); // [0.5, 1)
  if (m < 0.707106781186547524f) {
    e -= 1;
    m = m + m - 1.0f;
  } else {
    m = m - 1.0f;
  }
  const float z = m * m;
  float p = std::fma(m, 7.0376836292e-2f, -1.1514610310e-1f);
  p = std::fma(m, p, 1.1676998740e-1f);
  p = std::fma(m, p, -1.2420140846e-1f);
  p = std::fma(m, p, 1.4249322787e-1f);
  p = std::fma(m, p, -1.6668057665e-1f);
  p = std::fma(m, p, 2.0000714765e-1f);
  p = std::fma(m, p, -2.4999993993e-1f);
  p = std::fma(m, p, 3.3333331174e-1f);
  float y = (p * m) * z;
  const float fe = (float)e;
  y = std::fma(fe, -2.12194440e-4f, y);
  y = std::fma(z, -0.5f, y);
  float r = m + y;
  r = std::fma(fe, 0.693359375f, r);
  return r;
}

// ---- vectors ----------------------------------------------------------------

struct vec2f
{
  float x, y;
};
struct vec3f
{
  float x, y, z;
};
struct vec4f
{
  float x, y, z, w;
};
struct vec3ui
{
  uint32_t x, y, z;
};

inline vec2f v2(float x, float y) { return {x, y}; }
inline vec3f v3(float a) { return {a, a, a}; }
inline vec3f v3(float x, float y, float z) { return {x, y, z}; }
inline vec3f operator+(vec3f a, vec3f b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline vec3f operator-(vec3f a, vec3f b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline vec3f operator*(vec3f a, vec3f b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline vec3f operator*(float s, vec3f a) { return {s * a.x, s * a.y, s * a.z}; }
inline vec3f operator*(vec3f a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline vec3f operator/(vec3f a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline vec3f operator/(vec3f a, vec3f b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
inline vec3f neg(vec3f a) { return {-a.x, -a.y, -a.z}; }
inline float dot(vec3f a, vec3f b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline vec3f cross(vec3f a, vec3f b)
{
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline float length(vec3f a) { return std::sqrt(dot(a, a)); }
inline vec3f normalize(vec3f a) { return a * rsqrt(dot(a, a)); }
inline vec3f vabs(vec3f a) { return {std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)}; }
inline vec3f vmin(vec3f a, vec3f b)
{
  return {minf(a.x, b.x), minf(a.y, b.y), minf(a.z, b.z)};
}
inline vec3f vmax(vec3f a, vec3f b)
{
  return {maxf(a.x, b.x), maxf(a.y, b.y), maxf(a.z, b.z)};
}
inline float reduce_max(vec3f a) { return maxf(maxf(a.x, a.y), a.z); }
inline float reduce_min(vec3f a) { return minf(minf(a.x, a.y), a.z); }
inline float reduce_add(vec3f a) { return a.x + a.y + a.z; }
inline bool eq(vec3f a, vec3f b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
inline float luminance(vec3f c) { return 0.212671f * c.x + 0.715160f * c.y + 0.072169f * c.z; }
inline float distance(vec3f a, vec3f b) { return length(a - b); }

// column-major 3x3 (vx, vy, vz are the columns)
struct linear3f
{
  vec3f vx, vy, vz;
};
inline vec3f mul(const linear3f &m, vec3f v) { return m.vx * v.x + m.vy * v.y + m.vz * v.z; }
inline vec3f mulT(const linear3f &m, vec3f v) { return {dot(m.vx, v), dot(m.vy, v), dot(m.vz, v)}; }

struct affine3f
{
  linear3f l;
  vec3f p;
};
inline vec3f xfmPoint(const affine3f &a, vec3f v) { return mul(a.l, v) + a.p; }
inline vec3f xfmVector(const affine3f &a, vec3f v) { return mul(a.l, v); }

inline float det(const linear3f &m) { return dot(m.vx, cross(m.vy, m.vz)); }
inline linear3f transposed(const linear3f &m)
{
  return {{m.vx.x, m.vy.x, m.vz.x}, {m.vx.y, m.vy.y, m.vz.y}, {m.vx.z, m.vy.z, m.vz.z}};
}
// inverse via the adjoint: rows of the inverse are cross products of columns
inline linear3f inverse(const linear3f &m)
{
  const float d = rcp(det(m));
  const vec3f r0 = cross(m.vy, m.vz) * d;
  const vec3f r1 = cross(m.vz, m.vx) * d;
  const vec3f r2 = cross(m.vx, m.vy) * d;
  return transposed({r0, r1, r2});
}
inline affine3f inverse(const affine3f &a)
{
  const linear3f il = inverse(a.l);
  return {il, neg(mul(il, a.p))};
}

// Orthonormal frame around unit N: tangent from the axis cross product of
// larger magnitude (rkcommon frame(N) restated; parity unpinned).
inline linear3f frame(vec3f N)
{
  const vec3f dx0 = v3(0.0f, N.z, -N.y);
  const vec3f dx1 = v3(-N.z, 0.0f, N.x);
  const vec3f dx = normalize(dot(dx0, dx0) > dot(dx1, dx1) ? dx0 : dx1);
  const vec3f dy = normalize(cross(N, dx));
  return {dx, dy, N};
}

} // namespace orc
