// Scalar/vector arithmetic of the hip path tracer, usable on host and device.
//
// The path tracer's float arithmetic is written out operation by operation
// and the module is compiled with -ffp-contract=off and IEEE division/sqrt,
// so a given input produces the same bits on the MI355X and on a CPU.
// Conventions (the reference's ISPC stdlib / rkcommon equivalents are not
// available to restate exactly):
//   rcp(x) = 1/x, rsqrt(x) = 1/sqrt(x), normalize(v) = v * rsqrt(dot(v,v));
//   sin/cos/acos/exp/log: Cody-Waite reduction + Cephes single-precision
//   minimax polynomials (S. L. Moshier), Horner steps as explicit fmaf;
//   frame(N): tangent from the larger of (1,0,0)xN and (0,1,0)xN.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#define HPT_HD __host__ __device__ __forceinline__

namespace hpt {

constexpr float kInf = __builtin_inff();
constexpr float kPi = 3.14159265358979323846f;
constexpr float kTwoPi = 6.28318530717958647692f;
constexpr float kFourPi = 12.5663706143591729538f;
constexpr float kOneOverPi = 0.318309886183790671538f;
constexpr float kPiOver2 = 1.57079632679489661923f;
constexpr float kUlpEpsilon = 0.000000504636764526367187500f; // 0x1.0fp-21
constexpr uint32_t kInvalidID = 0xffffffffu;

HPT_HD float rcp(float x) { return 1.0f / x; }
HPT_HD float rsqrt(float x) { return 1.0f / sqrtf(x); }
HPT_HD float sqr(float x) { return x * x; }
HPT_HD float minf(float a, float b) { return a < b ? a : b; }
HPT_HD float maxf(float a, float b) { return a > b ? a : b; }
HPT_HD float clampf(float x, float lo = 0.f, float hi = 1.f) { return x < lo ? lo : (x > hi ? hi : x); }
HPT_HD uint32_t fbits(float f) { return __builtin_bit_cast(uint32_t, f); }
HPT_HD float bitsf(uint32_t u) { return __builtin_bit_cast(float, u); }
HPT_HD float to_float_unorm(uint32_t u) { return (float)u * 2.3283064365386963e-10f; }

HPT_HD uint32_t reverseBits(uint32_t x)
{
#ifdef __HIP_DEVICE_COMPILE__
  return __builtin_bitreverse32(x);
#else
  x = ((x >> 1) & 0x55555555u) | ((x & 0x55555555u) << 1);
  x = ((x >> 2) & 0x33333333u) | ((x & 0x33333333u) << 2);
  x = ((x >> 4) & 0x0F0F0F0Fu) | ((x & 0x0F0F0F0Fu) << 4);
  x = ((x >> 8) & 0x00FF00FFu) | ((x & 0x00FF00FFu) << 8);
  return (x >> 16) | (x << 16);
#endif
}

// ---- transcendental functions ------------------------------------------------

HPT_HD float reduce_pio2(float x, int &q)
{
  const float k = rintf(x * 0.636619772367581343076f);
  q = (int)k;
  float r = fmaf(-k, 1.57079637050628662109375f, x);
  r = fmaf(-k, -4.37113900018624283e-8f, r);
  r = fmaf(-k, -1.71512451438197e-15f, r);
  return r;
}

HPT_HD void fsincos(float x, float &s, float &c)
{
  int q;
  const float r = reduce_pio2(x, q);
  const float z = r * r;
  float ps = fmaf(z, -1.9515295891e-4f, 8.3321608736e-3f);
  ps = fmaf(z, ps, -1.6666654611e-1f);
  const float sr = fmaf(r * z, ps, r);
  float pc = fmaf(z, 2.443315711809948e-5f, -1.388731625493765e-3f);
  pc = fmaf(z, pc, 4.166664568298827e-2f);
  const float tc = fmaf(z, -0.5f, 1.0f);
  const float cr = fmaf(z * z, pc, tc);
  switch (q & 3) {
  case 0: s = sr; c = cr; break;
  case 1: s = cr; c = -sr; break;
  case 2: s = -sr; c = -cr; break;
  default: s = -cr; c = sr; break;
  }
}

HPT_HD float fasin_core(float x)
{
  const float z = x * x;
  float p = fmaf(z, 4.2163199048e-2f, 2.4181311049e-2f);
  p = fmaf(z, p, 4.5470025998e-2f);
  p = fmaf(z, p, 7.4953002686e-2f);
  p = fmaf(z, p, 1.6666752422e-1f);
  return fmaf(x * z, p, x);
}

HPT_HD float facos(float x)
{
  if (x < -0.5f)
    return kPi - 2.0f * fasin_core(sqrtf(0.5f * (1.0f + x)));
  if (x > 0.5f)
    return 2.0f * fasin_core(sqrtf(0.5f * (1.0f - x)));
  return kPiOver2 - fasin_core(x);
}

HPT_HD float fldexp(float x, int e)
{
  if (e > 127) {
    x *= bitsf(0x7f000000u);
    e -= 127;
    if (e > 127)
      e = 127;
  }
  if (e < -126) {
    x *= bitsf(0x00800000u);
    e += 126;
    if (e < -126)
      e = -126;
  }
  return x * bitsf((uint32_t)(e + 127) << 23);
}

HPT_HD float fexp(float x)
{
  if (x != x)
    return x;
  if (x > 88.72283905206835f)
    return kInf;
  if (x < -103.972077083991796f)
    return 0.0f;
  const float n = floorf(fmaf(x, 1.44269504088896341f, 0.5f));
  float r = fmaf(-n, 0.693359375f, x);
  r = fmaf(-n, -2.12194440e-4f, r);
  const float z = r * r;
  float p = fmaf(r, 1.9875691500e-4f, 1.3981999507e-3f);
  p = fmaf(r, p, 8.3334519073e-3f);
  p = fmaf(r, p, 4.1665795894e-2f);
  p = fmaf(r, p, 1.6666665459e-1f);
  p = fmaf(r, p, 5.0000001201e-1f);
  const float y = fmaf(p, z, r) + 1.0f;
  return fldexp(y, (int)n);
}

HPT_HD float flog(float x);

// x^y for x >= 0 as exp(y log x) (the power-cosine lobe of the OBJ specular
// BSDF, math/sampling.ih:70-87, Specular.ih:23-78)
HPT_HD float fpow(float x, float y)
{
  if (x == 0.0f)
    return y > 0.0f ? 0.0f : (y == 0.0f ? 1.0f : kInf);
  return fexp(y * flog(x));
}

HPT_HD float flog(float x)
{
  if (x != x || x < 0.0f)
    return __builtin_nanf("");
  if (x == 0.0f)
    return -kInf;
  if (x == kInf)
    return kInf;
  uint32_t u = fbits(x);
  int e = (int)((u >> 23) & 0xff);
  if (e == 0) {
    x *= 8388608.0f;
    u = fbits(x);
    e = (int)((u >> 23) & 0xff) - 23;
  }
  e -= 126;
  float m = bitsf((u & 0x807fffffu) | 0x3f000000u);
  if (m < 0.707106781186547524f) {
    e -= 1;
    m = m + m - 1.0f;
  } else {
    m = m - 1.0f;
  }
  const float z = m * m;
  float p = fmaf(m, 7.0376836292e-2f, -1.1514610310e-1f);
  p = fmaf(m, p, 1.1676998740e-1f);
  p = fmaf(m, p, -1.2420140846e-1f);
  p = fmaf(m, p, 1.4249322787e-1f);
  p = fmaf(m, p, -1.6668057665e-1f);
  p = fmaf(m, p, 2.0000714765e-1f);
  p = fmaf(m, p, -2.4999993993e-1f);
  p = fmaf(m, p, 3.3333331174e-1f);
  float y = (p * m) * z;
  const float fe = (float)e;
  y = fmaf(fe, -2.12194440e-4f, y);
  y = fmaf(z, -0.5f, y);
  float r = m + y;
  r = fmaf(fe, 0.693359375f, r);
  return r;
}

// ---- vectors ------------------------------------------------------------------

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

HPT_HD vec2f v2(float x, float y) { return {x, y}; }
HPT_HD vec3f v3(float a) { return {a, a, a}; }
HPT_HD vec3f v3(float x, float y, float z) { return {x, y, z}; }
HPT_HD vec3f operator+(vec3f a, vec3f b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
HPT_HD vec3f operator-(vec3f a, vec3f b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
HPT_HD vec3f operator*(vec3f a, vec3f b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
HPT_HD vec3f operator*(float s, vec3f a) { return {s * a.x, s * a.y, s * a.z}; }
HPT_HD vec3f operator*(vec3f a, float s) { return {a.x * s, a.y * s, a.z * s}; }
HPT_HD vec3f operator/(vec3f a, float s) { return {a.x / s, a.y / s, a.z / s}; }
HPT_HD vec3f neg(vec3f a) { return {-a.x, -a.y, -a.z}; }
HPT_HD float dot(vec3f a, vec3f b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
HPT_HD vec3f cross(vec3f a, vec3f b)
{
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
HPT_HD float length(vec3f a) { return sqrtf(dot(a, a)); }
HPT_HD vec3f normalize(vec3f a) { return a * rsqrt(dot(a, a)); }
HPT_HD vec3f vabs(vec3f a) { return {fabsf(a.x), fabsf(a.y), fabsf(a.z)}; }
HPT_HD vec3f vmin(vec3f a, vec3f b) { return {minf(a.x, b.x), minf(a.y, b.y), minf(a.z, b.z)}; }
HPT_HD vec3f vmax(vec3f a, vec3f b) { return {maxf(a.x, b.x), maxf(a.y, b.y), maxf(a.z, b.z)}; }
HPT_HD float reduce_max(vec3f a) { return maxf(maxf(a.x, a.y), a.z); }
HPT_HD float reduce_min(vec3f a) { return minf(minf(a.x, a.y), a.z); }
HPT_HD bool eq(vec3f a, vec3f b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
HPT_HD float luminance(vec3f c) { return 0.212671f * c.x + 0.715160f * c.y + 0.072169f * c.z; }
HPT_HD float distance(vec3f a, vec3f b) { return length(a - b); }
HPT_HD float comp(vec3f v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }
HPT_HD vec3f ld3(const float *p) { return {p[0], p[1], p[2]}; }

struct linear3f
{
  vec3f vx, vy, vz;
};
HPT_HD vec3f mul(const linear3f &m, vec3f v) { return m.vx * v.x + m.vy * v.y + m.vz * v.z; }
HPT_HD vec3f mulT(const linear3f &m, vec3f v) { return {dot(m.vx, v), dot(m.vy, v), dot(m.vz, v)}; }

HPT_HD linear3f frame(vec3f N)
{
  const vec3f dx0 = v3(0.0f, N.z, -N.y);
  const vec3f dx1 = v3(-N.z, 0.0f, N.x);
  const vec3f dx = normalize(dot(dx0, dx0) > dot(dx1, dx1) ? dx0 : dx1);
  const vec3f dy = normalize(cross(N, dx));
  return {dx, dy, N};
}

// affine inverse through the adjoint (rows of the inverse are cross products
// of the columns); host-side instance set-up
struct affine3f
{
  linear3f l;
  vec3f p;
};
HPT_HD affine3f inverse(const affine3f &a)
{
  const float d = rcp(dot(a.l.vx, cross(a.l.vy, a.l.vz)));
  const vec3f r0 = cross(a.l.vy, a.l.vz) * d;
  const vec3f r1 = cross(a.l.vz, a.l.vx) * d;
  const vec3f r2 = cross(a.l.vx, a.l.vy) * d;
  const linear3f il = {{r0.x, r1.x, r2.x}, {r0.y, r1.y, r2.y}, {r0.z, r1.z, r2.z}};
  return {il, neg(mul(il, a.p))};
}

} // namespace hpt
