// Sample generation of the hip path tracer: hashes, Owen-scrambled Sobol
// LD sampler and Distribution1D, following
//   modules/cpu/math/random.ih:15-64, 214-245, 262-394 (hashes, LDSampler)
//   modules/cpu/math/sobol.ih:38-80 (Sobol_revSample2/3/4)
//   modules/cpu/math/Distribution1D.ih:26-57, Distribution1D.ispc:14-44
//   modules/cpu/render/util.ih:17-60 (Z-order)
// The generator matrices (dimensions 2-5 of Joe-Kuo new-joe-kuo-6.21201,
// row k holding direction integer m_k with its k bits reversed -- the
// content of modules/cpu/math/sobol.inl:33) are generated at compile time.
#pragma once

#include "pt_math.h"

namespace hpt {

struct SobolRevTable
{
  uint32_t v[128];
};

constexpr SobolRevTable makeSobolRevTable()
{
  SobolRevTable t{};
  const int s[4] = {1, 2, 3, 3};
  const uint32_t a[4] = {0, 1, 1, 2};
  const uint32_t m0[4][3] = {{1, 0, 0}, {1, 3, 0}, {1, 3, 1}, {1, 1, 1}};
  for (int d = 0; d < 4; d++) {
    uint64_t m[33] = {};
    for (int k = 1; k <= s[d]; k++)
      m[k] = m0[d][k - 1];
    for (int k = s[d] + 1; k <= 32; k++) {
      const int sd = s[d];
      uint64_t v = m[k - sd] ^ (m[k - sd] << sd);
      for (int j = 1; j < sd; j++)
        if ((a[d] >> (sd - 1 - j)) & 1u)
          v ^= m[k - j] << j;
      m[k] = v;
    }
    for (int k = 1; k <= 32; k++) {
      uint32_t r = 0;
      for (int b = 0; b < k; b++)
        if ((m[k] >> b) & 1u)
          r |= 1u << (k - 1 - b);
      t.v[4 * (k - 1) + d] = r;
    }
  }
  return t;
}

#ifndef __HIP_DEVICE_COMPILE__
inline const uint32_t *hostSobol()
{
  static const SobolRevTable t = makeSobolRevTable();
  return t.v;
}
#endif

HPT_HD uint32_t tripleHash(uint32_t x)
{
  x ^= x >> 17;
  x *= 0xed5ad4bbu;
  x ^= x >> 11;
  x *= 0xac4c1b51u;
  x ^= x >> 15;
  x *= 0x31848babu;
  x ^= x >> 14;
  return x;
}

HPT_HD uint32_t OwenHash2(uint32_t x, uint32_t seed)
{
  x ^= x * 0x3d20adeau;
  x += seed;
  x *= (seed >> 16) | 1u;
  x ^= x * 0x05526c56u;
  x ^= x * 0x53a22864u;
  return x;
}

HPT_HD uint32_t OwenScramble4(uint32_t x, uint32_t seed)
{
  x = reverseBits(x);
  x ^= x * 0x3d20adeau;
  x ^= (x >> 1) & (x << 1) & 0x55555555u;
  x += seed;
  x *= (seed >> 16) | 1u;
  x ^= (x >> 1) & (x << 1) & 0x55555555u;
  x ^= x * 0x05526c56u;
  x ^= x * 0x53a22864u;
  return reverseBits(x);
}

HPT_HD uint32_t partitionZOrder8(uint32_t n)
{
  n &= 0xffu;
  n = (n | (n << 4)) & 0x0F0F0F0Fu;
  n = (n | (n << 2)) & 0x33333333u;
  n = (n | (n << 1)) & 0x55555555u;
  return n;
}

HPT_HD uint32_t interleaveZOrder16(uint32_t x, uint32_t y) { return partitionZOrder8(x) | (partitionZOrder8(y) << 1); }

#ifdef __HIP_DEVICE_COMPILE__
template <int N>
struct SobolRevResult
{
  uint32_t v[N];
};

// Fully unrolled over the 32 bits: the matrix rows fold into immediate
// masks, so there are no table loads and no data-dependent loop.  Out of
// line and returning by value: inlined, the compiler hoists the 32 N
// immediates into registers across the caller's loops.
template <int N>
__device__ __attribute__((noinline)) SobolRevResult<N> sobolRevSampleDevice(uint32_t revIndex)
{
  constexpr SobolRevTable T = makeSobolRevTable();
  SobolRevResult<N> r;
#pragma unroll
  for (int d = 0; d < N; d++)
    r.v[d] = 0;
#pragma unroll
  for (int k = 0; k < 32; k++) {
    const uint32_t sel = (uint32_t)((int32_t)(revIndex << k) >> 31); // all ones if bit 31 - k is set
#pragma unroll
    for (int d = 0; d < N; d++)
      r.v[d] ^= sel & T.v[4 * k + d];
  }
  return r;
}
#endif

// out[d] = XOR of the generator-matrix rows k of dimension d for which bit
// 31 - k of revIndex is set
template <int N>
HPT_HD void sobolRevSample(uint32_t revIndex, uint32_t *out)
{
#ifdef __HIP_DEVICE_COMPILE__
  const SobolRevResult<N> r = sobolRevSampleDevice<N>(revIndex);
#pragma unroll
  for (int d = 0; d < N; d++)
    out[d] = r.v[d];
#else
  const uint32_t *M = hostSobol();
  for (int d = 0; d < N; d++)
    out[d] = 0;
  for (uint32_t i = 0; revIndex; revIndex <<= 1, i += 4) {
    if (revIndex & 0x80000000u)
      for (int d = 0; d < N; d++)
        out[d] ^= M[i + d];
  }
#endif
}

constexpr uint32_t SCRAMBLE0 = 0x9374b0d3u;
constexpr uint32_t SCRAMBLE1 = 0x1ec38426u;
constexpr uint32_t SCRAMBLE2 = 0xdb81c29cu;
constexpr uint32_t SCRAMBLE3 = 0x5628372bu;
constexpr uint32_t SCRAMBLE4 = 0xbcb745bdu;

// LDSampler state; the masks are frame-uniform
struct LDSampler
{
  uint32_t revIndex, scramble, baseIndexLight;
  uint32_t idxMask, idxMaskLight;
};

HPT_HD float ldScramble(uint32_t sample, uint32_t scramble)
{
  return to_float_unorm(reverseBits(OwenHash2(sample, scramble)));
}

HPT_HD void ldNextGroup(LDSampler &s) { s.scramble = tripleHash(s.scramble); }

HPT_HD vec2f ldGet3LightSamples(const LDSampler &self, uint32_t idx, bool split, float &ss)
{
  uint32_t revIndex = split ? reverseBits(self.baseIndexLight + idx) : self.revIndex;
  revIndex = OwenHash2(revIndex, self.scramble);
  revIndex &= split ? self.idxMaskLight : self.idxMask;
  const uint32_t r0 = reverseBits(revIndex);
  uint32_t r12[2];
  sobolRevSample<2>(revIndex, r12);
  vec2f res;
  res.x = ldScramble(r0, self.scramble ^ SCRAMBLE0);
  res.y = ldScramble(r12[0], self.scramble ^ SCRAMBLE1);
  ss = ldScramble(r12[1], self.scramble ^ SCRAMBLE2);
  return res;
}

HPT_HD vec2f ldGetNext4Samples(LDSampler &self, float &ss, uint32_t &u)
{
  ldNextGroup(self);
  const uint32_t revIndex = OwenHash2(self.revIndex, self.scramble) & self.idxMask;
  const uint32_t r0 = reverseBits(revIndex);
  uint32_t r123[3];
  sobolRevSample<3>(revIndex, r123);
  vec2f res;
  res.x = ldScramble(r0, self.scramble ^ SCRAMBLE0);
  res.y = ldScramble(r123[0], self.scramble ^ SCRAMBLE1);
  ss = ldScramble(r123[1], self.scramble ^ SCRAMBLE2);
  u = r123[2];
  return res;
}

HPT_HD float ldFinalizeDim3(const LDSampler &self, uint32_t revRes)
{
  return ldScramble(revRes, self.scramble ^ SCRAMBLE3);
}

HPT_HD vec2f ldGetNext5Samples(LDSampler &self)
{
  ldNextGroup(self);
  const uint32_t revIndex = OwenHash2(self.revIndex, self.scramble) & self.idxMask;
  const uint32_t r0 = reverseBits(revIndex);
  uint32_t r1234[4];
  sobolRevSample<4>(revIndex, r1234);
  vec2f res;
  res.x = ldScramble(r0, self.scramble ^ SCRAMBLE0);
  res.y = ldScramble(r1234[0], self.scramble ^ SCRAMBLE1);
  return res;
}

struct Sample1D
{
  int idx;
  float frac, pdf, prob;
};

// the first index whose CDF value exceeds s (upper bound)
HPT_HD Sample1D distribution1DSample(int size, const float *cdf, int start, float s)
{
  int first = start;
  int len = size;
#ifdef __HIP_DEVICE_COMPILE__
  if (size <= 8) {
    // short CDFs (the light lists of most scenes): the CDF does not
    // decrease, so the upper bound is the count of entries <= s; the
    // entries load independently instead of one per search step
    int below = 0;
#pragma unroll
    for (int k = 0; k < 8; k++)
      below += (k < size && !(s < cdf[start + k])) ? 1 : 0;
    first = start + below;
    len = 0;
  }
#endif
  while (len > 0) {
    const int half = len >> 1;
    const int mid = first + half;
    if (s < cdf[mid]) {
      len = half;
    } else {
      first = mid + 1;
      len -= half + 1;
    }
  }
  Sample1D ret;
  ret.idx = first - start;
  const float bef = first == start ? 0.0f : cdf[first - 1];
  const float dpdf = cdf[first] - bef;
  ret.prob = dpdf;
  ret.pdf = dpdf * (float)size;
  ret.frac = (s - bef) * rcp(dpdf);
  return ret;
}

// host-side CDF construction (Distribution1D_create)
inline float distribution1DCreate(int size, float *cdf)
{
  float sum = cdf[0];
  for (int i = 1; i < size; i++)
    cdf[i] = sum += cdf[i];
  const float rcpSum = 1.0f / sum;
  const float nextAfter1 = 1.00000011920928955078125f;
  for (int i = 0; i < size; i++) {
    const float c = cdf[i];
    cdf[i] = c >= sum ? nextAfter1 : c * rcpSum;
  }
  return sum;
}

} // namespace hpt
