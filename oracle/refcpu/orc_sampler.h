// ORACLE — test infrastructure only.
//
// Integer sampling machinery of the reference path tracer, restated:
//   hashes                 modules/cpu/math/random.ih:15-64, 214-245
//   PCG32 RandomSampler    modules/cpu/math/random.ih:116-154
//   LDSampler (Owen-scrambled Sobol)  modules/cpu/math/random.ih:262-394
//   Sobol_revSample2/3/4   modules/cpu/math/sobol.ih:38-80
//   Distribution1D         modules/cpu/math/Distribution1D.ih:26-57,
//                          Distribution1D.ispc:14-44
//   Z-order helpers        modules/cpu/render/util.ih:17-60
// The Sobol generator matrices are regenerated from the Joe-Kuo direction
// numbers (new-joe-kuo-6.21201, dimensions 2-5), which is what the
// reference's table (modules/cpu/math/sobol.inl:33, "bit-reversed and
// interleaved") holds.
#pragma once

#include <cstdint>

#include "orc_math.h"

namespace orc {

inline uint32_t MurmurHash3_mix(uint32_t hash, uint32_t k)
{
  k *= 0xcc9e2d51u;
  k = (k << 15) | (k >> 17);
  k *= 0x1b873593u;
  hash ^= k;
  hash = ((hash << 13) | (hash >> 19)) * 5u + 0xe6546b64u;
  return hash;
}

inline uint32_t MurmurHash3_finalize(uint32_t h)
{
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}

inline uint32_t tripleHash(uint32_t x)
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

inline uint32_t OwenHash2(uint32_t x, uint32_t seed)
{
  x ^= x * 0x3d20adeau;
  x += seed;
  x *= (seed >> 16) | 1u;
  x ^= x * 0x05526c56u;
  x ^= x * 0x53a22864u;
  return x;
}

inline uint32_t OwenScramble4(uint32_t x, uint32_t seed)
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

inline uint32_t partitionZOrder8(uint32_t n)
{
  n &= 0xffu;
  n = (n | (n << 4)) & 0x0F0F0F0Fu;
  n = (n | (n << 2)) & 0x33333333u;
  n = (n | (n << 1)) & 0x55555555u;
  return n;
}
inline uint32_t interleaveZOrder16(uint32_t x, uint32_t y)
{
  return partitionZOrder8(x) | (partitionZOrder8(y) << 1);
}

// ---- PCG32 (XSH-RR) --------------------------------------------------------

struct RandomSampler
{
  uint64_t state;
  uint32_t stream;
};

inline uint32_t pcg32(RandomSampler &r)
{
  const uint64_t old = r.state;
  r.state = old * 6364136223846793005ULL + (uint64_t)(r.stream | 1u);
  const uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
  const uint32_t rot = (uint32_t)(old >> 59u);
  return (xorshifted >> rot) | (xorshifted << ((0u - rot) & 31u));
}

inline void RandomSampler_init(RandomSampler &r, uint32_t seed, uint32_t stream)
{
  r.state = 0;
  r.stream = (stream << 1u) | 1u;
  r.state = MurmurHash3_mix((uint32_t)r.state, seed);
  r.state = MurmurHash3_finalize((uint32_t)r.state);
  pcg32(r);
  r.state += seed;
  pcg32(r);
}

inline float RandomSampler_getFloat(RandomSampler &r) { return to_float_unorm(pcg32(r)); }

// ---- Sobol ------------------------------------------------------------------

// rows k = 0..31 of dimensions 1..4 (Joe-Kuo dims 2..5), each entry the
// direction integer m_{k+1} written with its k+1 bits reversed.
struct SobolTable
{
  uint32_t rev[4 * 32];
  SobolTable()
  {
    // Joe-Kuo: s (degree), a (polynomial coefficients), initial m values
    const int s[4] = {1, 2, 3, 3};
    const uint32_t a[4] = {0, 1, 1, 2};
    const uint32_t m0[4][3] = {{1, 0, 0}, {1, 3, 0}, {1, 3, 1}, {1, 1, 1}};
    for (int d = 0; d < 4; d++) {
      uint64_t m[33];
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
        rev[4 * (k - 1) + d] = r;
      }
    }
  }
};

const SobolTable &sobolTable();

template <int N>
inline void Sobol_revSample(uint32_t revIndex, uint32_t *out)
{
  const uint32_t *M = sobolTable().rev;
  for (int d = 0; d < N; d++)
    out[d] = 0;
  for (uint32_t i = 0; revIndex; revIndex <<= 1, i += 4) {
    if (revIndex & 0x80000000u)
      for (int d = 0; d < N; d++)
        out[d] ^= M[i + d];
  }
}

// ---- LDSampler --------------------------------------------------------------

static constexpr uint32_t SCRAMBLE0 = 0x9374b0d3u;
static constexpr uint32_t SCRAMBLE1 = 0x1ec38426u;
static constexpr uint32_t SCRAMBLE2 = 0xdb81c29cu;
static constexpr uint32_t SCRAMBLE3 = 0x5628372bu;
static constexpr uint32_t SCRAMBLE4 = 0xbcb745bdu;

struct LDSampler
{
  uint32_t revIndex;
  uint32_t scramble;
  uint32_t baseIndexLight;
  uint32_t idxMask;
  uint32_t idxMaskLight;
};

inline void LDSampler_init(LDSampler &self,
    uint32_t seed,
    uint32_t sampleIndex,
    uint32_t sampleOffsetLight,
    uint32_t bits,
    uint32_t bitsLight)
{
  self.revIndex = reverseBits(sampleIndex);
  self.scramble = seed;
  self.baseIndexLight = sampleOffsetLight;
  self.idxMask = bits > 31 ? 0xffffffffu : ~(0xffffffffu >> bits);
  self.idxMaskLight = bitsLight > 31 ? 0xffffffffu : ~(0xffffffffu >> bitsLight);
}

inline void LDSampler_nextGroup(LDSampler &self) { self.scramble = tripleHash(self.scramble); }

inline float LDSamplerScramble(uint32_t sample, uint32_t scramble)
{
  return to_float_unorm(reverseBits(OwenHash2(sample, scramble)));
}

inline vec2f LDSampler_get3LightSamples(const LDSampler &self, uint32_t idx, bool split, float &ss)
{
  uint32_t revIndex = split ? reverseBits(self.baseIndexLight + idx) : self.revIndex;
  revIndex = OwenHash2(revIndex, self.scramble);
  revIndex &= split ? self.idxMaskLight : self.idxMask;
  const uint32_t r0 = reverseBits(revIndex);
  uint32_t r12[2];
  Sobol_revSample<2>(revIndex, r12);
  vec2f res;
  res.x = LDSamplerScramble(r0, self.scramble ^ SCRAMBLE0);
  res.y = LDSamplerScramble(r12[0], self.scramble ^ SCRAMBLE1);
  ss = LDSamplerScramble(r12[1], self.scramble ^ SCRAMBLE2);
  return res;
}

inline vec2f LDSampler_getNext4Samples(LDSampler &self, float &ss, uint32_t &u)
{
  LDSampler_nextGroup(self);
  const uint32_t revIndex = OwenHash2(self.revIndex, self.scramble) & self.idxMask;
  const uint32_t r0 = reverseBits(revIndex);
  uint32_t r123[3];
  Sobol_revSample<3>(revIndex, r123);
  vec2f res;
  res.x = LDSamplerScramble(r0, self.scramble ^ SCRAMBLE0);
  res.y = LDSamplerScramble(r123[0], self.scramble ^ SCRAMBLE1);
  ss = LDSamplerScramble(r123[1], self.scramble ^ SCRAMBLE2);
  u = r123[2];
  return res;
}

inline float LDSampler_finalizeDim3(const LDSampler &self, uint32_t revRes)
{
  return LDSamplerScramble(revRes, self.scramble ^ SCRAMBLE3);
}

inline vec2f LDSampler_getNext5Samples(LDSampler &self, vec3ui &ss)
{
  LDSampler_nextGroup(self);
  const uint32_t revIndex = OwenHash2(self.revIndex, self.scramble) & self.idxMask;
  const uint32_t r0 = reverseBits(revIndex);
  uint32_t r1234[4];
  Sobol_revSample<4>(revIndex, r1234);
  vec2f res;
  res.x = LDSamplerScramble(r0, self.scramble ^ SCRAMBLE0);
  res.y = LDSamplerScramble(r1234[0], self.scramble ^ SCRAMBLE1);
  ss = {r1234[1], r1234[2], r1234[3]};
  return res;
}

inline vec2f LDSampler_finalizeDim23(const LDSampler &self, const vec3ui &rev)
{
  return {LDSamplerScramble(rev.x, self.scramble ^ SCRAMBLE2),
      LDSamplerScramble(rev.y, self.scramble ^ SCRAMBLE3)};
}

inline float LDSampler_finalizeDim4(const LDSampler &self, const vec3ui &rev)
{
  return LDSamplerScramble(rev.z, self.scramble ^ SCRAMBLE4);
}

// ---- Distribution1D ----------------------------------------------------------

// in place: function values -> normalised CDF; returns the sum
inline float Distribution1D_create(int size, float *cdf)
{
  float sum = cdf[0];
  for (int i = 1; i < size; i++)
    cdf[i] = sum += cdf[i];
  const float rcpSum = 1.0f / sum;
  const float nextAfter1 = 1.00000011920928955078125f; // 0x1.000002p+0
  for (int i = 0; i < size; i++) {
    const float c = cdf[i];
    cdf[i] = c >= sum ? nextAfter1 : c * rcpSum;
  }
  return sum;
}

struct Sample1D
{
  int idx;
  float frac;
  float pdf;
  float prob;
};

inline Sample1D Distribution1D_sample(int size, const float *cdf, int start, float s)
{
  int first = start;
  int len = size;
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

} // namespace orc
