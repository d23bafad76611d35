// ORACLE — test infrastructure only.
//
// Volumes of the path tracer: structured-regular grids with a piecewise-
// linear transfer function, sampled by delta tracking.  Restates
//   render/pathtracer/volumes/VolumeSampler.ispc:13-279 (delta tracking,
//     free-path sampling, transmittance),
//   volume/transferFunction/LinearTransferFunction.{ispc:11-60,cpp:108-160},
//   render/pathtracer/volumes/HenyeyGreenstein.ih:14-77,
//   with the per-path PCG32 RandomSampler of orc_sampler.h
//   (math/random.ih:116-154).
// The reference delegates grid sampling and interval iteration to OpenVKL
// 2.0.1 (vklComputeSampleV, vklIterateIntervalV; volume/Volume.cpp:60-108),
// which is not part of the reference tree: parity with OpenVKL is
// unpinned.  The oracle fixes its own definitions instead:
//   * samples: trilinear interpolation of vertex-centred voxels at
//     origin + i * spacing; NaN outside [origin, origin + (dims-1)*spacing];
//   * intervals: the ray's path through a grid of macrocells of kMacro^3
//     cells, each interval carrying the min/max voxel value of its
//     macrocell (a conservative majorant).
#pragma once

#include <cmath>
#include <cstdint>
#include <vector>

#include "orc_geom.h"
#include "orc_sampler.h"

namespace orc {

static constexpr int kMacro = 16; // cells per macrocell edge
static constexpr int kTfSubranges = 32; // PRECOMPUTED_OPACITY_SUBRANGE_COUNT

inline float lerpf(float t, float a, float b) { return (1.f - t) * a + t * b; }

struct VolumeData
{
  int64_t dims[3];
  const float *voxels; // x fastest
  vec3f origin, spacing, upper; // upper = origin + (dims-1) * spacing
  int mdims[3]; // macrocells per axis
  std::vector<float> mmin, mmax; // per macrocell
  // transfer function
  std::vector<vec3f> color;
  std::vector<float> opacity;
  float tfLower, tfUpper;
  float maxOpacityInRange[kTfSubranges][kTfSubranges];
  // volumetric model
  float densityScale, anisotropy;
  uint32_t userID;
};

// LinearTransferFunction::precomputeMaxOpacityRanges (LinearTransferFunction.cpp:108-160)
inline void precomputeMaxOpacity(VolumeData &v)
{
  const int maxOpacityDim = (int)v.opacity.size() - 1;
  const int maxPre = kTfSubranges - 1;
  for (int i = 0; i < kTfSubranges; i++) {
    const int lo = (int)std::floor(maxOpacityDim * (float)i / maxPre);
    const int hi = (int)std::ceil(maxOpacityDim * (float)i / maxPre);
    float m = v.opacity[lo];
    for (int k = lo; k <= hi; k++)
      m = std::max(m, v.opacity[k]);
    v.maxOpacityInRange[i][i] = m;
  }
  for (int i = 0; i < kTfSubranges; i++)
    for (int j = i + 1; j < kTfSubranges; j++) {
      const int lo = (int)std::floor(maxOpacityDim * (float)i / maxPre);
      const int hi = (int)std::ceil(maxOpacityDim * (float)j / maxPre);
      float m = v.maxOpacityInRange[i][i];
      for (int k = lo; k <= hi; k++)
        m = std::max(m, v.opacity[k]);
      v.maxOpacityInRange[i][j] = m;
    }
}

// macrocell value ranges: macrocell m covers cells [m*kMacro, (m+1)*kMacro),
// i.e. voxels [m*kMacro, (m+1)*kMacro] (clamped to the grid)
inline void buildMacrocells(VolumeData &v)
{
  for (int a = 0; a < 3; a++)
    v.mdims[a] = (int)((v.dims[a] - 1 + kMacro - 1) / kMacro);
  const size_t n = (size_t)v.mdims[0] * v.mdims[1] * v.mdims[2];
  v.mmin.assign(n, INFINITY);
  v.mmax.assign(n, -INFINITY);
  for (int mz = 0; mz < v.mdims[2]; mz++)
    for (int my = 0; my < v.mdims[1]; my++)
      for (int mx = 0; mx < v.mdims[0]; mx++) {
        float lo = INFINITY, hi = -INFINITY;
        for (int64_t z = (int64_t)mz * kMacro; z <= std::min<int64_t>(((int64_t)mz + 1) * kMacro, v.dims[2] - 1); z++)
          for (int64_t y = (int64_t)my * kMacro; y <= std::min<int64_t>(((int64_t)my + 1) * kMacro, v.dims[1] - 1); y++)
            for (int64_t x = (int64_t)mx * kMacro; x <= std::min<int64_t>(((int64_t)mx + 1) * kMacro, v.dims[0] - 1);
                 x++) {
              const float s = v.voxels[(z * v.dims[1] + y) * v.dims[0] + x];
              lo = std::min(lo, s);
              hi = std::max(hi, s);
            }
        const size_t m = ((size_t)mz * v.mdims[1] + my) * v.mdims[0] + mx;
        v.mmin[m] = lo;
        v.mmax[m] = hi;
      }
}

// trilinear sample at p (volume space); NaN outside the grid
inline float sampleVolume(const VolumeData &v, vec3f p)
{
  const float lx = (p.x - v.origin.x) / v.spacing.x;
  const float ly = (p.y - v.origin.y) / v.spacing.y;
  const float lz = (p.z - v.origin.z) / v.spacing.z;
  if (!(lx >= 0.f && ly >= 0.f && lz >= 0.f && lx <= (float)(v.dims[0] - 1) && ly <= (float)(v.dims[1] - 1)
          && lz <= (float)(v.dims[2] - 1)))
    return NAN;
  const int64_t ix = std::min<int64_t>((int64_t)lx, v.dims[0] - 2);
  const int64_t iy = std::min<int64_t>((int64_t)ly, v.dims[1] - 2);
  const int64_t iz = std::min<int64_t>((int64_t)lz, v.dims[2] - 2);
  const float fx = lx - (float)ix, fy = ly - (float)iy, fz = lz - (float)iz;
  const float *b = v.voxels + (iz * v.dims[1] + iy) * v.dims[0] + ix;
  const int64_t sy = v.dims[0], sz = v.dims[0] * v.dims[1];
  const float c00 = lerpf(fx, b[0], b[1]);
  const float c10 = lerpf(fx, b[sy], b[sy + 1]);
  const float c01 = lerpf(fx, b[sz], b[sz + 1]);
  const float c11 = lerpf(fx, b[sz + sy], b[sz + sy + 1]);
  const float c0 = lerpf(fy, c00, c10);
  const float c1 = lerpf(fy, c01, c11);
  return lerpf(fz, c0, c1);
}

// LinearTransferFunction_get (LinearTransferFunction.ispc:11-43)
inline vec4f tfGet(const VolumeData &v, float value)
{
  value = (value - v.tfLower) * rcp(v.tfUpper - v.tfLower);
  const float nextBefore1 = 0.999999940395355224609375f; // 0x1.fffffep-1
  value = !(value > 0.f) ? 0.f : (value > nextBefore1 ? nextBefore1 : value); // NaN -> 0
  const int maxIdxC = (int)v.color.size() - 1;
  const float idxCf = value * (float)maxIdxC;
  const float fracC = idxCf - std::floor(idxCf);
  const int idxC = (int)idxCf;
  const vec3f c0 = v.color[idxC], c1 = v.color[std::min(maxIdxC, idxC + 1)];
  const vec3f col = v3(lerpf(fracC, c0.x, c1.x), lerpf(fracC, c0.y, c1.y), lerpf(fracC, c0.z, c1.z));
  const int maxIdxO = (int)v.opacity.size() - 1;
  const float idxOf = value * (float)maxIdxO;
  const float fracO = idxOf - std::floor(idxOf);
  const int idxO = (int)idxOf;
  const float op = lerpf(fracO, v.opacity[idxO], v.opacity[std::min(maxIdxO, idxO + 1)]);
  return {col.x, col.y, col.z, op};
}

// LinearTransferFunction_getMaxOpacity (LinearTransferFunction.ispc:45-58)
inline float tfMaxOpacity(const VolumeData &v, float lower, float upper)
{
  const int maxDim = kTfSubranges - 1;
  const float denom = v.tfUpper - v.tfLower;
  auto cl = [](float x) { return !(x > 0.f) ? 0.f : (x > 1.f ? 1.f : x); }; // NaN -> 0
  const int i = (int)std::floor(cl((lower - v.tfLower) / denom) * (float)maxDim);
  const int j = (int)std::ceil(cl((upper - v.tfLower) / denom) * (float)maxDim);
  return v.maxOpacityInRange[std::min(i, maxDim)][std::min(j, maxDim)];
}

inline float safeRcp(float d) // rcp_safe (rkcommon math.ih): |d| >= 1e-18
{
  const float ad = std::fabs(d);
  return 1.0f / (ad < 1e-18f ? std::copysign(1e-18f, d) : d);
}

// the ray's parameter range inside the volume's box [origin, upper], clipped
// to [t0, t1]; false if empty (intersectBox, common/Intersect.ih:195-222, and
// Volume_intersect_kernel, volume/Volume.ispc:22-75)
inline bool boxInterval(const VolumeData &v, vec3f o, vec3f d, float t0, float t1, float &lo, float &hi)
{
  const vec3f inv = v3(safeRcp(d.x), safeRcp(d.y), safeRcp(d.z));
  const vec3f mins = (v.origin - o) * inv, maxs = (v.upper - o) * inv;
  const vec3f nears = vmin(mins, maxs), fars = vmax(mins, maxs);
  const float entry = maxf(maxf(nears.x, nears.y), nears.z);
  const float exit = minf(minf(fars.x, fars.y), fars.z);
  lo = maxf(t0, entry);
  hi = minf(t1, exit);
  return lo < hi;
}

// the ray's intervals through the volume's macrocells within [t0, t1] (a
// range inside the volume's box); fn(ta, tb, lo, hi) returns false to stop.
// The macrocell of a point is clamped to the grid, so rounding at the box
// faces cannot lose the first or last interval.
template <typename F>
inline void forEachInterval(const VolumeData &v, vec3f o, vec3f d, float t0, float t1, F &&fn)
{
  if (!(t0 < t1))
    return;
  const float msz[3] = {v.spacing.x * kMacro, v.spacing.y * kMacro, v.spacing.z * kMacro};
  const float oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
  const float ii[3] = {safeRcp(d.x), safeRcp(d.y), safeRcp(d.z)};
  const float org[3] = {v.origin.x, v.origin.y, v.origin.z};
  int m[3], step[3];
  float tnext[3];
  for (int a = 0; a < 3; a++) {
    const float p = oo[a] + t0 * dd[a];
    const float cf = std::floor((p - org[a]) / msz[a]);
    const int c = !(cf > 0.f) ? 0 : (cf >= (float)v.mdims[a] ? v.mdims[a] - 1 : (int)cf);
    m[a] = c;
    step[a] = dd[a] > 0.f ? 1 : (dd[a] < 0.f ? -1 : 0);
    tnext[a] = step[a] == 0 ? INFINITY : (org[a] + (float)(c + (step[a] > 0 ? 1 : 0)) * msz[a] - oo[a]) * ii[a];
  }
  float t = t0;
  // every step enters a new macrocell: at most mdims[0]+mdims[1]+mdims[2]
  for (int guard = v.mdims[0] + v.mdims[1] + v.mdims[2]; t < t1 && guard > 0; guard--) {
    const int a = tnext[0] < tnext[1] ? (tnext[0] < tnext[2] ? 0 : 2) : (tnext[1] < tnext[2] ? 1 : 2);
    const float tb = minf(tnext[a], t1);
    if (tb > t) {
      const size_t mi = ((size_t)m[2] * v.mdims[1] + m[1]) * v.mdims[0] + m[0];
      if (!fn(t, tb, v.mmin[mi], v.mmax[mi]))
        return;
      t = tb;
    }
    m[a] += step[a];
    if (m[a] < 0 || m[a] >= v.mdims[a])
      return;
    tnext[a] = (org[a] + (float)(m[a] + (step[a] > 0 ? 1 : 0)) * msz[a] - oo[a]) * ii[a];
  }
}

// tentative collisions per track before delta tracking gives up (the track
// then counts as passed): only reached when densityScale times the track
// length is ~1e4 or more
static constexpr int kMaxDeltaSteps = 1 << 16;

// delta_tracking (VolumeSampler.ispc:13-92) over [t0, t1] (inside the box):
// the distance of the first real collision (inf if none), its extinction and
// albedo
// The optical depth to the next tentative collision, tau ~ Exp(1), is drawn
// once per track and after each tentative collision, and carried across the
// intervals: an interval of majorant mu and length l uses up mu * l of it.
// The reference draws afresh in every interval; the exponential is
// memoryless, so the collision distances have the same distribution.
inline float deltaTracking(const VolumeData &v, vec3f o, vec3f d, float t0, float t1, RandomSampler &rs, float &mu_t,
    vec3f &albedo)
{
  float result = INFINITY;
  if (!(t0 < t1))
    return result;
  float tau = -flog(1.f - RandomSampler_getFloat(rs));
  int steps = 0;
  forEachInterval(v, o, d, t0, t1, [&](float ta, float tb, float lo, float hi) {
    float t = ta;
    const float maxOpacity = tfMaxOpacity(v, lo, hi);
    const float mu_max = v.densityScale * maxOpacity;
    if (mu_max <= 0.f)
      return true;
    for (;;) {
      const float seg = mu_max * (tb - t);
      if (tau > seg) { // the track leaves the interval
        tau -= seg;
        return true;
      }
      t += tau / mu_max; // tentative collision
      const float xi = RandomSampler_getFloat(rs);
      const vec3f p = o + t * d;
      const float smp = sampleVolume(v, p);
      if (!std::isnan(smp)) {
        const vec4f color = tfGet(v, smp);
        mu_t = v.densityScale * color.w;
        if (xi < mu_t / mu_max) {
          albedo = v3(color.x, color.y, color.z);
          result = t;
          return false;
        }
      }
      if (++steps >= kMaxDeltaSteps)
        return false;
      tau = -flog(1.f - RandomSampler_getFloat(rs));
    }
  });
  return result;
}

// Henyey-Greenstein phase function (HenyeyGreenstein.ih:14-77)
inline float hgPhase(float cosTheta, float g)
{
  const float denom = 1.f + g * g - 2.f * g * cosTheta;
  return rcp(4.f * kPi) * (1.f - g * g) * rcp(denom * std::sqrt(denom));
}
inline bool hgIsSmooth(float g) { return !(std::fabs(g) > 1.f - 1e-3f); }

} // namespace orc
