// Volume sampling of the hip path tracer: structured-regular grids,
// piecewise-linear transfer function, macrocell majorants, delta tracking
// and the Henyey-Greenstein phase function, following
//   render/pathtracer/volumes/VolumeSampler.ispc:13-279 (delta_tracking,
//     volumeSampleFreePath, volumeTransmittance)
//   volume/transferFunction/LinearTransferFunction.ispc:11-60
//   render/pathtracer/volumes/HenyeyGreenstein.ih:14-77
//   math/random.ih:116-154 (the per-path PCG32 RandomSampler)
// The reference leaves grid sampling and interval iteration to OpenVKL; here
// a sample is the trilinear interpolation of the vertex-centred voxels
// (NaN outside the box) and the intervals of a ray are its runs through
// the macrocells of kMacro^3 cells, each with the macrocell's (min, max)
// voxel value as value range (its majorant precomputed per model).
#pragma once

#include "pt_shade.h"
#include "pt_types.h"

namespace hpt {

// ---- PCG32 RandomSampler (math/random.ih:116-154) ---------------------------------

struct RandomSampler
{
  uint64_t state;
  uint32_t stream;
};

HPT_HD uint32_t murmur3Mix(uint32_t hash, uint32_t k)
{
  k *= 0xcc9e2d51u;
  k = (k << 15) | (k >> 17);
  k *= 0x1b873593u;
  hash ^= k;
  return ((hash << 13) | (hash >> 19)) * 5u + 0xe6546b64u;
}

HPT_HD uint32_t murmur3Finalize(uint32_t h)
{
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}

HPT_HD uint32_t pcg32(RandomSampler &r)
{
  const uint64_t old = r.state;
  r.state = old * 6364136223846793005ULL + (uint64_t)(r.stream | 1u);
  const uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
  const uint32_t rot = (uint32_t)(old >> 59u);
  return (xorshifted >> rot) | (xorshifted << ((0u - rot) & 31u));
}

HPT_HD void randomSamplerInit(RandomSampler &r, uint32_t seed, uint32_t stream)
{
  r.stream = (stream << 1u) | 1u;
  r.state = murmur3Finalize(murmur3Mix(0u, seed));
  pcg32(r);
  r.state += seed;
  pcg32(r);
}

HPT_HD float randomFloat(RandomSampler &r) { return to_float_unorm(pcg32(r)); }

// ---- grid and transfer function ------------------------------------------------------

HPT_HD float lerpf(float t, float a, float b) { return (1.f - t) * a + t * b; }

HPT_HD float safeRcp(float d) // |d| >= 1e-18
{
  const float ad = fabsf(d);
  return 1.0f / (ad < 1e-18f ? copysignf(1e-18f, d) : d);
}

// trilinear sample at volume-space p; NaN outside the grid's box
HPT_HD float volumeSample(const DVolume &v, vec3f p)
{
  const float lx = (p.x - v.origin[0]) / v.spacing[0];
  const float ly = (p.y - v.origin[1]) / v.spacing[1];
  const float lz = (p.z - v.origin[2]) / v.spacing[2];
  if (!(lx >= 0.f && ly >= 0.f && lz >= 0.f && lx <= (float)(v.dims[0] - 1) && ly <= (float)(v.dims[1] - 1)
          && lz <= (float)(v.dims[2] - 1)))
    return __builtin_nanf("");
  const int ix = (int)lx < v.dims[0] - 2 ? (int)lx : v.dims[0] - 2;
  const int iy = (int)ly < v.dims[1] - 2 ? (int)ly : v.dims[1] - 2;
  const int iz = (int)lz < v.dims[2] - 2 ? (int)lz : v.dims[2] - 2;
  const float fx = lx - (float)ix, fy = ly - (float)iy, fz = lz - (float)iz;
  const size_t sy = (size_t)v.dims[0], sz = (size_t)v.dims[0] * (size_t)v.dims[1];
  const float *b = v.voxels + (size_t)iz * sz + (size_t)iy * sy + (size_t)ix;
  const float c00 = lerpf(fx, b[0], b[1]);
  const float c10 = lerpf(fx, b[sy], b[sy + 1]);
  const float c01 = lerpf(fx, b[sz], b[sz + 1]);
  const float c11 = lerpf(fx, b[sz + sy], b[sz + sy + 1]);
  const float c0 = lerpf(fy, c00, c10);
  const float c1 = lerpf(fy, c01, c11);
  return lerpf(fz, c0, c1);
}

// LinearTransferFunction_get (LinearTransferFunction.ispc:11-43), split in
// the opacity (every tentative collision) and the colour (real collisions)
HPT_HD float tfUnit(const DVolume &v, float value)
{
  value = (value - v.tfLower) * rcp(v.tfUpper - v.tfLower);
  const float nextBefore1 = 0.999999940395355224609375f; // 0x1.fffffep-1
  return !(value > 0.f) ? 0.f : (value > nextBefore1 ? nextBefore1 : value); // NaN -> 0
}

HPT_HD float tfOpacity(const DVolume &v, float u)
{
  const int maxIdx = v.numOpacities - 1;
  const float idxf = u * (float)maxIdx;
  const float frac = idxf - floorf(idxf);
  const int idx = (int)idxf;
  return lerpf(frac, v.tfOpacity[idx], v.tfOpacity[idx + 1 < maxIdx ? idx + 1 : maxIdx]);
}

HPT_HD vec3f tfColor(const DVolume &v, float u)
{
  const int maxIdx = v.numColors - 1;
  const float idxf = u * (float)maxIdx;
  const float frac = idxf - floorf(idxf);
  const int idx = (int)idxf;
  const int idx1 = idx + 1 < maxIdx ? idx + 1 : maxIdx;
  const float *c0 = v.tfColor + 3 * idx, *c1 = v.tfColor + 3 * idx1;
  return v3(lerpf(frac, c0[0], c1[0]), lerpf(frac, c0[1], c1[1]), lerpf(frac, c0[2], c1[2]));
}

// LinearTransferFunction_getMaxOpacity (LinearTransferFunction.ispc:45-58)
// over a kTfSubranges^2 table (used on the host to build the majorant grid)
HPT_HD float tfMaxOpacity(const float *maxOpacity, float tfLower, float tfUpper, float lower, float upper)
{
  const int maxDim = kTfSubranges - 1;
  const float denom = tfUpper - tfLower;
  const float a = (lower - tfLower) / denom, b = (upper - tfLower) / denom;
  const int i = (int)floorf((!(a > 0.f) ? 0.f : (a > 1.f ? 1.f : a)) * (float)maxDim); // NaN -> 0
  const int j = (int)ceilf((!(b > 0.f) ? 0.f : (b > 1.f ? 1.f : b)) * (float)maxDim);
  return maxOpacity[(i < maxDim ? i : maxDim) * kTfSubranges + (j < maxDim ? j : maxDim)];
}

// the ray's parameter range inside the volume's box, clipped to [t0, t1]
// (intersectBox, common/Intersect.ih:195-222; Volume_intersect_kernel,
// volume/Volume.ispc:22-75)
HPT_HD bool volumeBox(const DVolume &v, vec3f o, vec3f d, float t0, float t1, float &lo, float &hi)
{
  const vec3f inv = v3(safeRcp(d.x), safeRcp(d.y), safeRcp(d.z));
  const vec3f mins = (ld3(v.origin) - o) * inv, maxs = (ld3(v.upper) - o) * inv;
  const float nx = minf(mins.x, maxs.x), ny = minf(mins.y, maxs.y), nz = minf(mins.z, maxs.z);
  const float fx = maxf(mins.x, maxs.x), fy = maxf(mins.y, maxs.y), fz = maxf(mins.z, maxs.z);
  const float entry = maxf(maxf(nx, ny), nz);
  const float exit = minf(minf(fx, fy), fz);
  lo = maxf(t0, entry);
  hi = minf(t1, exit);
  return lo < hi;
}

// tentative collisions per track before delta tracking gives up (the
// track then counts as passed)
constexpr int kMaxDeltaSteps = 1 << 16;

// work counters of delta tracking: voxel samples (tentative collisions),
// macrocell intervals entered, tracked segments
struct VolCount
{
  uint32_t samples = 0, cells = 0, tracks = 0;
};

// delta_tracking (VolumeSampler.ispc:13-92) over [t0, t1] (inside the box),
// walking the macrocells with a 3D DDA: the distance of the first real
// collision (inf if none) and its albedo.  The optical depth to the next
// tentative collision, tau ~ Exp(1), is drawn once per track and after each
// tentative collision, and carried across macrocell intervals: an interval
// of majorant mu and length l uses up mu * l of it.  The exponential is
// memoryless, so the collision distances have the distribution of a fresh
// draw per interval, with one logarithm per collision instead of one per
// interval.
HPT_HD float deltaTracking(
    const DVolume &v, vec3f o, vec3f d, float t0, float t1, RandomSampler &rs, vec3f &albedo, VolCount &cnt)
{
  if (!(t0 < t1))
    return kInf;
  cnt.tracks++;
  const float oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
  const float ii[3] = {safeRcp(d.x), safeRcp(d.y), safeRcp(d.z)};
  // the grid fields the macrocell steps read, in registers: each step then
  // waits on one load (its majorant) rather than also on the fields it
  // indexes by the step's axis; the tentative collisions read the fields
  // of the copy `vc`, loaded once per track
  const DVolume vc = v;
  const float *majorant = v.majorant;
  const int md[3] = {v.mdims[0], v.mdims[1], v.mdims[2]};
  const float gorg[3] = {v.origin[0], v.origin[1], v.origin[2]};
  float msz[3];
  int m[3], step[3];
  float tnext[3];
  for (int a = 0; a < 3; a++) {
    msz[a] = v.spacing[a] * (float)kMacro;
    const float p = oo[a] + t0 * dd[a];
    const float cf = floorf((p - gorg[a]) / msz[a]);
    const int c = !(cf > 0.f) ? 0 : (cf >= (float)md[a] ? md[a] - 1 : (int)cf);
    m[a] = c;
    step[a] = dd[a] > 0.f ? 1 : (dd[a] < 0.f ? -1 : 0);
    tnext[a] = step[a] == 0 ? kInf : (gorg[a] + (float)(c + (step[a] > 0 ? 1 : 0)) * msz[a] - oo[a]) * ii[a];
  }
  float t = t0;
  float tau = -flog(1.f - randomFloat(rs));
  int steps = 0;
  for (int guard = md[0] + md[1] + md[2]; t < t1 && guard > 0; guard--) {
    const int a = tnext[0] < tnext[1] ? (tnext[0] < tnext[2] ? 0 : 2) : (tnext[1] < tnext[2] ? 1 : 2);
    const float tb = minf(tnext[a], t1);
    if (tb > t) {
      // one interval [t, tb] of macrocell m
      const size_t mi = ((size_t)m[2] * (size_t)md[1] + (size_t)m[1]) * (size_t)md[0] + (size_t)m[0];
      cnt.cells++;
      const float mu_max = majorant[mi];
      if (mu_max > 0.f) {
        float tt = t;
        for (;;) {
          const float seg = mu_max * (tb - tt);
          if (tau > seg) { // the track leaves the interval
            tau -= seg;
            break;
          }
          tt += tau / mu_max; // tentative collision
          const float xi = randomFloat(rs);
          cnt.samples++;
          const vec3f p = o + tt * d;
          const float smp = volumeSample(vc, p);
          if (smp == smp) {
            const float u = tfUnit(vc, smp);
            const float mu_t = vc.densityScale * tfOpacity(vc, u);
            if (xi < mu_t / mu_max) {
              albedo = tfColor(vc, u);
              return tt;
            }
          }
          if (++steps >= kMaxDeltaSteps)
            return kInf;
          tau = -flog(1.f - randomFloat(rs));
        }
      }
      t = tb;
    }
    m[a] += step[a];
    if (m[a] < 0 || m[a] >= md[a])
      break;
    tnext[a] = (gorg[a] + (float)(m[a] + (step[a] > 0 ? 1 : 0)) * msz[a] - oo[a]) * ii[a];
  }
  return kInf;
}

HPT_HD void toVolumeSpace(const DInstance &I, vec3f worg, vec3f wdir, vec3f &o, vec3f &d)
{
  const linear3f rl = {{I.rcp[0], I.rcp[1], I.rcp[2]}, {I.rcp[3], I.rcp[4], I.rcp[5]}, {I.rcp[6], I.rcp[7], I.rcp[8]}};
  o = mul(rl, worg) + v3(I.rcp[9], I.rcp[10], I.rcp[11]);
  d = mul(rl, wdir);
}

// volumeSampleFreePath (VolumeSampler.ispc:96-199) for a ray [t0, tTrace]
// whose surface hit (if any) is at tTrace; [r0, r1] is the clipping
// pass's ray range.  The nearest collision wins and clips the later
// intervals.  Returns its distance (inf: none), pair index and albedo.
HPT_HD float volumeFreePath(const DScene &S, vec3f org, vec3f dir, float t0, float tTrace, bool surfaceHit, float r0,
    float r1, RandomSampler &rs, int &pair, vec3f &albedo, VolCount &cnt)
{
  float minDist = kInf;
  float tClip = tTrace;
  bool clipped = surfaceHit;
  pair = -1;
  for (int j = 0; j < S.numVolInsts; j++) {
    const DVolInst vi = S.volInsts[j];
    const DVolume &v = S.volumes[vi.vol];
    vec3f o, d;
    toVolumeSpace(S.instances[vi.inst], org, dir, o, d);
    float lo, hi;
    if (!volumeBox(v, o, d, t0, tTrace, lo, hi))
      continue;
    lo = maxf(0.f, lo);
    if (clipped)
      hi = minf(hi, tClip);
    if (hi < lo)
      continue;
    if (hi < r0 || r1 < lo)
      continue;
    vec3f alb;
    const float dist = deltaTracking(v, o, d, maxf(r0, lo), minf(r1, hi), rs, alb, cnt);
    if (dist < kInf && dist < minDist) {
      minDist = dist;
      pair = j;
      albedo = alb;
      tClip = dist;
      clipped = true;
    }
  }
  return minDist;
}

// volumeTransmittance (VolumeSampler.ispc:201-277): 1 if delta tracking
// passes every volume interval of [t0, t1] (within [r0, r1]), else 0
HPT_HD float volumeTransmittance(
    const DScene &S, vec3f org, vec3f dir, float t0, float t1, float r0, float r1, RandomSampler &rs, VolCount &cnt)
{
  float T = 1.f;
  for (int j = 0; j < S.numVolInsts && T > 0.f; j++) {
    const DVolInst vi = S.volInsts[j];
    const DVolume &v = S.volumes[vi.vol];
    vec3f o, d;
    toVolumeSpace(S.instances[vi.inst], org, dir, o, d);
    float lo, hi;
    if (!volumeBox(v, o, d, t0, t1, lo, hi))
      continue;
    if (hi < r0 || r1 < lo)
      continue;
    const float b = minf(r1, hi);
    vec3f alb;
    const float dist = deltaTracking(v, o, d, maxf(r0, lo), b, rs, alb, cnt);
    T = T * (dist < b ? 0.f : 1.f);
  }
  return T;
}

// ---- Henyey-Greenstein (volumes/HenyeyGreenstein.ih:14-77) -------------------------

HPT_HD float hgPhase(float cosTheta, float g)
{
  const float denom = 1.f + g * g - 2.f * g * cosTheta;
  return rcp(4.f * kPi) * (1.f - g * g) * rcp(denom * sqrtf(denom));
}

HPT_HD bool hgIsSmooth(float g) { return !(fabsf(g) > 1.f - 1e-3f); }

// the eval's pdf is the phase function whatever the anisotropy
HPT_HD EvalRes hgEval(float g, vec3f wo, vec3f wi)
{
  const float p = hgPhase(dot(neg(wo), wi), g);
  EvalRes r;
  r.value = v3(p);
  r.pdf = p;
  return r;
}

HPT_HD SampleRes hgSample(float g, vec3f wo, vec2f s)
{
  SampleRes res;
  res.weight = v3(1.f);
  if (!hgIsSmooth(g)) { // a delta function
    res.wi = g > 0 ? neg(wo) : wo;
    res.pdf = kInf;
    res.type = g > 0 ? SPEC_TRANS : SPEC_REFL;
    return res;
  }
  float cosTheta;
  if (fabsf(g) < 1e-3f)
    cosTheta = 1.f - 2.f * s.x;
  else {
    const float sqrTerm = (1.f - g * g) * rcp(1.f - g + 2.f * g * s.x);
    cosTheta = (1.f + g * g - sqrTerm * sqrTerm) * rcp(2.f * g);
  }
  res.type = S_SMOOTH;
  const float sinTheta = sqrtf(maxf(0.f, 1.f - cosTheta * cosTheta));
  const float phi = kTwoPi * s.y;
  res.wi = mul(frame(neg(wo)), cartesian(phi, sinTheta, cosTheta));
  res.pdf = hgPhase(cosTheta, g);
  return res;
}

} // namespace hpt
