// BVH traversal and ray/primitive tests of the hip path tracer (device).
//
// The reference hands these to Embree (rtcIntersectV,
// modules/cpu/common/World.ih:29-52, 247-286).  Here:
//   * two-level 4-wide BVH (TLAS over instances, one BLAS per group) in
//     128-byte nodes holding four child boxes; one lane per ray, children
//     visited nearest first, speculative while-while order (Aila and Laine,
//     "Understanding the efficiency of ray traversal on GPUs", 2009): a
//     lane holds a leaf it reaches and goes on with inner nodes until every
//     lane of the wave holds one;
//   * the traversal stack lives in LDS (LDS_DEPTH entries per lane, laid out
//     [depth][lane] so lanes hit distinct banks) and spills to a per-thread
//     global overflow row beyond that;
//   * triangles: watertight test (Woop, Benthin, Wald 2013) in float: an
//     edge function depends on the edge's two vertices only and is exactly
//     antisymmetric between the two triangles sharing the edge, so a ray
//     never passes between them; quads as two triangles (v0,v1,v3) and
//     (v2,v3,v1); spheres: Intersect.ih:29-80;
//   * closest hit = smallest t in (t0, tfar]; exact ties resolved by the
//     smallest (instance, model, primitive, half), so results do not depend
//     on BVH layout;
//   * near-ties of extension rays (TIES): when another hit's float t lies
//     within kTieEps (relative) of the closest one, a second traversal over
//     that window re-resolves the hit in FP64: every hit in the window gets
//     its distance recomputed in double precision from the world ray
//     (object ray formed in double from the instance's inverse transform)
//     and the float vertices, and the smallest (FP64 t, key) wins.  The
//     flag that triggers it is independent of traversal order: a hit is
//     compared with the current closest on arrival, hits up to kTieEps
//     beyond it are still tested, and the flag clears only when a hit
//     closer than the window's lower end arrives;
//   * node boxes only cull, conservatively (far plane scaled by 1+4e-7).
#pragma once

#include "pt_math.h"
#include "pt_types.h"

namespace hpt {

constexpr int kTraceBlock = 128;
constexpr int kLdsDepth = 24;
constexpr int kOverflowDepth = 104;
// transparent hits an occlusion traversal records for the transparency pass
constexpr int kTransRecHits = 4;
// words of scratch per resident traversal thread: the stack overflow, then
// kTransRecHits recorded hits of 4 words (one float4, 16-byte aligned),
// then their count (padded to 4 words)
constexpr int kThreadRow = kOverflowDepth + 4 * kTransRecHits + 4;
constexpr int32_t kMarker = 0x7fffffff;

struct Hit
{
  float t, u, v;
  vec3f Ng;
  uint32_t inst, model, prim, half;
  bool hit;
};

HPT_HD bool keyLess(uint32_t i0, uint32_t m0, uint32_t p0, uint32_t h0, const Hit &b)
{
  if (i0 != b.inst)
    return i0 < b.inst;
  if (m0 != b.model)
    return m0 < b.model;
  if (p0 != b.prim)
    return p0 < b.prim;
  return h0 < b.half;
}

HPT_HD bool better(float t, uint32_t i, uint32_t m, uint32_t p, uint32_t h, const Hit &b)
{
  return !b.hit || t < b.t || (t == b.t && keyLess(i, m, p, h, b));
}

HPT_HD float safeInv(float d)
{
  const float ad = fabsf(d);
  return 1.0f / (ad < 1e-18f ? copysignf(1e-18f, d) : d);
}

struct TriPre
{
  int kx, ky, kz;
  float Sx, Sy, Sz;
};

HPT_HD TriPre triPrecompute(vec3f dir)
{
  TriPre p;
  const vec3f ad = vabs(dir);
  p.kz = ad.x > ad.y ? (ad.x > ad.z ? 0 : 2) : (ad.y > ad.z ? 1 : 2);
  p.kx = p.kz == 2 ? 0 : p.kz + 1;
  p.ky = p.kx == 2 ? 0 : p.kx + 1;
  if (comp(dir, p.kz) < 0.0f) {
    const int t = p.kx;
    p.kx = p.ky;
    p.ky = t;
  }
  const float dz = comp(dir, p.kz);
  p.Sx = comp(dir, p.kx) / dz;
  p.Sy = comp(dir, p.ky) / dz;
  p.Sz = 1.0f / dz;
  return p;
}

HPT_HD bool intersectTriangle(vec3f org, float t0, const TriPre &pc, vec3f v0, vec3f v1, vec3f v2, float tmax,
    float &t, float &u, float &v)
{
  const vec3f A = v0 - org;
  const vec3f B = v1 - org;
  const vec3f C = v2 - org;
  const float Az_ = comp(A, pc.kz), Bz_ = comp(B, pc.kz), Cz_ = comp(C, pc.kz);
  const float Ax = comp(A, pc.kx) - pc.Sx * Az_;
  const float Ay = comp(A, pc.ky) - pc.Sy * Az_;
  const float Bx = comp(B, pc.kx) - pc.Sx * Bz_;
  const float By = comp(B, pc.ky) - pc.Sy * Bz_;
  const float Cx = comp(C, pc.kx) - pc.Sx * Cz_;
  const float Cy = comp(C, pc.ky) - pc.Sy * Cz_;
  const float U = Cx * By - Cy * Bx;
  const float V = Ax * Cy - Ay * Cx;
  const float W = Bx * Ay - By * Ax;
  if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f))
    return false;
  const float det = U + V + W;
  if (det == 0.0f)
    return false;
  const float Az = pc.Sz * Az_;
  const float Bz = pc.Sz * Bz_;
  const float Cz = pc.Sz * Cz_;
  const float T = U * Az + V * Bz + W * Cz;
  const float rcpDet = 1.0f / det;
  const float tt = T * rcpDet;
  if (!(tt > t0) || !(tt <= tmax))
    return false;
  t = tt;
  u = V * rcpDet;
  v = W * rcpDet;
  return true;
}

// both roots of the ray/sphere test, with the arithmetic of
// intersectSphere: bit 0 = entry (tE, NE) in (t0, tmax], bit 1 = exit
HPT_HD int sphereRoots(vec3f org, vec3f d, float t0, vec3f center, float radius, float tmax, float &tE, vec3f &NE,
    float &tX, vec3f &NX)
{
  const float rd2 = 1.0f / dot(d, d);
  const vec3f CO = center - org;
  const float projCO = dot(CO, d) * rd2;
  const vec3f perp = CO - projCO * d;
  const float l2 = dot(perp, perp);
  const float r2 = radius * radius;
  if (l2 > r2)
    return 0;
  const float td = sqrtf((r2 - l2) * rd2);
  tE = projCO - td;
  tX = projCO + td;
  NE = (-td) * d - perp;
  NX = td * d - perp;
  return (tE > t0 && tE <= tmax ? 1 : 0) | (tX > t0 && tX <= tmax ? 2 : 0);
}

HPT_HD bool intersectSphere(vec3f org, vec3f d, float t0, vec3f center, float radius, float tmax, float &t, vec3f &N)
{
  const float rd2 = 1.0f / dot(d, d);
  const vec3f CO = center - org;
  const float projCO = dot(CO, d) * rd2;
  const vec3f perp = CO - projCO * d;
  const float l2 = dot(perp, perp);
  const float r2 = radius * radius;
  if (l2 > r2)
    return false;
  const float td = sqrtf((r2 - l2) * rd2);
  const float tEntry = projCO - td;
  const float tExit = projCO + td;
  if (tEntry > t0 && tEntry <= tmax) {
    t = tEntry;
    N = (-td) * d - perp;
    return true;
  }
  if (tExit > t0 && tExit <= tmax) {
    t = tExit;
    N = td * d - perp;
    return true;
  }
  return false;
}

struct TraceStats
{
  uint32_t nodes, prims;
};

// relative window of float distances within which two extension-ray hits
// count as a near-tie (64 ulp at 1.0 is 7.6e-6; the float t of the
// watertight test is good to a few ulp away from grazing incidence)
constexpr float kTieEps = 1.0e-5f;
// lower end of the boxes the FP64 re-resolution opens, relative to the
// float closest hit
constexpr float kTieBoxLow = 0.9f;

HPT_HD bool nearTie(float t, float tBest) { return fabsf(t - tBest) <= kTieEps * tBest; }

// FP64 distance along the world ray (worg + t wdir) to the plane of
// triangle (a, b, c) of instance rcp (inverse transform, columns vx, vy,
// vz, p): the object ray is formed in double from the float world ray and
// the float transform, the plane from the float vertices
HPT_HD void fp64ObjectRay(const float *rcp, vec3f worg, vec3f wdir, double o[3], double d[3])
{
  const double wo[3] = {worg.x, worg.y, worg.z}, wd[3] = {wdir.x, wdir.y, wdir.z};
  for (int r = 0; r < 3; r++) {
    o[r] = (double)rcp[r] * wo[0] + (double)rcp[3 + r] * wo[1] + (double)rcp[6 + r] * wo[2] + (double)rcp[9 + r];
    d[r] = (double)rcp[r] * wd[0] + (double)rcp[3 + r] * wd[1] + (double)rcp[6 + r] * wd[2];
  }
}

HPT_HD double fp64TriangleT(const double o[3], const double d[3], vec3f a, vec3f b, vec3f c)
{
  const double e1[3] = {(double)b.x - a.x, (double)b.y - a.y, (double)b.z - a.z};
  const double e2[3] = {(double)c.x - a.x, (double)c.y - a.y, (double)c.z - a.z};
  const double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
  const double ao[3] = {(double)a.x - o[0], (double)a.y - o[1], (double)a.z - o[2]};
  const double num = ao[0] * n[0] + ao[1] * n[1] + ao[2] * n[2];
  const double den = d[0] * n[0] + d[1] * n[1] + d[2] * n[2];
  return num / den;
}

// FP64 entry (exitRoot 0) or exit (1) distance of a sphere
HPT_HD double fp64SphereT(const double o[3], const double d[3], vec3f center, float radius, int exitRoot)
{
  const double oc[3] = {o[0] - center.x, o[1] - center.y, o[2] - center.z};
  const double A = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  const double B = oc[0] * d[0] + oc[1] * d[1] + oc[2] * d[2];
  const double C = oc[0] * oc[0] + oc[1] * oc[1] + oc[2] * oc[2] - (double)radius * (double)radius;
  const double disc = B * B - A * C;
  const double sq = disc > 0.0 ? sqrt(disc) : 0.0;
  return (exitRoot ? (-B + sq) : (-B - sq)) / A;
}

// one 4-wide node: child boxes as per-axis float4s (lane c = child c)
struct Node4
{
  float4 lox, hix, loy, hiy, loz, hiz;
  int4 child;
};

__device__ __forceinline__ Node4 loadNode(const DNode4 *nodes, int idx)
{
  const float4 *np = reinterpret_cast<const float4 *>(nodes + idx);
  Node4 n;
  n.lox = np[0];
  n.hix = np[1];
  n.loy = np[2];
  n.hiy = np[3];
  n.loz = np[4];
  n.hiz = np[5];
  n.child = reinterpret_cast<const int4 *>(np)[6];
  return n;
}

// order (t, child) pairs by t
__device__ __forceinline__ void cswap(float &ta, int &ca, float &tb, int &cb)
{
  const bool sw = tb < ta;
  const float t = sw ? tb : ta;
  const int c = sw ? cb : ca;
  tb = sw ? ta : tb;
  cb = sw ? ca : cb;
  ta = t;
  ca = c;
}

__device__ __forceinline__ float slabNear(float lox, float loy, float loz, float hix, float hiy, float hiz,
    vec3f o, vec3f inv, float t0, float tb)
{
  const float x0 = (lox - o.x) * inv.x, x1 = (hix - o.x) * inv.x;
  const float y0 = (loy - o.y) * inv.y, y1 = (hiy - o.y) * inv.y;
  const float z0 = (loz - o.z) * inv.z, z1 = (hiz - o.z) * inv.z;
  const float tn = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), t0));
  const float tf = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), tb)) * 1.0000004f;
  return tn <= tf ? tn : kInf;
}

// Traversal modes: MODE_CLOSEST returns the closest hit; MODE_OCCLUSION
// returns as soon as a primitive of an opaque model is hit and otherwise
// reports in `sawTransparent` whether any (non-opaque) primitive was hit.
// With RECORD, an occlusion traversal also records every hit of a
// non-opaque primitive in (t0, tfar] (both roots of a sphere) in the
// thread's scratch row after the stack overflow (recordedHits): one float4
// (t, instance, index into S.prims, root: 1 = a sphere's exit) per hit, up
// to kTransRecHits, and in the word after them the count of all of them,
// also those beyond the capacity (0 when the traversal starts: the caller
// resets it after each ray that recorded hits; it lives in memory rather
// than a register since transparent hits are rare).
// hitFromRecord rebuilds the closest-hit record of one of them.

__device__ __forceinline__ float4 *recordedHits(uint32_t *ovf)
{
  return reinterpret_cast<float4 *>(ovf + kOverflowDepth);
}

__device__ __forceinline__ uint32_t *recordedCount(uint32_t *ovf)
{
  return ovf + kOverflowDepth + 4 * kTransRecHits;
}

__device__ __forceinline__ void recordHit(uint32_t *ovf, float t, uint32_t inst, uint32_t primIndex, uint32_t exitRoot)
{
  uint32_t *cnt = recordedCount(ovf);
  const uint32_t c = *cnt;
  if (c < (uint32_t)kTransRecHits)
    recordedHits(ovf)[c] = make_float4(t, __uint_as_float(inst), __uint_as_float(primIndex), __uint_as_float(exitRoot));
  *cnt = c + 1;
}
enum TraceMode
{
  MODE_CLOSEST = 0,
  MODE_OCCLUSION = 1
};

// Second traversal of a ray whose closest hit `best` has a near-tie: every
// hit whose float t is a near-tie of best.t gets its FP64 distance, and the
// smallest (FP64 t, key) becomes `best` (with its float hit record).  A
// plain stack traversal: the result is a minimum over a set, so the visit
// order does not matter.  Boxes are opened over [tLo, tHi]: every hit of
// the set has a float t within 1e-5 of tF, and tLo = 0.9 tF leaves a
// margin far above the float error of a hit distance, so only the boxes
// around the hit point are visited (the descent to it, not every box the
// ray crosses before it).
template <bool STATS>
__device__ void resolveNearTie(const DScene &S, vec3f worg, vec3f wdir, float t0, float tfar, Hit &best,
    uint32_t *lds, uint32_t *ovf, TraceStats &st)
{
  const int lane = threadIdx.x;
  const float tF = best.t;
  const float tHi = fminf(tF + kTieEps * tF, tfar);
  const float tLo = fmaxf(t0, kTieBoxLow * tF);
  const vec3f winv = v3(safeInv(wdir.x), safeInv(wdir.y), safeInv(wdir.z));
  Hit w = best;
  double wD = 0.0;
  bool have = false;
  int sp = 0;
  auto push = [&](int v) {
    if (sp < kLdsDepth)
      lds[sp * kTraceBlock + lane] = (uint32_t)v;
    else
      ovf[sp - kLdsDepth] = (uint32_t)v;
    sp++;
  };
  auto pop = [&]() -> int {
    sp--;
    if (sp < kLdsDepth)
      return (int)lds[sp * kTraceBlock + lane];
    return (int)ovf[sp - kLdsDepth];
  };
  // children of node `cur` whose boxes the ray enters in (t0, tHi]
  auto expand = [&](int cur, vec3f o, vec3f inv) -> bool {
    const Node4 nd = loadNode(S.nodes, cur);
    if (STATS)
      st.nodes++;
    const float tc[4] = {slabNear(nd.lox.x, nd.loy.x, nd.loz.x, nd.hix.x, nd.hiy.x, nd.hiz.x, o, inv, tLo, tHi),
        slabNear(nd.lox.y, nd.loy.y, nd.loz.y, nd.hix.y, nd.hiy.y, nd.hiz.y, o, inv, tLo, tHi),
        slabNear(nd.lox.z, nd.loy.z, nd.loz.z, nd.hix.z, nd.hiy.z, nd.hiz.z, o, inv, tLo, tHi),
        slabNear(nd.lox.w, nd.loy.w, nd.loz.w, nd.hix.w, nd.hiy.w, nd.hiz.w, o, inv, tLo, tHi)};
    const int ch[4] = {nd.child.x, nd.child.y, nd.child.z, nd.child.w};
    for (int c = 0; c < 4; c++)
      if (tc[c] != kInf) {
        if (sp >= kLdsDepth + kOverflowDepth) // cannot happen: the world commit checks the stack need
          return false;
        push(ch[c]);
      }
    return true;
  };
  push(S.tlasRoot);
  while (sp > 0) {
    const int cur = pop();
    if (cur >= 0) {
      if (!expand(cur, worg, winv))
        return;
      continue;
    }
    const uint32_t inst = ~(uint32_t)cur >> 3; // world leaf: one instance
    const DInstance &I = S.instances[inst];
    const linear3f rl = {{I.rcp[0], I.rcp[1], I.rcp[2]}, {I.rcp[3], I.rcp[4], I.rcp[5]}, {I.rcp[6], I.rcp[7], I.rcp[8]}};
    const vec3f o = mul(rl, worg) + v3(I.rcp[9], I.rcp[10], I.rcp[11]);
    const vec3f d = mul(rl, wdir);
    const vec3f inv = v3(safeInv(d.x), safeInv(d.y), safeInv(d.z));
    const TriPre pc = triPrecompute(d);
    double od[3], dd[3];
    fp64ObjectRay(I.rcp, worg, wdir, od, dd);
    const int base = sp;
    push(I.blasRoot);
    while (sp > base) {
      const int c = pop();
      if (c >= 0) {
        if (!expand(c, o, inv))
          return;
        continue;
      }
      const uint32_t enc = ~(uint32_t)c;
      const uint32_t first = enc >> 3, count = (enc & 7u) + 1u;
      for (uint32_t k = 0; k < count; k++) {
        const DPrim &P = S.prims[first + k];
        const float4 pa = reinterpret_cast<const float4 *>(&P)[0];
        const float4 pb = reinterpret_cast<const float4 *>(&P)[1];
        const float4 pcv = reinterpret_cast<const float4 *>(&P)[2];
        const uint32_t primID = __float_as_uint(pa.w);
        const uint32_t mk = __float_as_uint(pb.w);
        const uint32_t model = mk & 0x0fffffffu, kind = mk >> 28;
        if (STATS)
          st.prims++;
        float t, u = 0.f, v = 0.f;
        vec3f Ng;
        uint32_t half = 0;
        double tD;
        if (kind == PRIM_SPHERE) {
          float tE, tX;
          vec3f NE, NX;
          const int roots = sphereRoots(o, d, t0, v3(pa.x, pa.y, pa.z), pb.x, tHi, tE, NE, tX, NX);
          if (!roots)
            continue;
          const int exitRoot = (roots & 1) ? 0 : 1;
          t = exitRoot ? tX : tE;
          Ng = exitRoot ? NX : NE;
          if (!nearTie(t, tF))
            continue;
          tD = fp64SphereT(od, dd, v3(pa.x, pa.y, pa.z), pb.x, exitRoot);
        } else {
          const vec3f A = v3(pa.x, pa.y, pa.z), B = v3(pb.x, pb.y, pb.z), C = v3(pcv.x, pcv.y, pcv.z);
          if (!intersectTriangle(o, t0, pc, A, B, C, tHi, t, u, v) || !nearTie(t, tF))
            continue;
          Ng = cross(B - A, C - A);
          if (kind == PRIM_QUAD1) {
            half = 1;
            u = 1.0f - u;
            v = 1.0f - v;
          }
          tD = fp64TriangleT(od, dd, A, B, C);
        }
        if (!have || tD < wD || (tD == wD && keyLess(inst, model, primID, half, w))) {
          have = true;
          wD = tD;
          w.hit = true;
          w.t = t;
          w.u = u;
          w.v = v;
          w.Ng = Ng;
          w.inst = inst;
          w.model = model;
          w.prim = primID;
          w.half = half;
        }
      }
    }
  }
  if (have)
    best = w;
}

// near-tie handling of a closest-hit traversal (see the header comment):
// off; re-resolved in FP64 at the end of the traversal; or only reported
// (in sawTransparent, unused by closest-hit traversals otherwise) for a
// later pass to re-resolve with resolveNearTie, which keeps the FP64 code
// and its registers out of the traversal kernel
constexpr int TIES_OFF = 0, TIES_RESOLVE = 1, TIES_REPORT = 2;

// Per-lane state of one traversal.  A traversal runs as rounds of
// traceDescend (inner nodes, speculative while-while) and traceLeaf (the
// leaf the lane holds); traceRay runs one ray to its end, traceQueue (in
// the kernels) keeps the lanes of a wave busy with rays from a queue.
struct TraceState
{
  vec3f worg, wdir, winv; // the world ray
  vec3f o, d, inv;        // the current (world or object) ray
  TriPre pc;
  float t0, tfar;
  int inst, modelOffset, sp, cur, post;
  bool tie;            // a near-tie of the current closest hit has been seen (TIES != TIES_OFF)
  bool sawTransparent; // MODE_OCCLUSION: a non-opaque primitive was hit
};

// `cur` of a lane whose traversal has no node left
constexpr int kTraceDone = 0x7ffffffe;

__device__ __forceinline__ void traceBegin(
    TraceState &s, const DScene &S, vec3f worg, vec3f wdir, float t0, float tfar, Hit &best)
{
  best.hit = false;
  best.t = tfar;
  best.u = best.v = 0.f;
  best.Ng = v3(0.f);
  best.inst = best.model = best.prim = kInvalidID;
  best.half = 0;
  s.worg = worg;
  s.wdir = wdir;
  s.winv = v3(safeInv(wdir.x), safeInv(wdir.y), safeInv(wdir.z));
  s.o = worg;
  s.d = wdir;
  s.inv = s.winv;
  s.pc = {0, 1, 2, 0.f, 0.f, 0.f};
  s.t0 = t0;
  s.tfar = tfar;
  s.inst = -1;
  s.modelOffset = 0;
  s.sp = 0;
  s.cur = S.tlasRoot;
  s.post = 0;
  s.tie = false;
  s.sawTransparent = false;
}

// callers check sp < kLdsDepth + kOverflowDepth before pushing
__device__ __forceinline__ void tracePush(TraceState &s, uint32_t *lds, uint32_t *ovf, int v)
{
  if (s.sp < kLdsDepth)
    lds[s.sp * kTraceBlock + threadIdx.x] = (uint32_t)v;
  else
    ovf[s.sp - kLdsDepth] = (uint32_t)v;
  s.sp++;
}

// the LDS read is unconditional (a clamped slot), the overflow row is read
// only by the rare lanes whose stack is deeper than kLdsDepth (a volatile
// read, so the compiler keeps the two reads apart instead of merging them
// into one generic-address load)
__device__ __forceinline__ int tracePop(TraceState &s, const uint32_t *lds, const uint32_t *ovf)
{
  s.sp--;
  int v = (int)lds[(s.sp < kLdsDepth ? s.sp : kLdsDepth - 1) * kTraceBlock + threadIdx.x];
  if (s.sp >= kLdsDepth)
    v = (int)*reinterpret_cast<const volatile uint32_t *>(ovf + (s.sp - kLdsDepth));
  return v;
}

// Inner nodes of the speculative while-while traversal: a lane that reaches
// a leaf keeps it aside (post) and goes on with inner nodes until every
// lane of the wave holds a leaf or has no node left.  A held BLAS leaf
// stops the lane at the instance's exit marker (it needs the object ray); a
// held instance leaf is entered from the world level, with the next world
// node pushed below the marker.  Lanes with cur == kTraceDone and no held
// leaf (finished or idle) only take part in the wave vote.
template <int TIES, bool STATS>
__device__ __forceinline__ void traceDescend(
    TraceState &s, const DScene &S, const Hit &best, uint32_t *lds, uint32_t *ovf, TraceStats &st)
{
  for (;;) {
    if (s.cur >= 0 && s.cur < kTraceDone) {
      const Node4 nd = loadNode(S.nodes, s.cur);
      if (STATS)
        st.nodes++;
      // with TIES, boxes up to a near-tie beyond the closest hit stay open
      const float tb = TIES != TIES_OFF && best.hit ? fminf(best.t + kTieEps * best.t, s.tfar) : best.t;
      float t0c = slabNear(nd.lox.x, nd.loy.x, nd.loz.x, nd.hix.x, nd.hiy.x, nd.hiz.x, s.o, s.inv, s.t0, tb);
      float t1c = slabNear(nd.lox.y, nd.loy.y, nd.loz.y, nd.hix.y, nd.hiy.y, nd.hiz.y, s.o, s.inv, s.t0, tb);
      float t2c = slabNear(nd.lox.z, nd.loy.z, nd.loz.z, nd.hix.z, nd.hiy.z, nd.hiz.z, s.o, s.inv, s.t0, tb);
      float t3c = slabNear(nd.lox.w, nd.loy.w, nd.loz.w, nd.hix.w, nd.hiy.w, nd.hiz.w, s.o, s.inv, s.t0, tb);
      int c0 = nd.child.x, c1 = nd.child.y, c2 = nd.child.z, c3 = nd.child.w;
      cswap(t0c, c0, t1c, c1);
      cswap(t2c, c2, t3c, c3);
      cswap(t0c, c0, t2c, c2);
      cswap(t1c, c1, t3c, c3);
      cswap(t1c, c1, t2c, c2);
      if (t0c == kInf) {
        s.cur = s.sp == 0 ? kTraceDone : tracePop(s, lds, ovf);
      } else {
        const int npush = (t1c != kInf) + (t2c != kInf) + (t3c != kInf);
        if (s.sp + 3 <= kLdsDepth) {
          // the children after the nearest, farthest first: the sorted
          // boxes that were hit are a prefix, so the npush entries pushed
          // are c_npush .. c1; the slots above them get don't-care copies
          uint32_t *top = lds + s.sp * kTraceBlock + threadIdx.x;
          top[0] = (uint32_t)(npush == 3 ? c3 : (npush == 2 ? c2 : c1));
          top[kTraceBlock] = (uint32_t)(npush == 3 ? c2 : c1);
          top[2 * kTraceBlock] = (uint32_t)c1;
          s.sp += npush;
          s.cur = c0;
        } else if (s.sp + npush > kLdsDepth + kOverflowDepth) { // cannot happen: the world commit checks the stack need
          s.cur = kTraceDone;
          s.sp = 0;
        } else {
          if (t3c != kInf)
            tracePush(s, lds, ovf, c3);
          if (t2c != kInf)
            tracePush(s, lds, ovf, c2);
          if (t1c != kInf)
            tracePush(s, lds, ovf, c1);
          s.cur = c0;
        }
      }
    } else if (s.cur == kMarker && s.post == 0) {
      s.o = s.worg;
      s.d = s.wdir;
      s.inv = s.winv;
      s.inst = -1;
      s.cur = s.sp == 0 ? kTraceDone : tracePop(s, lds, ovf);
    }
    if (s.cur < 0 && s.post == 0) { // hold the leaf, go on
      s.post = s.cur;
      s.cur = s.sp == 0 ? kTraceDone : tracePop(s, lds, ovf);
    }
    const bool canStep = (s.cur >= 0 && s.cur < kTraceDone) || (s.cur == kMarker && s.post == 0);
    if (__all(!canStep || s.post != 0))
      break;
  }
}

// The leaf a lane holds (s.post != 0): an instance leaf enters the
// instance's BLAS with the object ray; a BLAS leaf tests its primitives.
// Returns true when the lane's traversal has ended: no node left, an
// opaque occluder (MODE_OCCLUSION, `occluded` set) or a full stack.
template <int MODE, bool STATS, bool RECORD, int TIES>
__device__ __forceinline__ bool traceLeaf(TraceState &s, const DScene &S, Hit &best, uint32_t *lds, uint32_t *ovf,
    TraceStats &st, bool &occluded)
{
  const uint32_t enc = ~(uint32_t)s.post;
  s.post = 0;
  const uint32_t first = enc >> 3;
  const uint32_t count = (enc & 7u) + 1u;
  if (s.inst < 0) {
    // world leaf: one instance; continue in its BLAS with the object ray
    s.inst = (int)first;
    const DInstance &I = S.instances[s.inst];
    const linear3f rl = {{I.rcp[0], I.rcp[1], I.rcp[2]}, {I.rcp[3], I.rcp[4], I.rcp[5]}, {I.rcp[6], I.rcp[7], I.rcp[8]}};
    s.o = mul(rl, s.worg) + v3(I.rcp[9], I.rcp[10], I.rcp[11]);
    s.d = mul(rl, s.wdir);
    s.inv = v3(safeInv(s.d.x), safeInv(s.d.y), safeInv(s.d.z));
    s.pc = triPrecompute(s.d);
    s.modelOffset = I.modelOffset;
    if (s.sp + 2 > kLdsDepth + kOverflowDepth)
      return true;
    if (s.cur != kTraceDone)
      tracePush(s, lds, ovf, s.cur);
    tracePush(s, lds, ovf, kMarker);
    s.cur = I.blasRoot;
    return false;
  }
  const float t0 = s.t0;
  for (uint32_t k = 0; k < count; k++) {
    const DPrim &P = S.prims[first + k];
    const float4 pa = reinterpret_cast<const float4 *>(&P)[0];
    const float4 pb = reinterpret_cast<const float4 *>(&P)[1];
    const float4 pcv = reinterpret_cast<const float4 *>(&P)[2];
    const uint32_t primID = __float_as_uint(pa.w);
    const uint32_t mk = __float_as_uint(pb.w);
    const uint32_t model = mk & 0x0fffffffu;
    const uint32_t kind = mk >> 28;
    if (STATS)
      st.prims++;
    if (MODE == MODE_OCCLUSION && kind == PRIM_SPHERE) {
      float tE, tX;
      vec3f NE, NX;
      const int roots = sphereRoots(s.o, s.d, t0, v3(pa.x, pa.y, pa.z), pb.x, best.t, tE, NE, tX, NX);
      if (!roots)
        continue;
      if (S.models[s.modelOffset + model].opaque) {
        occluded = true;
        return true;
      }
      s.sawTransparent = true;
      if (RECORD) {
        if (roots & 1)
          recordHit(ovf, tE, (uint32_t)s.inst, first + k, 0);
        if (roots & 2)
          recordHit(ovf, tX, (uint32_t)s.inst, first + k, 1);
      }
      continue;
    }
    float t, u = 0.f, v = 0.f;
    vec3f Ng;
    bool h;
    uint32_t half = 0;
    const float tmaxP = TIES != TIES_OFF && best.hit ? fminf(best.t + kTieEps * best.t, s.tfar) : best.t;
    if (kind == PRIM_SPHERE) {
      h = intersectSphere(s.o, s.d, t0, v3(pa.x, pa.y, pa.z), pb.x, tmaxP, t, Ng);
    } else {
      const vec3f A = v3(pa.x, pa.y, pa.z), B = v3(pb.x, pb.y, pb.z), C = v3(pcv.x, pcv.y, pcv.z);
      h = intersectTriangle(s.o, t0, s.pc, A, B, C, tmaxP, t, u, v);
      if (h) {
        Ng = cross(B - A, C - A);
        if (kind == PRIM_QUAD1) {
          half = 1;
          u = 1.0f - u;
          v = 1.0f - v;
        }
      }
    }
    if (!h)
      continue;
    if (MODE == MODE_OCCLUSION) {
      if (S.models[s.modelOffset + model].opaque) {
        occluded = true;
        return true;
      }
      s.sawTransparent = true;
      if (RECORD)
        recordHit(ovf, t, (uint32_t)s.inst, first + k, 0);
      continue;
    }
    const bool nt = TIES != TIES_OFF && best.hit && nearTie(t, best.t);
    if (better(t, (uint32_t)s.inst, model, primID, half, best)) {
      if (TIES != TIES_OFF)
        s.tie = nt; // a hit below the window of the previous closest one clears the flag
      best.hit = true;
      best.t = t;
      best.u = u;
      best.v = v;
      best.Ng = Ng;
      best.inst = (uint32_t)s.inst;
      best.model = model;
      best.prim = primID;
      best.half = half;
    } else if (nt) {
      s.tie = true;
    }
  }
  return s.cur == kTraceDone;
}

template <int MODE, bool STATS, bool RECORD = false, int TIES = TIES_OFF>
__device__ bool traceRay(const DScene &S, vec3f worg, vec3f wdir, float t0, float tfar, Hit &best,
    uint32_t *lds, uint32_t *ovf, bool &sawTransparent, TraceStats &st)
{
  TraceState s;
  traceBegin(s, S, worg, wdir, t0, tfar, best);
  bool occluded = false;
  for (;;) {
    traceDescend<TIES, STATS>(s, S, best, lds, ovf, st);
    if (s.post == 0) { // nothing held: the lane has no node left
      if (s.cur == kTraceDone)
        break;
      continue; // (cannot happen: a lane without a held leaf can step)
    }
    if (traceLeaf<MODE, STATS, RECORD, TIES>(s, S, best, lds, ovf, st, occluded))
      break;
  }
  sawTransparent = s.sawTransparent;
  if (MODE == MODE_OCCLUSION && occluded)
    return true;
  if (TIES == TIES_RESOLVE && s.tie)
    resolveNearTie<STATS>(S, worg, wdir, t0, tfar, best, lds, ovf, st);
  if (TIES == TIES_REPORT)
    sawTransparent = s.tie;
  return best.hit;
}

// The hit record the closest-hit traversal of (worg, wdir) returns for a
// hit recorded by the occlusion traversal: the same object ray (instance
// transform, triPrecompute) and the same primitive test give the same
// t, u, v and Ng bit for bit.
__device__ __forceinline__ void hitFromRecord(const DScene &S, vec3f worg, vec3f wdir, float4 r, Hit &h)
{
  const uint32_t inst = __float_as_uint(r.y);
  const DInstance &I = S.instances[inst];
  const linear3f rl = {{I.rcp[0], I.rcp[1], I.rcp[2]}, {I.rcp[3], I.rcp[4], I.rcp[5]}, {I.rcp[6], I.rcp[7], I.rcp[8]}};
  const vec3f o = mul(rl, worg) + v3(I.rcp[9], I.rcp[10], I.rcp[11]);
  const vec3f d = mul(rl, wdir);
  const DPrim &P = S.prims[__float_as_uint(r.z)];
  const float4 pa = reinterpret_cast<const float4 *>(&P)[0];
  const float4 pb = reinterpret_cast<const float4 *>(&P)[1];
  const float4 pcv = reinterpret_cast<const float4 *>(&P)[2];
  const uint32_t mk = __float_as_uint(pb.w);
  const uint32_t kind = mk >> 28;
  h.hit = true;
  h.t = r.x;
  h.inst = inst;
  h.model = mk & 0x0fffffffu;
  h.prim = __float_as_uint(pa.w);
  h.half = 0;
  h.u = h.v = 0.f;
  if (kind == PRIM_SPHERE) {
    float tE, tX;
    vec3f NE, NX;
    sphereRoots(o, d, -kInf, v3(pa.x, pa.y, pa.z), pb.x, kInf, tE, NE, tX, NX);
    h.Ng = __float_as_uint(r.w) ? NX : NE;
  } else {
    const vec3f A = v3(pa.x, pa.y, pa.z), B = v3(pb.x, pb.y, pb.z), C = v3(pcv.x, pcv.y, pcv.z);
    float t, u = 0.f, v = 0.f;
    intersectTriangle(o, -kInf, triPrecompute(d), A, B, C, kInf, t, u, v);
    h.Ng = cross(B - A, C - A);
    h.u = u;
    h.v = v;
    if (kind == PRIM_QUAD1) {
      h.half = 1;
      h.u = 1.0f - u;
      h.v = 1.0f - v;
    }
  }
}

} // namespace hpt
