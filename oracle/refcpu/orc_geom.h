// ORACLE — test infrastructure only.
//
// Ray/primitive intersection and closest-hit traversal for the CPU oracle.
// The reference delegates this to Embree 4.3.3 (rtcIntersectV at
// modules/cpu/common/World.ih:49-51), whose source is not part of the
// reference tree; parity with Embree itself is unpinned.  The oracle fixes a
// precise, reproducible definition instead:
//   * triangles: watertight test of Woop, Benthin and Wald (JCGT 2013) with
//     the double-precision fallback when an edge function is exactly zero;
//     u, v are the barycentric weights of v1, v2 and Ng = (v1-v0)x(v2-v0)
//     (Embree's conventions, used by TriangleMesh_postIntersect,
//     modules/cpu/geometry/Mesh.ispc:178-268);
//   * quads: split into (v0,v1,v3) and (v2,v3,v1); a hit on the second half
//     reports (1-u, 1-v) so (u,v) is the bilinear parameter consumed by
//     QuadMesh_postIntersect (Mesh.ispc:65-176);
//   * spheres: intersectSphere of modules/cpu/common/Intersect.ih:29-80;
//   * closest hit: smallest t in (tnear, tfar]; exact ties go to the
//     smallest (instance, model, primitive, half) key, so the answer does not
//     depend on the BVH layout or traversal order.
// Node boxes are only used for culling, conservatively, so any correct BVH
// returns the same hit.
#pragma once

#include <cstdint>
#include <vector>

#include "orc_math.h"
#include "orc_scene.h"

namespace orc {

static constexpr uint32_t kInvalidID = 0xffffffffu;

struct Ray
{
  vec3f org;
  vec3f dir;
  float t0; // tnear
  float t; // tfar on input, hit distance on output
};

struct Hit
{
  float t, u, v;
  uint32_t inst; // instance index in the world
  uint32_t model; // model index within the group
  uint32_t prim; // primitive index within the geometry
  uint32_t half; // quad half (0/1), 0 otherwise
  vec3f Ng; // unnormalised, object space
  bool hit;
};

inline bool keyLess(uint32_t i0, uint32_t m0, uint32_t p0, uint32_t h0, const Hit &b)
{
  if (i0 != b.inst)
    return i0 < b.inst;
  if (m0 != b.model)
    return m0 < b.model;
  if (p0 != b.prim)
    return p0 < b.prim;
  return h0 < b.half;
}

// candidate (t, key) replaces the current closest hit
inline bool better(float t, uint32_t i, uint32_t m, uint32_t p, uint32_t h, const Hit &b)
{
  return !b.hit || t < b.t || (t == b.t && keyLess(i, m, p, h, b));
}

struct TriPrecomp
{
  int kx, ky, kz;
  float Sx, Sy, Sz;
};

inline float comp(vec3f v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

inline TriPrecomp triPrecompute(vec3f dir)
{
  TriPrecomp p;
  const vec3f ad = vabs(dir);
  p.kz = ad.x > ad.y ? (ad.x > ad.z ? 0 : 2) : (ad.y > ad.z ? 1 : 2);
  p.kx = p.kz == 2 ? 0 : p.kz + 1;
  p.ky = p.kx == 2 ? 0 : p.kx + 1;
  if (comp(dir, p.kz) < 0.0f) {
    const int tmp = p.kx;
    p.kx = p.ky;
    p.ky = tmp;
  }
  const float dz = comp(dir, p.kz);
  p.Sx = comp(dir, p.kx) / dz;
  p.Sy = comp(dir, p.ky) / dz;
  p.Sz = 1.0f / dz;
  return p;
}

// returns true and (t,u,v) for a hit with t in (t0, tmax]
inline bool intersectTriangle(const Ray &ray,
    const TriPrecomp &pc,
    vec3f v0,
    vec3f v1,
    vec3f v2,
    float tmax,
    float &t,
    float &u,
    float &v)
{
  const vec3f A = v0 - ray.org;
  const vec3f B = v1 - ray.org;
  const vec3f C = v2 - ray.org;
  const float Az_ = comp(A, pc.kz), Bz_ = comp(B, pc.kz), Cz_ = comp(C, pc.kz);
  const float Ax = comp(A, pc.kx) - pc.Sx * Az_;
  const float Ay = comp(A, pc.ky) - pc.Sy * Az_;
  const float Bx = comp(B, pc.kx) - pc.Sx * Bz_;
  const float By = comp(B, pc.ky) - pc.Sy * Bz_;
  const float Cx = comp(C, pc.kx) - pc.Sx * Cz_;
  const float Cy = comp(C, pc.ky) - pc.Sy * Cz_;
  // in float only: an edge function depends on the edge's two vertices
  // and is exactly antisymmetric between the two triangles sharing it, so
  // no ray passes between them
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
  if (!(tt > ray.t0) || !(tt <= tmax))
    return false;
  t = tt;
  u = V * rcpDet;
  v = W * rcpDet;
  return true;
}

inline bool intersectSphere(const Ray &ray, vec3f center, float radius, float tmax, float &t, vec3f &N)
{
  const vec3f d = ray.dir;
  const float rd2 = 1.0f / dot(d, d);
  const vec3f CO = center - ray.org;
  const float projCO = dot(CO, d) * rd2;
  const vec3f perp = CO - projCO * d;
  const float l2 = dot(perp, perp);
  const float r2 = radius * radius;
  if (l2 > r2)
    return false;
  const float td = std::sqrt((r2 - l2) * rd2);
  const float tEntry = projCO - td;
  const float tExit = projCO + td;
  if (tEntry > ray.t0 && tEntry <= tmax) {
    t = tEntry;
    N = (-td) * d - perp;
    return true;
  }
  if (tExit > ray.t0 && tExit <= tmax) {
    t = tExit;
    N = td * d - perp;
    return true;
  }
  return false;
}

// ---- BVH --------------------------------------------------------------------

struct Box
{
  vec3f lo, hi;
};

struct BVHNode
{
  Box box;
  int32_t left; // inner: child indices; leaf: left = -1
  int32_t right;
  uint32_t first; // leaf: first item in BVH::items
  uint32_t count; // leaf: item count (0 for inner)
};

struct BVH
{
  std::vector<BVHNode> nodes;
  std::vector<uint32_t> items; // item ids in leaf order
  void build(const std::vector<Box> &boxes);
};

// primitive item packed as (model << 32 | prim << 1 | half) in a BLAS
struct BLASItem
{
  uint32_t model;
  uint32_t prim;
  uint32_t half;
};

struct GroupAccel
{
  BVH bvh;
  std::vector<BLASItem> items;
  Box bounds;
};

// Near-ties of extension rays: hits whose float distances lie within
// kTieEps (relative) of the closest one are re-resolved by their FP64
// distances (World::intersect with resolveTies).
static constexpr float kTieEps = 1.0e-5f;
inline bool nearTie(float t, float tBest) { return std::fabs(t - tBest) <= kTieEps * tBest; }

// the object ray of a world ray in double: the float transform applied in
// double precision (o = l * worg + p, d = l * wdir, columns in order)
inline void fp64ObjectRay(const affine3f &rx, vec3f worg, vec3f wdir, double o[3], double d[3])
{
  const float c[12] = {rx.l.vx.x, rx.l.vx.y, rx.l.vx.z, rx.l.vy.x, rx.l.vy.y, rx.l.vy.z,
      rx.l.vz.x, rx.l.vz.y, rx.l.vz.z, rx.p.x, rx.p.y, rx.p.z};
  const double wo[3] = {worg.x, worg.y, worg.z}, wd[3] = {wdir.x, wdir.y, wdir.z};
  for (int r = 0; r < 3; r++) {
    o[r] = (double)c[r] * wo[0] + (double)c[3 + r] * wo[1] + (double)c[6 + r] * wo[2] + (double)c[9 + r];
    d[r] = (double)c[r] * wd[0] + (double)c[3 + r] * wd[1] + (double)c[6 + r] * wd[2];
  }
}

// FP64 distance to the plane of triangle (a, b, c) along a double ray
inline double fp64TriangleT(const double o[3], const double d[3], vec3f a, vec3f b, vec3f c)
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
inline double fp64SphereT(const double o[3], const double d[3], vec3f center, float radius, int exitRoot)
{
  const double oc[3] = {o[0] - center.x, o[1] - center.y, o[2] - center.z};
  const double A = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  const double B = oc[0] * d[0] + oc[1] * d[1] + oc[2] * d[2];
  const double C = oc[0] * oc[0] + oc[1] * oc[1] + oc[2] * oc[2] - (double)radius * (double)radius;
  const double disc = B * B - A * C;
  const double sq = disc > 0.0 ? std::sqrt(disc) : 0.0;
  return (exitRoot ? (-B + sq) : (-B - sq)) / A;
}

struct World
{
  OrcScene desc; // shallow copy (arrays owned by the caller)
  std::vector<GroupAccel> groups;
  std::vector<affine3f> xfm, rcpXfm;
  BVH tlas;
  std::vector<Box> instBounds;

  void build(const OrcScene &scene);
  // closest hit; ray in world space; statistic counters optional.
  // resolveTies (extension rays): when another hit's float t is a near-tie
  // of the closest one's, the hits within that window are re-resolved by
  // their FP64 distances (smallest FP64 t, then key)
  Hit intersect(const Ray &ray, uint64_t *nodeVisits = nullptr, uint64_t *primTests = nullptr,
      bool resolveTies = false) const;
};

inline vec3f ld3(const float *p, int64_t i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }

// vertices of BLAS item (triangle or quad half) in object space
void itemTriangle(const OrcScene &s, const OrcGeometry &g, const BLASItem &it, vec3f &a, vec3f &b, vec3f &c);

// FP64 brute-force closest hit over every primitive (test helper)
Hit intersectBruteForceFP64(const World &w, const Ray &ray);

} // namespace orc
