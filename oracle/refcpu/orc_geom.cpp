// ORACLE — test infrastructure only.  See orc_geom.h for what is restated.
#include "orc_geom.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace orc {

static Box emptyBox() { return {v3(kInf), v3(-kInf)}; }
static void extend(Box &b, const Box &o)
{
  b.lo = vmin(b.lo, o.lo);
  b.hi = vmax(b.hi, o.hi);
}
static void extend(Box &b, vec3f p)
{
  b.lo = vmin(b.lo, p);
  b.hi = vmax(b.hi, p);
}
static float halfArea(const Box &b)
{
  const vec3f d = b.hi - b.lo;
  if (d.x < 0.f)
    return 0.f;
  return d.x * d.y + d.y * d.z + d.z * d.x;
}

// binned SAH (16 bins per axis, leaves of at most 4 items)
void BVH::build(const std::vector<Box> &boxes)
{
  nodes.clear();
  items.resize(boxes.size());
  for (size_t i = 0; i < boxes.size(); i++)
    items[i] = (uint32_t)i;
  if (boxes.empty()) {
    BVHNode leaf{emptyBox(), -1, -1, 0, 0};
    nodes.push_back(leaf);
    return;
  }
  std::vector<vec3f> cent(boxes.size());
  for (size_t i = 0; i < boxes.size(); i++)
    cent[i] = (boxes[i].lo + boxes[i].hi) * 0.5f;

  struct Task
  {
    uint32_t node, begin, end;
  };
  nodes.reserve(2 * boxes.size() / 2 + 1);
  nodes.push_back(BVHNode{});
  std::vector<Task> stack{{0, 0, (uint32_t)boxes.size()}};
  constexpr int kBins = 16;
  constexpr uint32_t kLeaf = 4;
  while (!stack.empty()) {
    Task t = stack.back();
    stack.pop_back();
    Box bb = emptyBox(), cb = emptyBox();
    for (uint32_t i = t.begin; i < t.end; i++) {
      extend(bb, boxes[items[i]]);
      extend(cb, cent[items[i]]);
    }
    const uint32_t n = t.end - t.begin;
    BVHNode &node = nodes[t.node];
    node.box = bb;
    auto makeLeaf = [&]() {
      nodes[t.node].left = -1;
      nodes[t.node].right = -1;
      nodes[t.node].first = t.begin;
      nodes[t.node].count = n;
    };
    if (n <= kLeaf) {
      makeLeaf();
      continue;
    }
    int bestAxis = -1, bestSplit = -1;
    float bestCost = kInf;
    const vec3f ext = cb.hi - cb.lo;
    for (int axis = 0; axis < 3; axis++) {
      const float lo = comp(cb.lo, axis), e = comp(ext, axis);
      if (!(e > 0.f))
        continue;
      Box binBox[kBins];
      uint32_t binCnt[kBins] = {0};
      for (int b = 0; b < kBins; b++)
        binBox[b] = emptyBox();
      const float scale = (float)kBins / e;
      for (uint32_t i = t.begin; i < t.end; i++) {
        int b = (int)((comp(cent[items[i]], axis) - lo) * scale);
        b = b < 0 ? 0 : (b >= kBins ? kBins - 1 : b);
        binCnt[b]++;
        extend(binBox[b], boxes[items[i]]);
      }
      float rightArea[kBins];
      uint32_t rightCnt[kBins];
      Box acc = emptyBox();
      uint32_t cnt = 0;
      for (int b = kBins - 1; b > 0; b--) {
        extend(acc, binBox[b]);
        cnt += binCnt[b];
        rightArea[b] = halfArea(acc);
        rightCnt[b] = cnt;
      }
      acc = emptyBox();
      cnt = 0;
      for (int b = 0; b < kBins - 1; b++) {
        extend(acc, binBox[b]);
        cnt += binCnt[b];
        if (cnt == 0 || rightCnt[b + 1] == 0)
          continue;
        const float cost = halfArea(acc) * (float)cnt + rightArea[b + 1] * (float)rightCnt[b + 1];
        if (cost < bestCost) {
          bestCost = cost;
          bestAxis = axis;
          bestSplit = b;
        }
      }
    }
    uint32_t mid;
    if (bestAxis < 0) {
      // all centroids coincide: split the item list in half
      if (n <= 8) {
        makeLeaf();
        continue;
      }
      mid = t.begin + n / 2;
    } else {
      const float leafCost = halfArea(bb) * (float)n;
      if (n <= 8 && leafCost <= bestCost + halfArea(bb)) {
        makeLeaf();
        continue;
      }
      const float lo = comp(cb.lo, bestAxis);
      const float scale = (float)kBins / comp(ext, bestAxis);
      auto pm = std::partition(items.begin() + t.begin, items.begin() + t.end, [&](uint32_t id) {
        int b = (int)((comp(cent[id], bestAxis) - lo) * scale);
        b = b < 0 ? 0 : (b >= kBins ? kBins - 1 : b);
        return b <= bestSplit;
      });
      mid = (uint32_t)(pm - items.begin());
      if (mid == t.begin || mid == t.end)
        mid = t.begin + n / 2;
    }
    const uint32_t l = (uint32_t)nodes.size();
    nodes.push_back(BVHNode{});
    nodes.push_back(BVHNode{});
    nodes[t.node].left = (int32_t)l;
    nodes[t.node].right = (int32_t)l + 1;
    nodes[t.node].count = 0;
    stack.push_back({l + 1, mid, t.end});
    stack.push_back({l, t.begin, mid});
  }
}

void itemTriangle(const OrcScene &s, const OrcGeometry &g, const BLASItem &it, vec3f &a, vec3f &b, vec3f &c)
{
  (void)s;
  if (g.type == ORC_GEOM_TRIANGLES) {
    const uint32_t *ix = g.index + 3 * (int64_t)it.prim;
    a = ld3(g.vertex, ix[0]);
    b = ld3(g.vertex, ix[1]);
    c = ld3(g.vertex, ix[2]);
  } else {
    const uint32_t *ix = g.index + 4 * (int64_t)it.prim;
    if (it.half == 0) {
      a = ld3(g.vertex, ix[0]);
      b = ld3(g.vertex, ix[1]);
      c = ld3(g.vertex, ix[3]);
    } else {
      a = ld3(g.vertex, ix[2]);
      b = ld3(g.vertex, ix[3]);
      c = ld3(g.vertex, ix[1]);
    }
  }
}

void World::build(const OrcScene &scene)
{
  desc = scene;
  groups.assign(scene.numGroups, GroupAccel{});
  for (int gi = 0; gi < scene.numGroups; gi++) {
    GroupAccel &ga = groups[gi];
    std::vector<Box> boxes;
    const OrcGroup &grp = scene.groups[gi];
    for (int mi = 0; mi < grp.numModels; mi++) {
      const OrcModel &m = scene.models[scene.groupModels[grp.firstModel + mi]];
      const OrcGeometry &g = scene.geometries[m.geometry];
      for (int64_t p = 0; p < g.numPrimitives; p++) {
        if (g.type == ORC_GEOM_SPHERES) {
          const float *sp = g.spheres + 4 * p;
          const vec3f c = v3(sp[0], sp[1], sp[2]);
          const float r = sp[3];
          ga.items.push_back({(uint32_t)mi, (uint32_t)p, 0});
          boxes.push_back({c - v3(r), c + v3(r)});
          continue;
        }
        const int halves = g.type == ORC_GEOM_QUADS ? 2 : 1;
        for (int h = 0; h < halves; h++) {
          BLASItem it{(uint32_t)mi, (uint32_t)p, (uint32_t)h};
          vec3f a, b, c;
          itemTriangle(scene, g, it, a, b, c);
          Box bx = {vmin(vmin(a, b), c), vmax(vmax(a, b), c)};
          ga.items.push_back(it);
          boxes.push_back(bx);
        }
      }
    }
    ga.bvh.build(boxes);
    ga.bounds = ga.bvh.nodes[0].box;
  }
  xfm.resize(scene.numInstances);
  rcpXfm.resize(scene.numInstances);
  instBounds.resize(scene.numInstances);
  for (int i = 0; i < scene.numInstances; i++) {
    const float *x = scene.instances[i].xfm;
    affine3f a{{{x[0], x[1], x[2]}, {x[3], x[4], x[5]}, {x[6], x[7], x[8]}}, {x[9], x[10], x[11]}};
    xfm[i] = a;
    rcpXfm[i] = inverse(a);
    const Box &ob = groups[scene.instances[i].group].bounds;
    Box wb = emptyBox();
    if (ob.lo.x <= ob.hi.x) {
      for (int c = 0; c < 8; c++) {
        const vec3f p = v3(c & 1 ? ob.hi.x : ob.lo.x, c & 2 ? ob.hi.y : ob.lo.y, c & 4 ? ob.hi.z : ob.lo.z);
        extend(wb, xfmPoint(a, p));
      }
      // transformed corners are rounded: pad by a relative margin
      const vec3f pad = (vabs(wb.lo) + vabs(wb.hi)) * 1e-6f;
      wb.lo = wb.lo - pad;
      wb.hi = wb.hi + pad;
    }
    instBounds[i] = wb;
  }
  tlas.build(instBounds);
}

struct BoxRay
{
  vec3f org, invd;
};

static inline float safeInv(float d)
{
  const float ad = std::fabs(d);
  return 1.0f / (ad < 1e-18f ? std::copysign(1e-18f, d) : d);
}

static inline BoxRay makeBoxRay(const Ray &r)
{
  return {r.org, v3(safeInv(r.dir.x), safeInv(r.dir.y), safeInv(r.dir.z))};
}

// conservative slab test: entry distance or +inf if the box is missed
static inline float boxEntry(const BoxRay &br, const Box &b, float t0, float tbest)
{
  const float x0 = (b.lo.x - br.org.x) * br.invd.x, x1 = (b.hi.x - br.org.x) * br.invd.x;
  const float y0 = (b.lo.y - br.org.y) * br.invd.y, y1 = (b.hi.y - br.org.y) * br.invd.y;
  const float z0 = (b.lo.z - br.org.z) * br.invd.z, z1 = (b.hi.z - br.org.z) * br.invd.z;
  const float tn = maxf(maxf(minf(x0, x1), minf(y0, y1)), maxf(minf(z0, z1), t0));
  float tf = minf(minf(maxf(x0, x1), maxf(y0, y1)), minf(maxf(z0, z1), tbest));
  tf *= 1.0000004f;
  if (tn > tf)
    return kInf;
  return tn;
}

// float hit of BLAS item `it` of instance `inst` for the object ray r
// (t in (r.t0, tmax]): t, u, v, Ng and, for a sphere, which root
static bool itemHit(const OrcScene &desc, const BLASItem &it, const OrcGeometry &g, const Ray &r, const TriPrecomp &pc,
    float tmax, float &t, float &u, float &v, vec3f &Ng, int &exitRoot)
{
  u = v = 0.f;
  exitRoot = 0;
  if (g.type == ORC_GEOM_SPHERES) {
    const float *sp4 = g.spheres + 4 * (int64_t)it.prim;
    const bool h = intersectSphere(r, v3(sp4[0], sp4[1], sp4[2]), sp4[3], tmax, t, Ng);
    if (h) {
      // intersectSphere takes the entry root when it is in range: recompute
      // the entry root with the same arithmetic to tell which one it was
      const vec3f d = r.dir;
      const float rd2 = 1.0f / dot(d, d);
      const vec3f CO = v3(sp4[0], sp4[1], sp4[2]) - r.org;
      const float projCO = dot(CO, d) * rd2;
      const vec3f perp = CO - projCO * d;
      const float l2 = dot(perp, perp);
      const float r2 = sp4[3] * sp4[3];
      const float td = std::sqrt((r2 - l2) * rd2);
      const float tE = projCO - td;
      exitRoot = (tE > r.t0 && tE <= tmax) ? 0 : 1;
    }
    return h;
  }
  vec3f a, b, c;
  itemTriangle(desc, g, it, a, b, c);
  const bool h = intersectTriangle(r, pc, a, b, c, tmax, t, u, v);
  if (h) {
    Ng = cross(b - a, c - a);
    if (it.half) {
      u = 1.0f - u;
      v = 1.0f - v;
    }
  }
  return h;
}

Hit World::intersect(const Ray &rayIn, uint64_t *nodeVisits, uint64_t *primTests, bool resolveTies) const
{
  Hit best;
  best.hit = false;
  best.t = rayIn.t;
  best.inst = best.model = best.prim = kInvalidID;
  best.half = 0;
  best.u = best.v = 0.f;
  best.Ng = v3(0.f);
  if (tlas.nodes.empty() || desc.numInstances == 0)
    return best;
  bool tie = false; // a near-tie of the current closest hit has been seen
  // with resolveTies, boxes and primitives up to a near-tie beyond the
  // closest hit stay open
  auto tLimit = [&]() { return resolveTies && best.hit ? minf(best.t + kTieEps * best.t, rayIn.t) : best.t; };
  const BoxRay wbr = makeBoxRay(rayIn);
  uint32_t tstack[64];
  int tsp = 0;
  tstack[tsp++] = 0;
  while (tsp) {
    const BVHNode &n = tlas.nodes[tstack[--tsp]];
    if (nodeVisits)
      (*nodeVisits)++;
    if (boxEntry(wbr, n.box, rayIn.t0, tLimit()) == kInf)
      continue;
    if (n.left >= 0) {
      tstack[tsp++] = (uint32_t)n.right;
      tstack[tsp++] = (uint32_t)n.left;
      continue;
    }
    for (uint32_t k = 0; k < n.count; k++) {
      const uint32_t inst = tlas.items[n.first + k];
      const GroupAccel &ga = groups[desc.instances[inst].group];
      const affine3f &rx = rcpXfm[inst];
      Ray r;
      r.org = xfmPoint(rx, rayIn.org);
      r.dir = xfmVector(rx, rayIn.dir);
      r.t0 = rayIn.t0;
      r.t = rayIn.t;
      const BoxRay br = makeBoxRay(r);
      const TriPrecomp pc = triPrecompute(r.dir);
      uint32_t stack[128];
      int sp = 0;
      stack[sp++] = 0;
      while (sp) {
        const BVHNode &bn = ga.bvh.nodes[stack[--sp]];
        if (nodeVisits)
          (*nodeVisits)++;
        if (boxEntry(br, bn.box, r.t0, tLimit()) == kInf)
          continue;
        if (bn.left >= 0) {
          stack[sp++] = (uint32_t)bn.right;
          stack[sp++] = (uint32_t)bn.left;
          continue;
        }
        for (uint32_t j = 0; j < bn.count; j++) {
          const BLASItem &it = ga.items[ga.bvh.items[bn.first + j]];
          const OrcModel &m = desc.models[desc.groupModels[desc.groups[desc.instances[inst].group].firstModel + it.model]];
          const OrcGeometry &g = desc.geometries[m.geometry];
          if (primTests)
            (*primTests)++;
          float t, u, v;
          vec3f Ng;
          int exitRoot;
          if (!itemHit(desc, it, g, r, pc, tLimit(), t, u, v, Ng, exitRoot))
            continue;
          const bool nt = resolveTies && best.hit && nearTie(t, best.t);
          if (better(t, inst, it.model, it.prim, it.half, best)) {
            // a hit below the window of the previous closest one clears the flag
            tie = nt;
            best.hit = true;
            best.t = t;
            best.u = u;
            best.v = v;
            best.inst = inst;
            best.model = it.model;
            best.prim = it.prim;
            best.half = it.half;
            best.Ng = Ng;
          } else if (nt) {
            tie = true;
          }
        }
      }
    }
  }
  if (!tie)
    return best;
  // FP64 re-resolution over the window around the closest hit: the
  // smallest (FP64 t, key) among every hit whose float t is a near-tie
  const float tF = best.t;
  const float tHi = minf(tF + kTieEps * tF, rayIn.t);
  Hit w = best;
  double wD = 0.0;
  bool have = false;
  tsp = 0;
  tstack[tsp++] = 0;
  while (tsp) {
    const BVHNode &n = tlas.nodes[tstack[--tsp]];
    if (nodeVisits)
      (*nodeVisits)++;
    if (boxEntry(wbr, n.box, rayIn.t0, tHi) == kInf)
      continue;
    if (n.left >= 0) {
      tstack[tsp++] = (uint32_t)n.right;
      tstack[tsp++] = (uint32_t)n.left;
      continue;
    }
    for (uint32_t k = 0; k < n.count; k++) {
      const uint32_t inst = tlas.items[n.first + k];
      const GroupAccel &ga = groups[desc.instances[inst].group];
      const affine3f &rx = rcpXfm[inst];
      Ray r;
      r.org = xfmPoint(rx, rayIn.org);
      r.dir = xfmVector(rx, rayIn.dir);
      r.t0 = rayIn.t0;
      r.t = rayIn.t;
      const BoxRay br = makeBoxRay(r);
      const TriPrecomp pc = triPrecompute(r.dir);
      double od[3], dd[3];
      fp64ObjectRay(rx, rayIn.org, rayIn.dir, od, dd);
      uint32_t stack[128];
      int sp = 0;
      stack[sp++] = 0;
      while (sp) {
        const BVHNode &bn = ga.bvh.nodes[stack[--sp]];
        if (nodeVisits)
          (*nodeVisits)++;
        if (boxEntry(br, bn.box, r.t0, tHi) == kInf)
          continue;
        if (bn.left >= 0) {
          stack[sp++] = (uint32_t)bn.right;
          stack[sp++] = (uint32_t)bn.left;
          continue;
        }
        for (uint32_t j = 0; j < bn.count; j++) {
          const BLASItem &it = ga.items[ga.bvh.items[bn.first + j]];
          const OrcModel &m = desc.models[desc.groupModels[desc.groups[desc.instances[inst].group].firstModel + it.model]];
          const OrcGeometry &g = desc.geometries[m.geometry];
          if (primTests)
            (*primTests)++;
          float t, u, v;
          vec3f Ng;
          int exitRoot;
          if (!itemHit(desc, it, g, r, pc, tHi, t, u, v, Ng, exitRoot) || !nearTie(t, tF))
            continue;
          double tD;
          if (g.type == ORC_GEOM_SPHERES) {
            const float *sp4 = g.spheres + 4 * (int64_t)it.prim;
            tD = fp64SphereT(od, dd, v3(sp4[0], sp4[1], sp4[2]), sp4[3], exitRoot);
          } else {
            vec3f a, b, c;
            itemTriangle(desc, g, it, a, b, c);
            tD = fp64TriangleT(od, dd, a, b, c);
          }
          if (!have || tD < wD || (tD == wD && keyLess(inst, it.model, it.prim, it.half, w))) {
            have = true;
            wD = tD;
            w.hit = true;
            w.t = t;
            w.u = u;
            w.v = v;
            w.inst = inst;
            w.model = it.model;
            w.prim = it.prim;
            w.half = it.half;
            w.Ng = Ng;
          }
        }
      }
    }
  }
  return have ? w : best;
}

// double-precision closest hit by brute force (no BVH)
Hit intersectBruteForceFP64(const World &w, const Ray &ray)
{
  Hit best;
  best.hit = false;
  best.t = ray.t;
  best.inst = best.model = best.prim = kInvalidID;
  best.half = 0;
  best.u = best.v = 0.f;
  best.Ng = v3(0.f);
  double bestT = ray.t;
  const OrcScene &s = w.desc;
  for (int i = 0; i < s.numInstances; i++) {
    const affine3f &x = w.xfm[i];
    // invert in double
    const double m[9] = {x.l.vx.x, x.l.vy.x, x.l.vz.x, x.l.vx.y, x.l.vy.y, x.l.vz.y, x.l.vx.z, x.l.vy.z, x.l.vz.z};
    const double det = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) + m[2] * (m[3] * m[7] - m[4] * m[6]);
    double inv[9];
    inv[0] = (m[4] * m[8] - m[5] * m[7]) / det;
    inv[1] = (m[2] * m[7] - m[1] * m[8]) / det;
    inv[2] = (m[1] * m[5] - m[2] * m[4]) / det;
    inv[3] = (m[5] * m[6] - m[3] * m[8]) / det;
    inv[4] = (m[0] * m[8] - m[2] * m[6]) / det;
    inv[5] = (m[2] * m[3] - m[0] * m[5]) / det;
    inv[6] = (m[3] * m[7] - m[4] * m[6]) / det;
    inv[7] = (m[1] * m[6] - m[0] * m[7]) / det;
    inv[8] = (m[0] * m[4] - m[1] * m[3]) / det;
    const double po[3] = {(double)ray.org.x - x.p.x, (double)ray.org.y - x.p.y, (double)ray.org.z - x.p.z};
    double o[3], d[3];
    for (int r = 0; r < 3; r++) {
      o[r] = inv[3 * r] * po[0] + inv[3 * r + 1] * po[1] + inv[3 * r + 2] * po[2];
      d[r] = inv[3 * r] * ray.dir.x + inv[3 * r + 1] * ray.dir.y + inv[3 * r + 2] * ray.dir.z;
    }
    const OrcGroup &grp = s.groups[s.instances[i].group];
    for (int mi = 0; mi < grp.numModels; mi++) {
      const OrcModel &mdl = s.models[s.groupModels[grp.firstModel + mi]];
      const OrcGeometry &g = s.geometries[mdl.geometry];
      for (int64_t p = 0; p < g.numPrimitives; p++) {
        if (g.type == ORC_GEOM_SPHERES) {
          const float *sp = g.spheres + 4 * p;
          const double oc[3] = {o[0] - sp[0], o[1] - sp[1], o[2] - sp[2]};
          const double A = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
          const double B = oc[0] * d[0] + oc[1] * d[1] + oc[2] * d[2];
          const double C = oc[0] * oc[0] + oc[1] * oc[1] + oc[2] * oc[2] - (double)sp[3] * sp[3];
          const double disc = B * B - A * C;
          if (disc < 0)
            continue;
          const double sq = std::sqrt(disc);
          double t = (-B - sq) / A;
          if (!(t > ray.t0 && t <= bestT))
            t = (-B + sq) / A;
          if (!(t > ray.t0 && t <= bestT))
            continue;
          if (!best.hit || t < bestT || (t == bestT && keyLess(i, mi, (uint32_t)p, 0, best))) {
            bestT = t;
            best.hit = true;
            best.t = (float)t;
            best.inst = i;
            best.model = mi;
            best.prim = (uint32_t)p;
            best.half = 0;
          }
          continue;
        }
        const int halves = g.type == ORC_GEOM_QUADS ? 2 : 1;
        for (int h = 0; h < halves; h++) {
          BLASItem it{(uint32_t)mi, (uint32_t)p, (uint32_t)h};
          vec3f a, b, c;
          itemTriangle(s, g, it, a, b, c);
          const double e1[3] = {(double)b.x - a.x, (double)b.y - a.y, (double)b.z - a.z};
          const double e2[3] = {(double)c.x - a.x, (double)c.y - a.y, (double)c.z - a.z};
          const double pv[3] = {d[1] * e2[2] - d[2] * e2[1], d[2] * e2[0] - d[0] * e2[2], d[0] * e2[1] - d[1] * e2[0]};
          const double dt = e1[0] * pv[0] + e1[1] * pv[1] + e1[2] * pv[2];
          if (dt == 0.0)
            continue;
          const double tv[3] = {o[0] - a.x, o[1] - a.y, o[2] - a.z};
          const double u = (tv[0] * pv[0] + tv[1] * pv[1] + tv[2] * pv[2]) / dt;
          if (u < 0.0 || u > 1.0)
            continue;
          const double qv[3] = {tv[1] * e1[2] - tv[2] * e1[1], tv[2] * e1[0] - tv[0] * e1[2], tv[0] * e1[1] - tv[1] * e1[0]};
          const double v = (d[0] * qv[0] + d[1] * qv[1] + d[2] * qv[2]) / dt;
          if (v < 0.0 || u + v > 1.0)
            continue;
          const double t = (e2[0] * qv[0] + e2[1] * qv[1] + e2[2] * qv[2]) / dt;
          if (!(t > ray.t0 && t <= bestT))
            continue;
          if (!best.hit || t < bestT || (t == bestT && keyLess(i, mi, (uint32_t)p, h, best))) {
            bestT = t;
            best.hit = true;
            best.t = (float)t;
            best.u = (float)(h ? 1.0 - u : u);
            best.v = (float)(h ? 1.0 - v : v);
            best.inst = i;
            best.model = mi;
            best.prim = (uint32_t)p;
            best.half = h;
          }
        }
      }
    }
  }
  return best;
}

} // namespace orc
