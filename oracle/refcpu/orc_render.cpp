// ORACLE — test infrastructure only.  See orc_render.h for the reference
// files each part follows; function-level citations below.
#include "orc_render.h"

// ORC_PLANT = k builds the oracle with planted error k (a plausible mistake
// in one pinned function); tests/test_oracle_planted.py builds each one and
// checks that the pinning tests reject it.  0: no planted error.
#ifndef ORC_PLANT
#define ORC_PLANT 0
#endif

#include <algorithm>
#include <atomic>
#include <cmath>
#include <stdexcept>
#include <thread>

namespace orc {

const SobolTable &sobolTable()
{
  static SobolTable t;
  return t;
}

float hostExp(float x) { return fexp(x); }

// ---- scattering types (render/pathtracer/Scattering.ih:5-45) -----------------
enum : uint32_t
{
  SPEC_REFL = 1u << 0,
  GLOSSY_REFL = 1u << 1,
  DIFF_REFL = 1u << 2,
  SPEC_TRANS = 1u << 3,
  GLOSSY_TRANS = 1u << 4,
  DIFF_TRANS = 1u << 5,
  S_DIFFUSE = DIFF_REFL | DIFF_TRANS,
  S_GLOSSY = GLOSSY_REFL | GLOSSY_TRANS,
  S_SPECULAR = SPEC_REFL | SPEC_TRANS,
  S_REFLECTION = DIFF_REFL | GLOSSY_REFL | SPEC_REFL,
  S_TRANSMISSION = DIFF_TRANS | GLOSSY_TRANS | SPEC_TRANS,
  S_SMOOTH = S_DIFFUSE | S_GLOSSY
};

static constexpr float kMaxRouletteContProb = 0.95f; // PathTracerDefines.ih:6
static constexpr uint32_t kMaxLightSamples = 1024u; // PathTracer.ispc:35
static constexpr float kCosAngleMax = 0.99999988f; // DirectionalLightShared.h:14

inline float deg2rad(float x) { return x * 0.01745329251994329577f; }
inline float cos2sin(float c) { return std::sqrt(maxf(0.f, 1.f - c * c)); }
inline vec3f cartesian(float phi, float sinTheta, float cosTheta)
{
  float s, c;
  fsincos(phi, s, c);
  return v3(c * sinTheta, s * sinTheta, cosTheta);
}
inline vec3f cosineSampleHemisphere(vec2f s) // math/sampling.ih:43-49
{
  const float phi = kTwoPi * s.x;
  const float cosTheta = std::sqrt(s.y);
  const float sinTheta = std::sqrt(1.0f - s.y);
  return cartesian(phi, sinTheta, cosTheta);
}
inline vec3f uniformSampleSphere(float radius, vec2f s) // sampling.ih:191-197
{
  const float phi = kTwoPi * s.x;
  const float cosTheta = radius * (1.f - 2.f * s.y);
  const float sinTheta = 2.f * radius * std::sqrt(s.y * (1.f - s.y));
  return cartesian(phi, sinTheta, cosTheta);
}
inline vec3f uniformSampleCone(float cosAngle, vec2f s) // sampling.ih:92-97
{
  const float phi = kTwoPi * s.x;
  const float u = ORC_PLANT == 5 ? s.y * s.y : s.y; // plant 5: cos(theta) not uniform
  const float cosTheta = 1.0f - u * (1.0f - cosAngle);
  return cartesian(phi, cos2sin(cosTheta), cosTheta);
}
inline float rcp_safe(float f) { return rcp(std::fabs(f) < 1e-8f ? std::copysign(1e-8f, f) : f); }
inline vec3f reflect(vec3f I, vec3f N) { return (2.0f * dot(I, N)) * N - I; } // Optics.ih:11-20

// ---- setup --------------------------------------------------------------------

static constexpr int kLightGeometry = 3;

inline vec2f uniformSampleTriangleUV(vec2f s) // math/sampling.ih:165-173
{
  const float su = std::sqrt(s.x);
  return v2(s.y * su, (1.0f - s.y) * su);
}

// the vertices a, b, c, d of a mesh primitive as Mesh_sampleArea reads them
// (a triangle repeats its last vertex: index (x, y, z, z))
static bool meshPrim(const OrcGeometry &g, int64_t p, vec3f &a, vec3f &b, vec3f &c, vec3f &d)
{
  bool quad = false;
  uint32_t ix[4];
  if (g.type == ORC_GEOM_TRIANGLES) {
    const uint32_t *q = g.index + 3 * p;
    ix[0] = q[0], ix[1] = q[1], ix[2] = q[2], ix[3] = q[2];
  } else {
    const uint32_t *q = g.index + 4 * p;
    ix[0] = q[0], ix[1] = q[1], ix[2] = q[2], ix[3] = q[3];
    quad = ix[2] != ix[3];
  }
  a = ld3(g.vertex, ix[0]);
  b = ld3(g.vertex, ix[1]);
  d = ld3(g.vertex, ix[3]);
  c = quad ? ld3(g.vertex, ix[2]) : d;
  return quad;
}

// world-space area of a primitive (Mesh_getAreas, geometry/Mesh.ispc:377-415;
// Spheres_getAreas, geometry/Spheres.ispc:76-103)
static float primArea(const OrcScene &, const OrcGeometry &g, int64_t p, const affine3f &xfm)
{
  if (g.type == ORC_GEOM_SPHERES) {
    const vec3f sc = v3(length(xfm.l.vx), length(xfm.l.vy), length(xfm.l.vz));
    const float scale = 4.0f * kPi * sqr((sc.x + sc.y + sc.z) / 3.f);
    return scale * sqr(g.spheres[4 * p + 3]);
  }
  vec3f a, b, c, d;
  const bool quad = meshPrim(g, p, a, b, c, d);
  const vec3f e1 = xfmVector(xfm, a - d);
  const vec3f e2 = xfmVector(xfm, b - d);
  float area = length(cross(e1, e2));
  if (quad) {
    const vec3f e3 = xfmVector(xfm, c - d);
    area += length(cross(e2, e3));
  }
  return area * 0.5f;
}

struct SampleAreaRes
{
  vec3f pos, normal;
  float epsilon;
};

// Mesh_sampleArea (geometry/Mesh.ispc:270-351), Spheres_sampleArea
// (geometry/Spheres.ispc:44-74)
static SampleAreaRes sampleArea(const OrcGeometry &g, int64_t p, const affine3f &xfm, const affine3f &rcpXfm, vec2f s)
{
  SampleAreaRes res;
  const float xfmScale = maxf(std::fabs(xfm.l.vx.x), maxf(std::fabs(xfm.l.vy.y), std::fabs(xfm.l.vz.z)));
  if (g.type == ORC_GEOM_SPHERES) {
    const float *sp = g.spheres + 4 * p;
    const vec3f localNormal = uniformSampleSphere(sp[3], s);
    res.pos = xfmPoint(xfm, v3(sp[0], sp[1], sp[2]) + localNormal);
    res.normal = normalize(mulT(rcpXfm.l, localNormal)); // transposed(rcp_xfm.l) * n
    res.epsilon = sp[3] * kUlpEpsilon;
    res.epsilon *= xfmScale;
    return res;
  }
  vec3f a, b, c, d;
  const bool quad = meshPrim(g, p, a, b, c, d);
  res.epsilon = reduce_max(vmax(vmax(vabs(a), vabs(b)), vmax(vabs(c), vabs(d)))) * kUlpEpsilon;
  res.epsilon *= xfmScale;
  const vec3f e1 = xfmVector(xfm, a - d);
  const vec3f e2 = xfmVector(xfm, b - d);
  const vec3f m1 = cross(e1, e2);
  res.normal = m1;
  vec3f v0 = a, v1 = b, v2 = d;
  vec2f sp = s;
  if (quad) {
    const vec3f e3 = xfmVector(xfm, c - d);
    const vec3f m2 = cross(e2, e3);
    const float a1 = length(m1);
    const float a2 = length(m2);
    const float p1 = a1 * rcp(a1 + a2);
    if (s.x < p1) {
      sp.x = sp.x * rcp(p1);
    } else {
      sp.x = (s.x - p1) * rcp(1.f - p1);
      v0 = c;
      v1 = d;
      v2 = b;
      res.normal = m2;
    }
  }
  const vec2f uv = uniformSampleTriangleUV(sp);
  const vec3f localPos = v0 + uv.x * (v1 - v0) + uv.y * (v2 - v0);
  res.normal = normalize(res.normal);
  res.pos = xfmPoint(xfm, localPos);
  return res;
}

static void setupMaterial(MaterialData &md, const OrcMaterial &m)
{
  {
    md.type = m.type;
    if (m.type == ORC_MAT_OBJ) { // materials/OBJ.cpp:32-57
      md.Kd = v3(m.kd[0], m.kd[1], m.kd[2]);
      md.Ks = v3(m.ks[0], m.ks[1], m.ks[2]);
      md.Tf = v3(m.tf[0], m.tf[1], m.tf[2]);
      md.d = m.d;
      md.Ns = m.ns;
      const float total = reduce_max(md.Kd + md.Ks + md.Tf);
      if (total > 1.f) {
        md.Kd = md.Kd / total;
        md.Ks = md.Ks / total;
        md.Tf = md.Tf / total;
      }
    } else if (m.type == ORC_MAT_METAL) { // materials/Metal.cpp:30-66
      md.eta = v3(m.eta[0], m.eta[1], m.eta[2]);
      md.k = v3(m.k[0], m.k[1], m.k[2]);
      md.roughness = m.roughness;
      md.spectral = m.ior != nullptr;
      if (md.spectral) { // Metal.cpp:34-57: linear resampling to 430 + 35 l nm
        if (m.numIor <= 0)
          throw std::runtime_error("metal: empty ior");
        int64_t j = 0;
        int64_t prev = 0;
        for (int l = 0; l < 8; l++) {
          const float wl = 430.f + 35.f * (float)l;
          while (j < m.numIor && m.ior[3 * j] < wl)
            prev = j++;
          const int64_t next = j < m.numIor ? j : prev; // past the table: the last entry holds
          const float *a = m.ior + 3 * prev, *c = m.ior + 3 * next;
          if (c[0] == a[0]) {
            md.etaS[l] = a[1];
            md.kS[l] = a[2];
          } else {
            const float f = (wl - a[0]) / (c[0] - a[0]);
            md.etaS[l] = (1.f - f) * a[1] + f * c[1];
            md.kS[l] = (1.f - f) * a[2] + f * c[2];
          }
        }
      }
    } else if (m.type == ORC_MAT_LUMINOUS) { // materials/Luminous.cpp:37-48, Luminous.h:25-28
      md.radiance = v3(m.color[0], m.color[1], m.color[2]) * m.intensity;
      md.transparency = m.transparency;
      md.emissive = reduce_max(md.radiance) > 0.f;
    } else { // materials/Glass.cpp:36-62
      const float dist = maxf(m.attenuationDistance, 1e-5f);
      md.inside.ior = m.etaInside;
      md.inside.attenuation = v3(std::log(m.attenuationColorInside[0]), std::log(m.attenuationColorInside[1]),
                                  std::log(m.attenuationColorInside[2])) / dist;
      md.outside.ior = m.etaOutside;
      md.outside.attenuation = v3(std::log(m.attenuationColorOutside[0]), std::log(m.attenuationColorOutside[1]),
                                   std::log(m.attenuationColorOutside[2])) / dist;
    }
  }
}

void setupScene(Scene &s, const OrcScene &d)
{
  s.world.build(d);
  s.materials.resize(d.numMaterials);
  for (int i = 0; i < d.numMaterials; i++)
    setupMaterial(s.materials[i], d.materials[i]);
  // geometry lights (PathTracerData.cpp:121-162): every emissive model of
  // every instance, in instance and group order; the model's areaPDF is the
  // one of the last instance that gathers it (GeometricModel.ispc:14-55)
  s.modelAreaPDF.assign(d.numModels, 0.f);
  std::vector<LightData> geo;
  for (int i = 0; i < d.numInstances; i++) {
    const OrcGroup &grp = d.groups[d.instances[i].group];
    for (int k = 0; k < grp.numModels; k++) {
      const int om = d.groupModels[grp.firstModel + k];
      const OrcModel &model = d.models[om];
      if (model.material < 0 || !s.materials[model.material].emissive)
        continue;
      const OrcGeometry &g = d.geometries[model.geometry];
      if (g.numPrimitives <= 0)
        continue;
      LightData ld{};
      ld.type = kLightGeometry;
      ld.visible = true;
      ld.inst = i;
      ld.model = om;
      ld.primCDF.resize((size_t)g.numPrimitives);
      for (int64_t p = 0; p < g.numPrimitives; p++)
        ld.primCDF[p] = primArea(d, g, p, s.world.xfm[i]);
      ld.geoPdf = 1.f / Distribution1D_create((int)g.numPrimitives, ld.primCDF.data());
      s.modelAreaPDF[om] = ld.geoPdf;
      geo.push_back(std::move(ld));
    }
  }
  s.numGeoLights = (uint32_t)geo.size();
  // world lights in array order (PathTracerData.cpp:45-50)
  s.lights.resize(d.numLights);
  for (int i = 0; i < d.numLights; i++) {
    const OrcLight &l = d.lights[i];
    LightData &ld = s.lights[i];
    ld.type = l.type;
    ld.visible = l.visible != 0;
    const vec3f colored = v3(l.color[0], l.color[1], l.color[2]) * l.intensity; // Light.cpp:20-24
    if (l.type == ORC_LIGHT_QUAD) { // QuadLight.cpp:68-105, QuadLight.ispc:15-33
      ld.position = v3(l.position[0], l.position[1], l.position[2]);
      ld.edge1 = v3(l.edge1[0], l.edge1[1], l.edge1[2]);
      ld.edge2 = v3(l.edge2[0], l.edge2[1], l.edge2[2]);
      const int q = l.intensityQuantity < 0 ? 0 : l.intensityQuantity;
      const float area = length(cross(ld.edge1, ld.edge2));
      if (q == 2) // INTENSITY
        ld.radiance = colored / area;
      else if (q == 3) // POWER
        ld.radiance = colored / (kPi * area);
      else if (q == 0) // RADIANCE
        ld.radiance = colored;
      else
        ld.radiance = v3(0.f);
      const vec3f nd = cross(ld.edge2, ld.edge1);
      ld.ppdf = rcp(length(nd));
      ld.nnormal = nd * ld.ppdf;
    } else if (l.type == ORC_LIGHT_DISTANT) { // DirectionalLight.cpp:20-106
      const vec3f dir = neg(normalize(v3(l.direction[0], l.direction[1], l.direction[2])));
      const float ad = clampf(l.angularDiameter, 0.f, 180.f);
      ld.cosAngle = std::cos(deg2rad(0.5f * ad));
      const int q = l.intensityQuantity < 0 ? 1 : l.intensityQuantity;
      if (q == 0)
        ld.irradiance = (2.0f * kPi * (1.0f - ld.cosAngle)) * colored;
      else if (q == 1)
        ld.irradiance = colored;
      else
        ld.irradiance = v3(0.f);
      ld.frame = frame(dir);
      ld.pdf = ld.cosAngle < kCosAngleMax ? rcp(kTwoPi * (1.0f - ld.cosAngle)) : kInf;
    } else { // AmbientLight.cpp:41-60
      const int q = l.intensityQuantity < 0 ? 0 : l.intensityQuantity;
      if (q == 1)
        ld.radiance = colored / kPi;
      else if (q == 0)
        ld.radiance = colored;
      else
        ld.radiance = v3(0.f);
    }
  }
  s.worldLightsCDF.assign(d.numLights, 1.f);
  if (d.numLights)
    Distribution1D_create(d.numLights, s.worldLightsCDF.data());
  s.lights.insert(s.lights.begin(), std::make_move_iterator(geo.begin()), std::make_move_iterator(geo.end()));
  s.lightsCDF.assign(s.lights.size(), 1.f);
  if (!s.lights.empty())
    Distribution1D_create((int)s.lights.size(), s.lightsCDF.data());
  // volumetric models (VolumetricModel.cpp:31-82, TransferFunction.cpp:19,
  // LinearTransferFunction.cpp:18-31)
  s.volumes.resize(d.numVolumes);
  for (int i = 0; i < d.numVolumes; i++) {
    const OrcVolume &ov = d.volumes[i];
    VolumeData &v = s.volumes[i];
    for (int a = 0; a < 3; a++)
      v.dims[a] = ov.dims[a];
    v.voxels = ov.voxels;
    v.origin = v3(ov.gridOrigin[0], ov.gridOrigin[1], ov.gridOrigin[2]);
    v.spacing = v3(ov.gridSpacing[0], ov.gridSpacing[1], ov.gridSpacing[2]);
    v.upper = v.origin + v3((float)(ov.dims[0] - 1), (float)(ov.dims[1] - 1), (float)(ov.dims[2] - 1)) * v.spacing;
    v.color.resize(ov.numColors);
    for (int k = 0; k < ov.numColors; k++)
      v.color[k] = v3(ov.tfColor[3 * k], ov.tfColor[3 * k + 1], ov.tfColor[3 * k + 2]);
    v.opacity.assign(ov.tfOpacity, ov.tfOpacity + ov.numOpacities);
    v.tfLower = ov.tfValueRange[0];
    v.tfUpper = ov.tfValueRange[1];
    v.densityScale = ov.densityScale;
    v.anisotropy = ov.anisotropy;
    v.userID = ov.userID;
    precomputeMaxOpacity(v);
    buildMacrocells(v);
  }
  s.volInsts.clear();
  for (int i = 0; i < d.numInstances; i++) {
    const OrcGroup &grp = d.groups[d.instances[i].group];
    for (int k = 0; k < grp.numVolumes; k++)
      s.volInsts.push_back({i, d.groupVolumes[grp.firstVolume + k], k});
  }
  s.hasVolumes = !s.volInsts.empty();
}

void setupCamera(CameraData &c, const OrcCamera &d) // PerspectiveCamera.cpp:30-110
{
  const vec3f pos = v3(d.position[0], d.position[1], d.position[2]);
  const vec3f dir = v3(d.direction[0], d.direction[1], d.direction[2]);
  const vec3f up = v3(d.up[0], d.up[1], d.up[2]);
  const float imgY = 2.f * std::tan(deg2rad((ORC_PLANT == 1 ? 1.0f : 0.5f) * d.fovy)); // plant 1: full angle
  const float imgX = imgY * d.aspect;
  c.org = pos;
  c.direction = dir;
  c.dir_00 = normalize(dir);
  c.du_size = normalize(cross(c.dir_00, up));
  c.dv_up = cross(c.du_size, c.dir_00);
  c.du_size = c.du_size * imgX;
  c.dv_up = c.dv_up * imgY;
  c.dir_00 = c.dir_00 - 0.5f * c.du_size - 0.5f * c.dv_up;
  c.nearClip = maxf(d.nearClip, 0.f);
}

void setupRenderer(RendererData &r, const OrcRenderer &d) // Renderer.cpp:38-92, PathTracer.cpp:29-51
{
  r.geometryLights = d.geometryLights != 0;
  r.spp = d.pixelSamples < 1 ? 1 : d.pixelSamples;
  r.maxDepth = d.maxPathLength < 0 ? 0 : d.maxPathLength;
  r.rouletteDepth = d.roulettePathLength;
  r.maxScatteringEvents = d.maxScatteringEvents;
  r.numLightSamples = d.lightSamples;
  r.limitIndirectLightSamples = d.limitIndirectLightSamples != 0;
  r.minContribution = d.minContribution;
  r.maxRadiance = d.maxContribution;
  r.bgColor = {d.backgroundColor[0], d.backgroundColor[1], d.backgroundColor[2], d.backgroundColor[3]};
  r.errorThreshold = d.varianceThreshold;
  r.lutSize = 0;
  if (d.pixelFilter == 0) {
    r.filterType = 0;
    r.filterWidth = 0.f;
  } else if (d.pixelFilter == 1) {
    r.filterType = 1;
    r.filterWidth = 1.f;
  } else {
    // Gaussian LUT, width 3, 5 bins per pixel (PixelFilter.cpp:19-56,
    // LUTPixelFilter.ispc:17-120).  Other LUT filters are not restated.
    r.filterType = 2;
    r.filterWidth = 3.f;
    const int n = (int)std::ceil(r.filterWidth) * 5;
    r.lutSize = n;
    r.lutCdfX.assign(n * n, 0.f);
    r.lutCdfY.assign(n, 0.f);
    const float sigma = (0.5f * r.filterWidth) / (ORC_PLANT == 2 ? 2.0f : 3.0f); // plant 2: wrong sigma
    const float sigmaSqr = sigma * sigma;
    for (int y = 0; y < n; y++) {
      const float fy = (y + 0.5f) * (1.0f / (float)n);
      for (int x = 0; x < n; x++) {
        const float fx = (x + 0.5f) * (1.0f / (float)n);
        const float cx = (fx - 0.5f) * r.filterWidth, cy = (fy - 0.5f) * r.filterWidth;
        const float distSqr = cx * cx + cy * cy;
        const float e = fexp(-0.5f * (distSqr / sigmaSqr));
        r.lutCdfX[y * n + x] = maxf(0.0f, e);
      }
      r.lutCdfY[y] = Distribution1D_create(n, r.lutCdfX.data() + y * n);
    }
    Distribution1D_create(n, r.lutCdfY.data());
  }
}

void FrameBuffer::clear() // FrameBuffer.cpp:95-103
{
  frameID = -1;
  mtGen.seed(43);
  varianceValid = false;
}

void initFrameBuffer(FrameBuffer &fb, int w, int h, int format, uint32_t ch)
{
  fb.w = w;
  fb.h = h;
  fb.format = format;
  if (!(ch & 4u)) // OSP_FB_ACCUM; API.cpp:568-582 drops VARIANCE without ACCUM
    ch &= ~8u;
  fb.channels = ch;
  fb.targetFrames = (ch & 4u) ? 0 : 1; // FrameBuffer.cpp:85-93
  const size_t n = (size_t)w * h;
  if (ch & 1u)
    fb.color.assign(n, {0, 0, 0, 0});
  if (ch & 8u)
    fb.variance.assign(n, {0, 0, 0, 0});
  if (ch & 2u)
    fb.depth.assign(n, kInf);
  if (ch & 16u)
    fb.normal.assign(n, v3(0.f));
  if (ch & 32u)
    fb.albedo.assign(n, v3(0.f));
  if (ch & 64u)
    fb.primID.assign(n, kInvalidID);
  if (ch & 128u)
    fb.objID.assign(n, kInvalidID);
  if (ch & 256u)
    fb.instID.assign(n, kInvalidID);
  if (ch & 512u)
    fb.firstNormal.assign(n, v3(0.f));
  if (ch & 1024u)
    fb.position.assign(n, v3(0.f));
  const int tx = (w + 7) / 8, ty = (h + 7) / 8;
  fb.taskVariance.assign((size_t)tx * ty, 0.f);
  fb.clear();
}

// ---- camera and pixel filter ---------------------------------------------------

vec2f pixelFilterSample(const RendererData &r, vec2f s) // PixelFilterDispatch.ispc:7-19
{
  if (r.filterType == 0)
    return v2(0.f, 0.f);
  if (r.filterType == 1)
    return v2((s.x + -0.5f) * r.filterWidth, (s.y + -0.5f) * r.filterWidth);
  // Distribution2D_sample (math/Distribution2D.ispc:9-25)
  const int n = r.lutSize;
  const Sample1D sy = Distribution1D_sample(n, r.lutCdfY.data(), 0, s.y);
  const int x0 = sy.idx * n;
  const Sample1D sx = Distribution1D_sample(n, r.lutCdfX.data(), x0, s.x);
  const float rcpN = 1.f / (float)n;
  const vec2f uv = v2(((float)sx.idx + sx.frac) * rcpN, ((float)sy.idx + sy.frac) * rcpN);
  return v2((uv.x + -0.5f) * r.filterWidth, (uv.y + -0.5f) * r.filterWidth);
}

void cameraRay(const CameraData &c, vec2f screen, Ray &ray) // PerspectiveCamera.ispc:16-64
{
  // Camera_subRegion with the default image region [0,1]^2
  const float sx = (1.f - screen.x) * 0.f + screen.x * 1.f;
  const float sy = (1.f - screen.y) * 0.f + screen.y * 1.f;
  const vec3f dir = c.dir_00 + sx * c.du_size + sy * c.dv_up;
  ray.org = c.org;
  ray.dir = normalize(dir);
  ray.t0 = c.nearClip;
  ray.t = kInf;
}

// ---- post-intersect (common/World.ih:389-465, Instance.ih:44-105,
//      GeometricModel.ih:46-76, Mesh.ispc:65-268, Spheres.ispc:24-42) -----------

static void postIntersect(const Scene &s, const Ray &ray, const Hit &hit, DG &dg, bool tangents)
{
  const World &w = s.world;
  const OrcScene &d = w.desc;
  const OrcInstance &inst = d.instances[hit.inst];
  const OrcGroup &grp = d.groups[inst.group];
  const int om = d.groupModels[grp.firstModel + hit.model];
  const OrcModel &model = d.models[om];
  const OrcGeometry &g = d.geometries[model.geometry];

  dg.primID = hit.prim;
  dg.areaPDF = s.modelAreaPDF[om];
  dg.material = -1;
  dg.color = {1.f, 1.f, 1.f, 1.f};
  dg.dPds = v3(1.f, 0.f, 0.f);
  dg.dPdt = v3(0.f, 1.f, 0.f);
  dg.P = ray.org + hit.t * ray.dir;
  dg.epsilon = 0.f;

  // geometry
  dg.Ng = dg.Ns = hit.Ng;
  if (g.type == ORC_GEOM_TRIANGLES) {
    const uint32_t *ix = g.index + 3 * (int64_t)hit.prim;
    const float u = hit.u, v = hit.v;
    const vec3f uv = v3(1.0f - u - v, u, v);
    if (g.normal) {
      const vec3f a = ld3(g.normal, ix[0]), b = ld3(g.normal, ix[1]), c = ld3(g.normal, ix[2]);
      dg.Ns = uv.x * a + uv.y * b + uv.z * c;
    }
    const vec3f va = ld3(g.vertex, ix[0]), vb = ld3(g.vertex, ix[1]), vc = ld3(g.vertex, ix[2]);
    dg.epsilon = reduce_max(vmax(vmax(vabs(va), vabs(vb)), vabs(vc))) * kUlpEpsilon;
    if (tangents) {
      dg.dPds = vb - va;
      dg.dPdt = vc - va;
    }
  } else if (g.type == ORC_GEOM_QUADS) {
    const uint32_t *ix = g.index + 4 * (int64_t)hit.prim;
    const float u = hit.u, v = hit.v;
    vec4f uv;
    if (ix[2] == ix[3])
      uv = {1.f - u - v, u, v, 0.f};
    else
      uv = {(1 - v) * (1 - u), (1 - v) * u, v * u, v * (1 - u)};
    if (g.normal) {
      const vec3f a = ld3(g.normal, ix[0]), b = ld3(g.normal, ix[1]), c = ld3(g.normal, ix[2]),
                  e = ld3(g.normal, ix[3]);
      dg.Ns = uv.x * a + uv.y * b + uv.z * c + uv.w * e;
    }
    const vec3f va = ld3(g.vertex, ix[0]), vb = ld3(g.vertex, ix[1]), vc = ld3(g.vertex, ix[2]),
                vd = ld3(g.vertex, ix[3]);
    dg.epsilon = reduce_max(vmax(vmax(vabs(va), vabs(vb)), vmax(vabs(vc), vabs(vd)))) * kUlpEpsilon;
    if (tangents) {
      dg.dPds = vb - va;
      dg.dPdt = vd - va;
    }
  } else {
    const float *sp = g.spheres + 4 * (int64_t)hit.prim;
    dg.epsilon = sp[3] * kUlpEpsilon;
  }
  // geometric model
  dg.objID = model.userID == kInvalidID ? hit.model : model.userID;
  dg.material = model.material;
  if (model.invertNormals)
    dg.Ns = neg(dg.Ns);
  // instance
  dg.instID = inst.userID == kInvalidID ? hit.inst : inst.userID;
  const affine3f &x = w.xfm[hit.inst];
  const affine3f &rx = w.rcpXfm[hit.inst];
  dg.Ns = mulT(rx.l, dg.Ns);
  dg.Ng = mulT(rx.l, dg.Ng);
  dg.epsilon *= maxf(std::fabs(x.l.vx.x), maxf(std::fabs(x.l.vy.y), std::fabs(x.l.vz.z)));
  if (tangents) {
    dg.dPds = xfmVector(x, dg.dPds);
    dg.dPdt = xfmVector(x, dg.dPdt);
  }
  // world
  const float tEps = reduce_max(vabs(ray.dir)) * hit.t;
  dg.epsilon = maxf(dg.epsilon, maxf(reduce_max(vabs(dg.P)), tEps) * kUlpEpsilon);
  vec3f ffnng = normalize(dg.Ng);
  dg.Ng = ffnng;
  dg.Ns = normalize(dg.Ns);
  const bool flip = dot(ray.dir, dg.Ng) >= 0.f;
  if (flip) {
    ffnng = neg(ffnng);
    dg.Ng = neg(dg.Ng);
  }
  if (dot(dg.Ng, dg.Ns) < 0.f)
    dg.Ns = neg(dg.Ns);
  dg.P = dg.P + dg.epsilon * ffnng;
}

// ---- BSDFs -----------------------------------------------------------------------

struct BSDF
{
  int kind; // 0 none, 1 OBJ (multi), 2 conductor, 3 microfacet conductor, 4 dielectric
  uint32_t scatteringType;
  vec3f albedo;
  linear3f frame;
  // OBJ children: 0 Lambert, 1 Transmission, 2 Specular
  float weight[3], importance[3], importanceSum;
  vec3f childAlbedo[3];
  // Specular (Specular.ih:14-19): normalised reflectivity, sample weight,
  // Blinn exponent
  vec3f specKs, specWeight;
  float specNs;
  // conductors
  vec3f eta, k;
  const MaterialData *spectralMat; // a spectral conductor, else null
  float roughness;
  vec2f alpha;
  // dielectric
  float etaRel;
  vec3f Ng, Ns;
};

struct EvalRes
{
  vec3f value;
  float pdf;
};
struct SampleRes
{
  vec3f weight;
  vec3f wi;
  float pdf;
  uint32_t type;
};

// ---- the OBJ material's specular lobe: a normalised Blinn-Phong BRDF
// f = r (n+2) / (4 pi (2 - 2^(-n/2))) cos(HN)^n / cos(LN) with the Phong
// exponent converted to Blinn (x4), sampled by the power-cosine
// distribution of the half vector (bsdfs/Specular.ih:14-78,
// math/sampling.ih:70-87)
static constexpr float kOneOverFourPi = 0.0795774715459476678844f;
static constexpr float kLnTwoOverTwo = 0.346573590279972654709f;

static float powerCosineSampleHemispherePDF(float cosTheta, float n)
{
  return (n + 1.0f) * kOneOverTwoPi * fpow(cosTheta, n);
}

static vec3f powerCosineSampleHemisphere(float n, vec2f s)
{
  const float phi = kTwoPi * s.x;
  const float cosTheta = fpow(s.y, 1.0f / (n + 1.0f));
  return cartesian(phi, cos2sin(cosTheta), cosTheta);
}

// Specular_Constructor and its MultiBSDF_add as child 2 (OBJ.ispc:86-95)
static void specularLobe(const MaterialData &m, float d, BSDF &b)
{
  const float Ns = m.Ns;
  const vec3f Ks = d * m.Ks;
  if (!(reduce_max(Ks) > 0.0f))
    return;
  b.specNs = Ns * 4.f;
  const float tmp = (b.specNs + 2.f) * rcp(2.f - fexp(b.specNs * -kLnTwoOverTwo));
  b.specKs = Ks * (kOneOverFourPi * tmp);
  b.specWeight = Ks * (2.f * tmp * rcp(b.specNs + 1.f));
  b.childAlbedo[2] = Ks;
  b.albedo = b.albedo + Ks * 1.f;
  b.scatteringType |= GLOSSY_REFL;
  b.weight[2] = 1.f;
  b.importance[2] = luminance(Ks);
  b.importanceSum += b.importance[2];
}

static EvalRes specularEval(const BSDF &b, vec3f wo, vec3f wi) // Specular.ih:23-41
{
  const vec3f N = b.frame.vz;
  const float cosThetaO = dot(wo, N);
  const float cosThetaI = dot(wi, N);
  if (cosThetaO <= 0.0f || cosThetaI <= 0.0f)
    return {v3(0.f), 0.f};
  const vec3f wh = normalize(wo + wi);
  const float cosThetaH = dot(wh, N);
  EvalRes res;
  res.pdf = powerCosineSampleHemispherePDF(cosThetaH, b.specNs) * rcp(4.0f * dot(wo, wh));
  res.value = b.specKs * fpow(cosThetaH, b.specNs);
  return res;
}

static SampleRes specularSample(const BSDF &b, vec3f wo, vec2f s) // Specular.ih:43-66
{
  SampleRes res;
  const vec3f whLocal = powerCosineSampleHemisphere(b.specNs, s);
  const vec3f wh = mul(b.frame, whLocal);
  const float cosThetaOH = dot(wo, wh);
  res.wi = (2.0f * cosThetaOH) * wh - wo; // reflect(wo, wh, cosThetaOH), Optics.ih:13-16
  const float cosThetaI = dot(res.wi, b.frame.vz);
  if (cosThetaI <= 0.0f) {
    res.weight = v3(0.f);
    res.wi = v3(0.f);
    res.pdf = 0.f;
    res.type = 0;
    return res;
  }
  const float cosThetaH = whLocal.z;
  res.pdf = powerCosineSampleHemispherePDF(cosThetaH, b.specNs) * rcp(4.0f * cosThetaOH);
  res.type = GLOSSY_REFL;
  res.weight = b.specWeight * cosThetaOH;
  return res;
}

static float fresnelConductor1(float cosI, float eta, float k) // Fresnel.ih:75-84
{
  const float tmp = sqr(eta) + sqr(k);
  const float Rpar = (tmp * sqr(cosI) - eta * (2.0f * cosI) + 1.f) * rcp(tmp * sqr(cosI) + eta * (2.0f * cosI) + 1.f);
  const float Rper = (tmp - 2.0f * eta * cosI + sqr(cosI)) * rcp(tmp + 2.0f * eta * cosI + sqr(cosI));
  return 0.5f * (Rpar + Rper);
}
// spectrum_sRGB (math/spectrum.ispc:11-38): linear sRGB of each of the 8
// wavelength samples
static const float kSpectrumRGB[3][8] = {
    {0.0598548f, -0.0234574f, -0.220138f, -0.238902f, 0.316327f, 0.738315f, 0.323302f, 0.0446981f},
    {-0.0567346f, -0.0160361f, 0.223861f, 0.531185f, 0.337221f, 0.0149718f, -0.0296053f, -0.00486239f},
    {0.420693f, 0.616597f, 0.0796766f, -0.0496266f, -0.0473149f, -0.0167536f, -0.00295686f, -0.000314818f}};

// Fresnel_dispatch_eval for FresnelConductorRGBUniform (Fresnel.ih:131-141)
// and FresnelConductorSpectral (Fresnel.ih:83-90)
static vec3f fresnelRGB(const BSDF &b, float cosI)
{
  if (b.spectralMat) {
    vec3f rgb = v3(0.f);
    for (int l = 0; l < 8; l++)
      rgb = rgb + fresnelConductor1(cosI, b.spectralMat->etaS[l], b.spectralMat->kS[l]) *
          v3(kSpectrumRGB[0][l], kSpectrumRGB[1][l], kSpectrumRGB[2][l]);
    return v3(clampf(rgb.x), clampf(rgb.y), clampf(rgb.z));
  }
  return v3(fresnelConductor1(cosI, b.eta.x, b.k.x), fresnelConductor1(cosI, b.eta.y, b.k.y),
      fresnelConductor1(cosI, b.eta.z, b.k.z));
}
static float fresnelDielectric(float cosI, float eta) // Fresnel.ih:19-40
{
  const float sqrcost = 1.0f - sqr(eta) * (1.0f - sqr(cosI));
  if (sqrcost < 0.0f)
    return 1.0f;
  const float cosT = std::sqrt(sqrcost);
  const float Rper = (eta * cosI - cosT) * rcp(eta * cosI + cosT);
  const float Rpar = (cosI - eta * cosT) * rcp(cosI + eta * cosT);
  return 0.5f * (sqr(Rpar) + sqr(Rper));
}

// GGX (bsdfs/GGXDistribution.ih)
static float ggxEval(vec2f a, vec3f wh)
{
  const float cosTheta = wh.z;
  const float cosTheta2 = sqr(cosTheta);
  const float e = (sqr(wh.x / a.x) + sqr(wh.y / a.y)) / cosTheta2;
  return rcp(kPi * a.x * a.y * sqr(cosTheta2 * (1.f + e)));
}
static float ggxLambda(vec2f a, vec3f wo)
{
  const float c2 = sqr(wo.z);
  const float invA2 = (sqr(wo.x * a.x) + sqr(wo.y * a.y)) / c2;
  return 0.5f * (-1.f + std::sqrt(1.f + invA2));
}
static float ggxG1(vec2f a, vec3f wo, float cosThetaOH)
{
  if (wo.z * cosThetaOH <= 0.f)
    return 0.f;
  return rcp(1.f + ggxLambda(a, wo));
}
static float ggxG2(vec2f a, vec3f wo, vec3f wi, float cosOH, float cosIH)
{
  if (wo.z * cosOH <= 0.f || wi.z * cosIH <= 0.f)
    return 0.f;
  return rcp(1.f + ggxLambda(a, wo) + ggxLambda(a, wi));
}
static float ggxEvalVisible(vec2f a, vec3f wh, vec3f wo, float cosOH, float &pdf)
{
  const float D = ggxEval(a, wh);
  pdf = ggxG1(a, wo, cosOH) * std::fabs(cosOH) * D / std::fabs(wo.z);
  return D;
}
static vec3f ggxSampleVisible(vec2f a, vec3f wo, float &pdf, vec2f s)
{
  const vec3f V = normalize(v3(a.x * wo.x, a.y * wo.y, wo.z));
  const vec3f T1 = (V.z < 0.9999f) ? normalize(cross(V, v3(0.f, 0.f, 1.f))) : v3(1.f, 0.f, 0.f);
  const vec3f T2 = cross(T1, V);
  const float aa = 1.f / (1.f + V.z);
  const float r = std::sqrt(s.x);
  const float phi = (s.y < aa) ? s.y / aa * kPi : kPi + (s.y - aa) / (1.f - aa) * kPi;
  float sinPhi, cosPhi;
  fsincos(phi, sinPhi, cosPhi);
  const float P1 = r * cosPhi;
  const float P2 = r * sinPhi * ((s.y < aa) ? 1.f : V.z);
  vec3f wh = P1 * T1 + P2 * T2 + std::sqrt(maxf(0.f, 1.f - P1 * P1 - P2 * P2)) * V;
  wh = normalize(v3(a.x * wh.x, a.y * wh.y, maxf(0.f, wh.z)));
  const float cwh = dot(wo, wh);
  pdf = ggxG1(a, wo, cwh) * std::fabs(cwh) * ggxEval(a, wh) / std::fabs(wo.z);
  return wh;
}

// Material_dispatch_getBSDF (render/MaterialDispatch.ispc:23-72)
static bool getBSDF(const Scene &s, const DG &dg, const Medium &medium, BSDF &b)
{
  if (dg.material < 0)
    return false;
  const MaterialData &m = s.materials[dg.material];
  if (m.type == ORC_MAT_LUMINOUS) { // Luminous.ispc:15-32: Transmission or no BSDF
    if (!(m.transparency > 0.0f))
      return false;
    b.kind = 5;
    b.albedo = v3(m.transparency);
    b.scatteringType = SPEC_TRANS;
    b.frame = frame(dg.Ns);
    return true;
  }
  if (m.type == ORC_MAT_OBJ) { // OBJ.ispc:29-98
    b.kind = 1;
    b.albedo = v3(0.f);
    b.scatteringType = 0;
    b.importanceSum = 0.f;
    for (int i = 0; i < 3; i++)
      b.weight[i] = b.importance[i] = 0.f;
    // makeShadingFrame (render/Material.ih:25-37)
    {
      const vec3f N = dg.Ns;
      vec3f U = normalize(dg.dPds);
      vec3f V = cross(N, U);
      if (dot(V, V) > 0.f) {
        V = normalize(V);
        U = cross(V, N);
        b.frame = {U, V, N};
      } else {
        b.frame = frame(N);
      }
    }
    const float d = m.d * dg.color.w;
    const vec3f Kd = m.Kd * d * v3(dg.color.x, dg.color.y, dg.color.z);
    if (reduce_max(Kd) > 0.0f) {
      b.childAlbedo[0] = Kd;
      b.albedo = b.albedo + Kd * 1.f;
      b.scatteringType |= DIFF_REFL;
      b.weight[0] = 1.f;
      b.importance[0] = luminance(Kd);
      b.importanceSum += b.importance[0];
    }
    const vec3f T = m.Tf * d + v3(1.f - d);
    if (reduce_max(T) > 0.0f) {
      b.childAlbedo[1] = T;
      b.albedo = b.albedo + T * 1.f;
      b.scatteringType |= SPEC_TRANS;
      b.weight[1] = 1.f;
      b.importance[1] = luminance(T);
      b.importanceSum += b.importance[1];
    }
    specularLobe(m, d, b);
    return true;
  }
  if (m.type == ORC_MAT_METAL) { // Metal.ispc:18-46
    b.frame = frame(dg.Ns);
    b.eta = m.eta;
    b.spectralMat = m.spectral ? &m : nullptr;
    b.k = m.k;
    // FresnelConductorRGBUniform has no evalAvg (Fresnel.ih:307-319): Favg = 0,
    // hence albedo 0 and no Kulla-Conty energy compensation term.
    b.albedo = v3(0.f);
    if (m.roughness == 0.0f) {
      b.kind = 2;
      b.scatteringType = SPEC_REFL;
    } else {
      b.kind = 3;
      b.scatteringType = GLOSSY_REFL;
      b.roughness = m.roughness;
      // roughnessToAlpha(roughness, 0) (MicrofacetDistribution.ih:17-22)
      const float aspect = std::sqrt(1.f - 0.9f * 0.f);
      b.alpha = v2(maxf(sqr(m.roughness) / aspect, 0.001f), maxf(sqr(m.roughness) * aspect, 0.001f));
    }
    return true;
  }
  // Glass.ispc:18-35 -> RobustDielectric (bsdfs/RobustDielectric.ih:162-185)
  b.kind = 4;
  b.etaRel = eq(medium, m.outside) ? m.outside.ior * rcp(m.inside.ior) : m.inside.ior * rcp(m.outside.ior);
  b.frame = frame(dg.Ns);
  b.albedo = v3(1.f);
  b.scatteringType = S_SPECULAR;
  b.Ns = dg.Ns;
  b.Ng = dg.Ng;
  return true;
}

static EvalRes lambertEval(const BSDF &b, vec3f albedo, vec3f wi) // Lambert.ih:11-19
{
  const float cosThetaI = maxf(dot(wi, b.frame.vz), 0.f);
  return {albedo * kOneOverPi * cosThetaI, cosThetaI * kOneOverPi};
}

// BSDF_dispatch_eval (bsdfs/BSDF.ispc:22-56)
static EvalRes bsdfEval(const BSDF &b, vec3f wo, vec3f wi)
{
  EvalRes z{v3(0.f), 0.f};
  if (b.kind == 1) { // MULTIBSDF_EVAL (MultiBSDF.ih)
    EvalRes e = z;
    if (b.importance[0] > 0.f) {
      EvalRes cur = lambertEval(b, b.childAlbedo[0], wi);
      cur.value = cur.value * b.weight[0];
      e.value = e.value + cur.value;
      e.pdf += cur.pdf * b.importance[0];
    }
    // child 1 (transmission) has no evaluation (OBJ.ispc:120-130)
    if (b.importance[2] > 0.f) {
      EvalRes cur = specularEval(b, wo, wi);
      cur.value = cur.value * b.weight[2];
      e.value = e.value + cur.value;
      e.pdf += cur.pdf * b.importance[2];
    }
    e.pdf *= rcp(b.importanceSum);
    return e;
  }
  if (b.kind == 3) { // MicrofacetConductor_eval (MicrofacetConductor.ih:27-67)
    const vec3f N = b.frame.vz;
    const float cosThetaO = dot(wo, N);
    const float cosThetaI = dot(wi, N);
    if (cosThetaO <= 0.f || cosThetaI <= 0.f)
      return z;
    const vec3f wh = normalize(wi + wo);
    const float cosThetaOH = dot(wo, wh);
    const float cosThetaIH = dot(wi, wh);
    const vec3f wo0 = mulT(b.frame, wo), wi0 = mulT(b.frame, wi), wh0 = mulT(b.frame, wh);
    const vec3f F = fresnelRGB(b, cosThetaOH);
    float whPdf;
    const float D = ggxEvalVisible(b.alpha, wh0, wo0, cosThetaOH, whPdf);
    const float G = ggxG2(b.alpha, wo0, wi0, cosThetaOH, cosThetaIH);
    const vec3f fms = v3(0.f);
    EvalRes r;
    r.pdf = whPdf * rcp(4.f * std::fabs(cosThetaOH));
    r.value = F * (D * G * rcp(4.f * cosThetaO)) + fms;
    return r;
  }
  return z;
}

// BSDF_dispatch_sample (bsdfs/BSDF.ispc:61-112)
static SampleRes bsdfSample(const BSDF &b, vec3f wo, vec2f s, float ss)
{
  if (b.kind == 5) { // Transmission_sample (bsdfs/Transmission.ih:11-20)
    SampleRes res;
    res.wi = neg(wo);
    res.pdf = kInf;
    res.type = SPEC_TRANS;
    res.weight = b.albedo;
    return res;
  }
  SampleRes zero;
  zero.weight = v3(0.f);
  zero.wi = v3(0.f);
  zero.pdf = 0.f;
  zero.type = 0;
  if (b.kind == 1) { // MULTIBSDF_SAMPLE (MultiBSDF.ih)
    SampleRes res = zero;
    const float x = ss * b.importanceSum;
    int choice = 0;
    float prefixSum = 0.f;
    if (b.importanceSum == 0.f)
      return res;
    if (x < (prefixSum += b.importance[0])) {
      choice = 0;
      ss = (x + b.importance[0] - prefixSum) * rcp(b.importance[0]);
      const vec3f localDir = cosineSampleHemisphere(s); // Lambert_sample, Lambert.ih:21-30
      res.wi = mul(b.frame, localDir);
      res.pdf = localDir.z * kOneOverPi;
      res.type = DIFF_REFL;
      res.weight = b.childAlbedo[0];
      res.weight = res.weight * b.weight[0];
    } else if (x < (prefixSum += b.importance[1])) {
      choice = 1;
      ss = (x + b.importance[1] - prefixSum) * rcp(b.importance[1]);
      res.wi = neg(wo); // Transmission_sample, Transmission.ih:11-20
      res.pdf = kInf;
      res.type = SPEC_TRANS;
      res.weight = b.childAlbedo[1];
      res.weight = res.weight * b.weight[1];
    } else if (x < (prefixSum += b.importance[2])) {
      choice = 2;
      ss = (x + b.importance[2] - prefixSum) * rcp(b.importance[2]);
      res = specularSample(b, wo, s);
      res.weight = res.weight * b.weight[2];
    }
    if (eq(res.weight, v3(0.f)) || res.pdf == 0.f)
      return zero;
    if (res.type & S_SPECULAR) {
      res.weight = res.weight * (b.importanceSum * rcp(b.importance[choice]));
      return res;
    }
    EvalRes e;
    e.value = res.weight * res.pdf;
    res.pdf *= b.importance[choice];
    e.pdf = res.pdf;
    const vec3f wi = res.wi;
    if (b.importance[0] > 0.f && choice != 0) {
      EvalRes cur = lambertEval(b, b.childAlbedo[0], wi);
      cur.value = cur.value * b.weight[0];
      e.value = e.value + cur.value;
      e.pdf += cur.pdf * b.importance[0];
    }
    if (b.importance[2] > 0.f && choice != 2) {
      EvalRes cur = specularEval(b, wo, wi);
      cur.value = cur.value * b.weight[2];
      e.value = e.value + cur.value;
      e.pdf += cur.pdf * b.importance[2];
    }
    res.pdf = e.pdf;
    res.pdf *= rcp(b.importanceSum);
    res.weight = e.value * rcp(res.pdf);
    return res;
  }
  if (b.kind == 2) { // Conductor_sample (Conductor.ih:17-27)
    SampleRes res;
    const vec3f N = b.frame.vz;
    res.wi = reflect(wo, N);
    res.pdf = kInf;
    res.type = SPEC_REFL;
    res.weight = fresnelRGB(b, dot(wo, N));
    return res;
  }
  if (b.kind == 3) { // MicrofacetConductor_sample (MicrofacetConductor.ih:69-110)
    const vec3f N = b.frame.vz;
    const float cosThetaO = dot(wo, N);
    if (cosThetaO <= 0.f)
      return zero;
    const vec3f wo0 = mulT(b.frame, wo);
    float whPdf;
    const vec3f wh = mul(b.frame, ggxSampleVisible(b.alpha, wo0, whPdf, s));
    SampleRes res;
    res.wi = reflect(wo, wh);
    const float cosThetaI = dot(res.wi, N);
    if (cosThetaI <= 0.f)
      return zero;
    const float cosThetaOH = dot(wo, wh);
    const float cosThetaIH = dot(res.wi, wh);
    const vec3f wi0 = mulT(b.frame, res.wi);
    const vec3f F = fresnelRGB(b, cosThetaOH);
    const float G = ggxG2(b.alpha, wo0, wi0, cosThetaOH, cosThetaIH);
    const vec3f fms = v3(0.f);
    res.type = GLOSSY_REFL;
    res.pdf = whPdf * rcp(4.f * std::fabs(cosThetaOH));
    res.weight = F * (G * rcp_safe(ggxG1(b.alpha, wo0, cosThetaOH))) + (fms * rcp(res.pdf));
    return res;
  }
  if (b.kind == 4) { // RobustDielectric_sample (RobustDielectric.ih:87-160)
    SampleRes res;
    res.pdf = kInf;
    const vec3f incomingDir = neg(wo);
    vec3f shadingNormal = b.Ns;
    vec3f geometricNormal = b.Ng;
    const float eta = 1.f / b.etaRel;
    const float glassEps = 1e-6f;
    const bool canReflect = (eta < 1.f - glassEps || eta > 1.f + glassEps);
    if (dot(incomingDir, shadingNormal) > 0.f)
      shadingNormal = neg(shadingNormal);
    if (dot(incomingDir, geometricNormal) > 0.f)
      geometricNormal = neg(geometricNormal);
    // ReflectionDirection (RobustDielectric.ih:73-85)
    const vec3f par = dot(incomingDir, shadingNormal) * shadingNormal;
    const vec3f dirR = (incomingDir - par) - par;
    // RefractionDirection (RobustDielectric.ih:31-71)
    vec3f dirT;
    float FresnelR;
    bool mustReflect = false;
    {
      const float cos1 = dot(incomingDir, shadingNormal);
      float cos1_abs = cos1;
      if (cos1_abs < 0.f)
        cos1_abs = -cos1_abs;
      const bool headOn = (1.f - 1e-6f < cos1_abs && cos1_abs < 1.f - 1e-6f);
      if (headOn) {
        FresnelR = sqr((eta - 1.f) / (eta + 1.f));
        dirT = incomingDir;
      } else {
        const vec3f dirPar = cos1 * shadingNormal;
        const vec3f dirPerp = incomingDir - dirPar;
        const vec3f bb = normalize(dirPerp);
        const float sin1 = dot(dirPerp, bb);
        const float sin2 = sin1 / eta;
        if (sin2 > 1.0f) {
          dirT = v3(0.f);
          FresnelR = 1.f;
          mustReflect = true;
        } else {
          const float cos2 = std::sqrt(clampf(1.f - sin2 * sin2));
          const vec3f N2 = normalize(dirPar);
          dirT = N2 * cos2 + bb * sin2;
          const float c1 = cos1_abs, c2 = cos2;
          const float r1 = (eta * c1 - c2) / (eta * c1 + c2);
          const float r2 = (c1 - eta * c2) / (c1 + eta * c2);
          FresnelR = 0.5f * (r1 * r1 + r2 * r2);
        }
      }
      FresnelR = clampf(FresnelR);
    }
    res.weight = v3(1.f);
    if ((ss < FresnelR && canReflect) || mustReflect) {
      res.wi = dirR;
      res.type = SPEC_REFL;
      const float gc = dot(res.wi, geometricNormal);
      if (gc >= 0.f) {
        if (gc < glassEps) {
          res.wi = res.wi + glassEps * geometricNormal;
          res.wi = normalize(res.wi);
        }
      } else {
        res.wi = res.wi - 2.f * geometricNormal * gc;
      }
    } else {
      res.wi = dirT;
      res.type = SPEC_TRANS;
      const float gc = dot(res.wi, geometricNormal);
      if (gc <= 0.f) {
        if (gc > -glassEps) {
          res.wi = res.wi - glassEps * geometricNormal;
          res.wi = normalize(res.wi);
        }
      } else {
        res.type = SPEC_REFL;
      }
    }
    return res;
  }
  return zero;
}

// Material_dispatch_getTransparency (render/MaterialDispatch.ispc:72-104)
static vec3f getTransparency(const Scene &s, const DG &dg, const Ray &ray, const Medium &medium)
{
  if (dg.material < 0)
    return v3(0.f);
  const MaterialData &m = s.materials[dg.material];
  if (m.type == ORC_MAT_OBJ) { // OBJ.ispc:100-117
    const float d = m.d * dg.color.w;
    return m.Tf * d + v3(1.f - d);
  }
  if (m.type == ORC_MAT_LUMINOUS) // Luminous.ispc:34-44
    return v3(m.transparency);
  if (m.type == ORC_MAT_GLASS) { // Glass.ispc:37-52
    const float eta = eq(medium, m.outside) ? m.outside.ior * rcp(m.inside.ior) : m.inside.ior * rcp(m.outside.ior);
    const float cosThetaO = ORC_PLANT == 6 ? 1.0f : maxf(-dot(ray.dir, dg.Ns), 0.0f); // plant 6: normal incidence
    return v3(1.0f - fresnelDielectric(cosThetaO, eta));
  }
  return v3(0.f);
}

static void selectNextMedium(const Scene &s, const DG &dg, Medium &medium) // Glass.ispc:54-64
{
  if (dg.material < 0)
    return;
  const MaterialData &m = s.materials[dg.material];
  if (m.type != ORC_MAT_GLASS)
    return;
  medium = eq(medium, m.outside) ? m.inside : m.outside;
}

// ---- lights (lights/*.ispc) ---------------------------------------------------------

struct LightSample
{
  vec3f weight, dir;
  float dist, pdf;
};
struct LightEval
{
  vec3f radiance;
  float pdf;
};

struct SphericalQuad
{
  vec3f e0, e1, z;
  float S, k, n0z, n0z2, n2z, x0, x1, y0, y02, y1, z0, z02;
};

static SphericalQuad sphericalQuadCreate(vec3f q0, vec3f e0, vec3f e1, vec3f n, vec3f P) // SphericalQuadSampling.ih:15-78
{
  const float e0l = length(e0);
  const float e1l = length(e1);
  SphericalQuad q;
  q.e0 = e0 / e0l;
  q.e1 = e1 / e1l;
  q.z = n;
  const vec3f d = q0 - P;
  q.z0 = dot(d, q.z);
  q.z = q.z0 > 0.f ? neg(q.z) : q.z;
  q.z0 = -std::fabs(q.z0);
  q.z02 = sqr(q.z0);
  q.x0 = dot(d, q.e0);
  q.y0 = dot(d, q.e1);
  q.y02 = sqr(q.y0);
  q.x1 = q.x0 + e0l;
  q.y1 = q.y0 + e1l;
  const vec3f v00 = v3(q.x0, q.y0, q.z0), v01 = v3(q.x0, q.y1, q.z0), v10 = v3(q.x1, q.y0, q.z0),
              v11 = v3(q.x1, q.y1, q.z0);
  const vec3f n0 = normalize(cross(v00, v10));
  q.n0z = n0.z;
  q.n0z2 = n0.z * n0.z;
  const vec3f n1 = normalize(cross(v10, v11));
  const vec3f n2 = normalize(cross(v11, v01));
  q.n2z = n2.z;
  const vec3f n3 = normalize(cross(v01, v00));
  const float g0 = facos(clampf(-dot(n0, n1), -1.f, 1.f));
  const float g1 = facos(clampf(-dot(n1, n2), -1.f, 1.f));
  const float g2 = facos(clampf(-dot(n2, n3), -1.f, 1.f));
  const float g3 = facos(clampf(-dot(n3, n0), -1.f, 1.f));
  q.k = kTwoPi - g2 - g3;
  q.S = maxf(0.f, g0 + g1 - q.k);
  return q;
}

static inline float signf(float x) { return (float)(0.f < x) - (float)(x < 0.f); }

static vec3f sampleSphericalQuad(const SphericalQuad &q, vec2f s) // SphericalQuadSampling.ih:86-111
{
  const float au = s.x * q.S + q.k;
  float sinau, cosau;
  fsincos(au, sinau, cosau);
  const float fu = (cosau * q.n0z - q.n2z) * rcp(sinau);
  const float cu = clampf(rsqrt(fu * fu + sqr(q.n0z)) * signf(fu), -1.f, 1.f);
  const float x0 = -(cu * q.z0) * rsqrt(1.f - cu * cu);
  const float xu = clampf(x0, q.x0, q.x1);
  const float ds = std::sqrt(sqr(xu) + sqr(q.z0));
  const float ds2 = sqr(ds);
  const float h0 = q.y0 * rsqrt(ds2 + sqr(q.y0));
  const float h1 = q.y1 * rsqrt(ds2 + sqr(q.y1));
  const float hv = s.y * (h1 - h0) + h0;
  const float hv2 = sqr(hv);
  const float yv = (hv2 < 1.f - 1e-4f) ? (hv * ds) * rsqrt(1.f - hv2) : q.y1;
  return xu * q.e0 + yv * q.e1 + q.z0 * q.z;
}

static float solveQuadratic(float A, float B, float C) // SphericalQuadSampling.ih:113-135
{
  if (A == 0.f)
    return -C / B;
  float rad = B * B - 4.f * A * C;
  if (rad < 0.f)
    return 0.f;
  rad = std::sqrt(rad);
  float root;
  if (B >= 0.f) {
    root = (-B - rad) / (2.f * A);
    if (root < 0.f || root > 1.f)
      root = 2.f * C / (-B - rad);
  } else {
    root = 2.f * C / (-B + rad);
    if (root < 0.f || root > 1.f)
      root = (-B + rad) / (2.f * A);
  }
  return root;
}

static float fuWarp(float r, float W00, float W01, float W10, float W11)
{
  const float A = W10 + W11 - W01 - W00;
  const float B = 2.f * (W01 + W00);
  if (A + B == 0.f)
    return r;
  const float C = -(W00 + W01 + W10 + W11) * r;
  return solveQuadratic(A, B, C);
}

static float fvWarp(float r, float u, float W00, float W01, float W10, float W11)
{
  const float A = (1.f - u) * W01 + u * W11 - (1.f - u) * W00 - u * W10;
  const float B = 2.f * (1.f - u) * W00 + 2.f * u * W10;
  if (A + B == 0.f)
    return r;
  const float C = -((1.f - u) * W01 + u * W11 + (1.f - u) * W00 + u * W10) * r;
  return solveQuadratic(A, B, C);
}

static vec3f cosineWeightedRNG(vec3f q0, vec3f e0, vec3f e1, vec3f P, vec3f Ng, vec2f s) // SphericalQuadSampling.ih:157-196
{
  vec3f wp[4];
  float prob[4];
  wp[0] = q0;
  wp[1] = q0 + e0;
  wp[2] = q0 + e1;
  wp[3] = wp[1] + e1;
  for (int i = 0; i < 4; i++) {
    const vec3f wpn = normalize(wp[i] - P);
    prob[i] = maxf(0.f, dot(wpn, Ng));
  }
  if ((prob[0] + prob[2]) == 0.f)
    return v3(s.x, s.y, 1.f);
  const float totProb = prob[0] + prob[2] + prob[1] + prob[3];
  const float u = fuWarp(s.x, prob[0], prob[2], prob[1], prob[3]);
  const float v = fvWarp(s.y, u, prob[0], prob[2], prob[1], prob[3]);
  const float pdf = 4.f * ((1.f - u) * (1.f - v) * prob[0] + u * (1.f - v) * prob[1] + u * v * prob[3] + (1.f - u) * v * prob[2]) / totProb;
  return v3(u, v, pdf);
}

static LightSample lightSample(const Scene &sc, const LightData &l, const DG &dg, vec2f s)
{
  LightSample res;
  if (l.type == kLightGeometry) { // GeometryLight_sample (render/pathtracer/GeometryLight.ispc:60-123)
    const OrcScene &d = sc.world.desc;
    const OrcModel &model = d.models[l.model];
    const OrcGeometry &g = d.geometries[model.geometry];
    const Sample1D smp = Distribution1D_sample((int)l.primCDF.size(), l.primCDF.data(), 0, s.x);
    const SampleAreaRes as = sampleArea(g, smp.idx, sc.world.xfm[l.inst], sc.world.rcpXfm[l.inst], v2(smp.frac, s.y));
    const vec3f dir = as.pos - dg.P;
    const float dist = length(dir);
    res.dir = dir / dist;
    const float rcosd = rcp(std::fabs(dot(as.normal, res.dir)));
    res.pdf = l.geoPdf * sqr(dist) * rcosd;
    // calcEpsilon(P, t) (common/DifferentialGeometry.ih:48-51)
    const float epsP = maxf(reduce_max(vabs(as.pos)), dist) * kUlpEpsilon;
    res.dist = dist - maxf(as.epsilon, epsP) * rcosd;
    res.weight = sc.materials[model.material].radiance * rcp(res.pdf);
    return res;
  }
  if (l.type == ORC_LIGHT_QUAD) { // QuadLight.ispc:82-136
    if (s.x == 0.0f && s.y == 0.0f) { // SampleArea at the centre
      const vec3f p = l.position + l.edge1 * 0.5f + l.edge2 * 0.5f;
      const vec3f dir = p - dg.P;
      const float dist = length(dir);
      res.dir = dir / dist;
      res.dist = dist;
      const float cosd = dot(l.nnormal, res.dir);
      res.pdf = l.ppdf * sqr(dist) / std::fabs(cosd);
      const float weight = cosd > 0.f ? rcp(res.pdf) : 0.f;
      res.weight = l.radiance * weight;
      return res;
    }
    const SphericalQuad quad = sphericalQuadCreate(l.position, l.edge1, l.edge2, neg(l.nnormal), dg.P);
    const vec3f cosW = cosineWeightedRNG(l.position, l.edge1, l.edge2, dg.P, dg.Ng, s);
    const vec3f dir = sampleSphericalQuad(quad, v2(cosW.x, cosW.y));
    const float dist = length(dir);
    res.dir = dir / dist;
    res.dist = dist;
    const float cosd = dot(l.nnormal, res.dir);
    res.pdf = quad.S == 0.f ? 0.f : cosW.z / quad.S;
    res.weight = (cosd > 0.f) && (res.pdf != 0.f) ? l.radiance * rcp(res.pdf) : v3(0.f);
    return res;
  }
  if (l.type == ORC_LIGHT_DISTANT) { // DirectionalLight.ispc:14-27
    res.dir = l.frame.vz;
    res.dist = kInf;
    res.pdf = l.pdf;
    if (l.cosAngle < kCosAngleMax)
      res.dir = mul(l.frame, uniformSampleCone(l.cosAngle, s));
    res.weight = l.irradiance;
    return res;
  }
  // AmbientLight.ispc:17-38
  if (reduce_max(vabs(dg.Ns)) > 0.f) {
    const vec3f localDir = cosineSampleHemisphere(s);
    res.dir = mul(frame(dg.Ns), localDir);
    res.pdf = localDir.z * kOneOverPi;
  } else {
    res.dir = uniformSampleSphere(1.f, s);
    res.pdf = rcp(kFourPi * sqr(1.f));
  }
  res.dist = kInf;
  res.weight = l.radiance * rcp(res.pdf);
  return res;
}

static LightEval lightEval(const LightData &l, const DG &dg, vec3f dir, float minDist, float maxDist)
{
  LightEval res;
  res.radiance = v3(0.f);
  res.pdf = 0.f;
  if (l.type == ORC_LIGHT_QUAD) { // QuadLight.ispc:167-209
    const float cosd = dot(l.nnormal, dir);
    if (cosd <= 0.f)
      return res;
    const vec3f c = l.position - dg.P;
    const vec3f r = cross(c, dir);
    const float rcosd = rcp(cosd);
    const float u = dot(r, l.edge1) * rcosd;
    const float v = -dot(r, l.edge2) * rcosd;
    if (minf(u, v) < 0.f || maxf(u, v) * l.ppdf > 1.0f)
      return res;
    const float dist = dot(l.nnormal, c) * rcosd;
    if (dist <= minDist || dist > maxDist)
      return res;
    const SphericalQuad quad = sphericalQuadCreate(l.position, l.edge1, l.edge2, neg(l.nnormal), dg.P);
    if (quad.S > 0.f) {
      res.radiance = l.radiance;
      res.pdf = rcp(quad.S);
    }
    return res;
  }
  if (l.type == ORC_LIGHT_DISTANT) { // DirectionalLight.ispc:58-70
    if (kInf <= maxDist && l.cosAngle < kCosAngleMax && dot(l.frame.vz, dir) > l.cosAngle) {
      res.radiance = l.irradiance * l.pdf;
      res.pdf = l.pdf;
    }
    return res;
  }
  // AmbientLight.ispc:40-55
  res.radiance = kInf <= maxDist ? l.radiance : v3(0.f);
  if (reduce_max(vabs(dg.Ns)) > 0.f)
    res.pdf = maxf(dot(dg.Ns, dir), 0.f) * kOneOverPi;
  else
    res.pdf = rcp(kFourPi * sqr(1.f));
  return res;
}

void lightSampleHook(const Scene &s, int light, vec3f P, vec3f Ng, vec3f Ns, vec2f smp, float out[8])
{
  DG dg{};
  dg.P = P;
  dg.Ng = Ng;
  dg.Ns = Ns;
  dg.material = -1;
  const LightSample ls = lightSample(s, s.lights[(size_t)light], dg, smp);
  const float v[8] = {ls.dir.x, ls.dir.y, ls.dir.z, ls.dist, ls.weight.x, ls.weight.y, ls.weight.z, ls.pdf};
  for (int k = 0; k < 8; k++)
    out[k] = v[k];
}

// ---- path tracer -----------------------------------------------------------------

enum VertexType
{
  V_SURFACE,
  V_VOLUME,
  V_CAMERA,
  V_LIGHT,
  V_ENVIRONMENT
};

struct PathVertex
{
  DG dg;
  vec3f wo, albedo;
  bool hasBsdf;
  BSDF bsdf;
  float pdf_w;
  int type;
  uint32_t numLightSamples;
  int volInst; // V_VOLUME: index into Scene::volInsts
};

struct PathState
{
  vec3f throughput, contribution;
  Medium currentMedium;
  bool firstBounceLight, straightPath, specularTransmissionPath;
  uint32_t scatteringEvents, depth;
  LDSampler ld;
  RandomSampler rs; // volume sampling (PathStructs.ih:83)
};

struct ScreenSample
{
  vec3f rgb;
  float alpha, z;
  vec3f position, firstNormal, normal, albedo;
  uint32_t primID, geomID, instID;
};

struct PathCtx
{
  const Scene *scene;
  const RendererData *r;
  const LightData *lights; // [geometry lights (numGeoLights), world lights]
  const float *lightsCDF;
  uint32_t numLights, numGeoLights, numFirst, numIndirect;
  vec3f cameraDir;
  bool projectedDepth;
};

static inline float misHeuristic(float pdf1, float pdf2) // PathTracerUtil.ih:30-46
{
  const float p = sqr(pdf1) * rcp(sqr(pdf1) + sqr(pdf2));
  return pdf1 > 1e17f ? 1.0f : p;
}

// closest hit of a shadow-ray segment (TransparentShadow.ispc:36-84)
static Hit trace(const Scene &s, const Ray &ray) { return s.world.intersect(ray); }
// closest hit of a camera or extension ray, near-ties re-resolved in FP64
static Hit traceExtension(const Scene &s, const Ray &ray) { return s.world.intersect(ray, nullptr, nullptr, true); }

// TransparentShadow.ispc:18-85
// On return shadowRay spans [t0 of the last segment traced, original t]:
// the segment that volumeTransmittance then covers.
static vec3f transparentShadow(const PathCtx &c, vec3f lightContrib, Ray &shadowRay, Medium medium)
{
  uint32_t maxDepth = (uint32_t)c.r->maxDepth;
  const float tOriginal = shadowRay.t;
  while (true) {
    const Hit h = trace(*c.scene, shadowRay);
    if (!h.hit)
      return lightContrib;
    DG dg;
    Ray hr = shadowRay;
    hr.t = h.t;
    postIntersect(*c.scene, hr, h, dg, false);
    const vec3f transparency = getTransparency(*c.scene, dg, hr, medium);
    lightContrib = lightContrib * transparency;
    if (!eq(medium.attenuation, v3(0.f))) {
      const vec3f a = medium.attenuation * (h.t - shadowRay.t0);
      lightContrib = lightContrib * v3(fexp(a.x), fexp(a.y), fexp(a.z));
    }
    if (reduce_max(lightContrib) <= c.r->minContribution || --maxDepth <= 0)
      return v3(0.f);
    selectNextMedium(*c.scene, dg, medium);
    shadowRay.t0 = h.t + dg.epsilon;
    shadowRay.t = tOriginal;
  }
}

// ---- volumes (render/pathtracer/volumes/VolumeSampler.ispc) -------------------------

// volumeSampleFreePath (VolumeSampler.ispc:96-199): delta tracking through
// every volume interval of the ray up to its surface hit (ray.t), the ray
// range of the clipping pass being [r0, r1].  The nearest collision wins;
// once one is found, later intervals are clipped to it.  Returns the
// collision distance (inf: none) and sets ray.t, volInst and the albedo.
static float volumeSampleFreePath(const Scene &s, Ray &ray, bool surfaceHit, float r0, float r1, RandomSampler &rs,
    int &volInst, vec3f &albedo)
{
  float minDist = kInf;
  volInst = -1;
  const float tTrace = ray.t;
  bool clipped = surfaceHit;
  for (size_t j = 0; j < s.volInsts.size(); j++) {
    const Scene::VolInst &vi = s.volInsts[j];
    const VolumeData &v = s.volumes[vi.vol];
    const affine3f &rx = s.world.rcpXfm[vi.inst];
    const vec3f o = xfmPoint(rx, ray.org), d = xfmVector(rx, ray.dir);
    float lo, hi;
    if (!boxInterval(v, o, d, ray.t0, tTrace, lo, hi))
      continue;
    lo = maxf(0.f, lo);
    if (clipped)
      hi = minf(hi, ray.t);
    if (hi < lo)
      continue;
    if (hi < r0 || r1 < lo)
      continue;
    float mu;
    vec3f alb;
    const float dist = deltaTracking(v, o, d, maxf(r0, lo), minf(r1, hi), rs, mu, alb);
    if (dist < kInf && dist < minDist) {
      minDist = dist;
      volInst = (int)j;
      albedo = alb;
      ray.t = dist;
      clipped = true;
    }
  }
  return minDist;
}

// volumeTransmittance (VolumeSampler.ispc:201-277): 1 if delta tracking
// passes every volume interval of [ray.t0, ray.t] (within [r0, r1]), else 0
static float volumeTransmittance(const Scene &s, const Ray &ray, float r0, float r1, RandomSampler &rs)
{
  float T = 1.f;
  for (size_t j = 0; j < s.volInsts.size() && T > 0.f; j++) {
    const Scene::VolInst &vi = s.volInsts[j];
    const VolumeData &v = s.volumes[vi.vol];
    const affine3f &rx = s.world.rcpXfm[vi.inst];
    const vec3f o = xfmPoint(rx, ray.org), d = xfmVector(rx, ray.dir);
    float lo, hi;
    if (!boxInterval(v, o, d, ray.t0, ray.t, lo, hi))
      continue;
    if (hi < r0 || r1 < lo)
      continue;
    const float b = minf(r1, hi);
    float mu;
    vec3f alb;
    const float dist = deltaTracking(v, o, d, maxf(r0, lo), b, rs, mu, alb);
    T = T * (dist < b ? 0.f : 1.f);
  }
  return T;
}

// HenyeyGreenstein_eval / _sample (volumes/HenyeyGreenstein.ih:28-77); the
// eval's pdf is the phase function whatever the anisotropy
static EvalRes hgEval(float g, vec3f wo, vec3f wi)
{
  const float p = hgPhase(dot(neg(wo), wi), g);
  return EvalRes{v3(p), p};
}

static SampleRes hgSample(float g, vec3f wo, vec2f s)
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
  if (std::fabs(g) < 1e-3f)
    cosTheta = 1.f - 2.f * s.x;
  else {
    const float sqrTerm = (1.f - g * g) * rcp(1.f - g + 2.f * g * s.x);
    cosTheta = (1.f + g * g - sqrTerm * sqrTerm) * rcp(2.f * g);
  }
  res.type = S_SMOOTH;
  const float sinTheta = std::sqrt(maxf(0.f, 1.f - cosTheta * cosTheta));
  const float phi = kTwoPi * s.y;
  res.wi = mul(frame(neg(wo)), cartesian(phi, sinTheta, cosTheta));
  res.pdf = hgPhase(cosTheta, g);
  return res;
}

float volumeTransmittanceHook(const Scene &s, const Ray &ray, float r0, float r1, RandomSampler &rs)
{
  return volumeTransmittance(s, ray, r0, r1, rs);
}

float volumeFreePathHook(const Scene &s, Ray &ray, float r0, float r1, RandomSampler &rs, int &volInst, vec3f &albedo)
{
  return volumeSampleFreePath(s, ray, false, r0, r1, rs, volInst, albedo);
}

void bsdfHook(const OrcMaterial &om, vec3f wo, vec3f wi, vec2f s, float ss, float out[13])
{
  Scene sc;
  sc.materials.resize(1);
  setupMaterial(sc.materials[0], om);
  DG dg{};
  dg.Ng = dg.Ns = v3(0.f, 0.f, 1.f);
  dg.dPds = v3(1.f, 0.f, 0.f);
  dg.dPdt = v3(0.f, 1.f, 0.f);
  dg.color = {1.f, 1.f, 1.f, 1.f};
  dg.material = 0;
  BSDF b{};
  if (!getBSDF(sc, dg, sc.materials[0].outside, b)) {
    for (int k = 0; k < 13; k++)
      out[k] = 0.f;
    return;
  }
  const EvalRes e = bsdfEval(b, wo, wi);
  const SampleRes r = bsdfSample(b, wo, s, ss);
  const float v[13] = {e.value.x, e.value.y, e.value.z, e.pdf, r.wi.x, r.wi.y, r.wi.z, r.pdf, r.weight.x,
      r.weight.y, r.weight.z, (float)r.type, (float)b.scatteringType};
  for (int k = 0; k < 13; k++)
    out[k] = v[k];
}

void specularHook(vec3f ks, float ns, vec3f wo, vec3f wi, vec2f s, float out[11])
{
  MaterialData m{};
  m.Ks = ks;
  m.Ns = ns;
  BSDF b{};
  b.frame = {v3(1.f, 0.f, 0.f), v3(0.f, 1.f, 0.f), v3(0.f, 0.f, 1.f)};
  specularLobe(m, 1.f, b);
  const EvalRes e = specularEval(b, wo, wi);
  const SampleRes r = specularSample(b, wo, s);
  const float v[11] = {e.value.x, e.value.y, e.value.z, e.pdf, r.wi.x, r.wi.y, r.wi.z, r.pdf, r.weight.x, r.weight.y,
      r.weight.z};
  for (int k = 0; k < 11; k++)
    out[k] = v[k];
}

void hgSampleHook(float g, vec3f wo, vec2f s, vec3f &wi, float &pdf, uint32_t &type)
{
  const SampleRes r = hgSample(g, wo, s);
  wi = r.wi;
  pdf = r.pdf;
  type = r.type;
}

float hgEvalHook(float g, vec3f wo, vec3f wi) { return hgEval(g, wo, wi).pdf; }

// NextEventEstimation.ispc:25-160
static vec3f nextEventEstimation(const PathCtx &c, PathState &ps, PathVertex &pv)
{
  vec3f L = v3(0.f);
  const uint32_t n = ps.firstBounceLight ? c.numFirst : c.numIndirect;
  pv.numLightSamples = n;
  for (uint32_t i = 0; i < n; i++) {
    float ss;
    const vec2f s = LDSampler_get3LightSamples(ps.ld, i, n > 1, ss);
    const Sample1D sel = Distribution1D_sample((int)c.numLights, c.lightsCDF, 0, ss);
    const float selProb = ORC_PLANT == 4 ? sel.prob : sel.prob * (float)n; // plant 4: sample count dropped
    LightSample ls = lightSample(*c.scene, c.lights[sel.idx], pv.dg, s);
    ls.weight = ls.weight / selProb;
    ls.pdf *= selProb;
    if (reduce_max(ls.weight) <= 0.0f || ls.pdf <= 0.0f)
      continue;
    EvalRes fe{v3(0.f), 0.f};
    if (pv.type == V_SURFACE && pv.hasBsdf)
      fe = bsdfEval(pv.bsdf, pv.wo, ls.dir);
    else if (pv.type == V_VOLUME)
      fe = hgEval(c.scene->volumes[c.scene->volInsts[pv.volInst].vol].anisotropy, pv.wo, ls.dir);
    if (reduce_max(fe.value) <= 0.0f)
      continue;
    vec3f org = pv.dg.P;
    if (pv.type != V_VOLUME && dot(pv.dg.Ng, ls.dir) < 0.f)
      org = org - (2.0f * pv.dg.epsilon) * pv.dg.Ng;
    Ray sr{org, ls.dir, 0.f, ls.dist};
    const vec3f throughput = ps.throughput * fe.value;
    const vec3f unshaded = throughput * ls.weight;
    const vec3f lc = transparentShadow(c, unshaded, sr, ps.currentMedium);
    if (reduce_max(lc) > 0) {
      if (c.scene->hasVolumes) {
        // two independent transmittance estimates: one weights the sample,
        // the other enters the MIS weight (NextEventEstimation.ispc:126-149)
        const float T = volumeTransmittance(*c.scene, sr, 0.f, ls.dist, ps.rs);
        if (T > 0.f) {
          const float Tmis = volumeTransmittance(*c.scene, sr, 0.f, ls.dist, ps.rs);
          L = L + (T * lc) * misHeuristic(ls.pdf, fe.pdf * Tmis);
        }
      } else
        L = L + lc * misHeuristic(ls.pdf, fe.pdf);
    }
  }
  return L;
}

// VirtualLight.ispc:38-107
static vec3f evaluateVirtualLights(const PathCtx &c, PathState &ps, const PathVertex &last, const PathVertex &pv, const Ray &ray)
{
  vec3f L = v3(0.f);
  if (c.numLights <= c.numGeoLights)
    return L;
  const float lower = distance(last.dg.P, ray.org);
  const float upper = pv.type == V_ENVIRONMENT ? kInf : distance(last.dg.P, ray.org + ray.t * ray.dir);
  const float selectionPDF = (float)last.numLightSamples / (float)c.numLights;
  for (uint32_t i = c.numGeoLights; i < c.numLights; i++) {
    const LightData &l = c.lights[i];
    if (!ps.straightPath || l.visible) {
      const LightEval le = lightEval(l, last.dg, ray.dir, lower, upper);
      if (reduce_max(le.radiance) > 0.0f) {
        float mis = 1.0f;
        if (selectionPDF != 0.f) {
          // transmittance from the last vertex over the light's distance
          // range (VirtualLight.ispc:76-99)
          float T = 1.f;
          if (c.scene->hasVolumes) {
            const Ray sr{last.dg.P, ray.dir, lower, upper};
            T = volumeTransmittance(*c.scene, sr, lower, upper, ps.rs);
          }
          mis = misHeuristic(last.pdf_w * T, selectionPDF * le.pdf);
        }
        L = L + ps.throughput * le.radiance * mis;
      }
    }
  }
  return L;
}

// evaluateGeometryLights (render/pathtracer/GeometryLight.ispc:22-56): the
// emission of a hit luminous surface, MIS-weighted against NEE
static vec3f evaluateGeometryLights(const PathCtx &c, const PathState &ps, const PathVertex &last, const PathVertex &pv, const Ray &ray)
{
  if (pv.dg.material < 0)
    return v3(0.f);
  const MaterialData &m = c.scene->materials[pv.dg.material];
  if (m.type != ORC_MAT_LUMINOUS || !(reduce_max(m.radiance) > 0.f))
    return v3(0.f);
  float misWeight = 1.0f;
  if (c.numGeoLights > 0 && last.numLightSamples > 0) {
    const float lightSelectionPDF = (float)last.numLightSamples / (float)c.numLights;
    const float cosd = dot(pv.dg.Ng, ray.dir);
    const vec3f vdis = last.dg.P - pv.dg.P;
    const float lePdf = pv.dg.areaPDF * dot(vdis, vdis) / std::fabs(cosd);
    misWeight = misHeuristic(last.pdf_w, lightSelectionPDF * lePdf);
  }
  return ps.throughput * m.radiance * misWeight;
}

// postIntersect of a volume vertex (PathSampler.ispc:51-68,
// VolumetricModel.ih:16-33)
static void volumePostIntersect(const Scene &s, const Ray &ray, PathVertex &pv)
{
  const Scene::VolInst &vi = s.volInsts[pv.volInst];
  DG &dg = pv.dg;
  dg.P = ray.org + ray.t * ray.dir;
  const uint32_t instUser = s.world.desc.instances[vi.inst].userID;
  dg.instID = instUser != kInvalidID ? instUser : (uint32_t)vi.inst;
  dg.Ng = dg.Ns = dg.dPds = dg.dPdt = v3(0.f);
  dg.color = vec4f{0.f, 0.f, 0.f, 0.f};
  dg.primID = 0;
  const uint32_t volUser = s.volumes[vi.vol].userID;
  dg.objID = volUser != kInvalidID ? volUser : (uint32_t)vi.local;
  dg.areaPDF = 0.f;
  dg.epsilon = 0.f;
  dg.material = -1;
}

// PathSampler.ispc:109-415
static void samplePath(const PathCtx &c, PathState &ps, Ray ray, ScreenSample &sample)
{
  const RendererData &r = *c.r;
  PathVertex last;
  last.type = V_CAMERA;
  last.pdf_w = kInf;
  last.dg.P = ray.org;
  last.dg.epsilon = maxf(reduce_max(vabs(ray.org)), 0.f) * kUlpEpsilon;
  last.dg.Ns = ray.dir;
  last.dg.Ng = ray.dir;
  last.numLightSamples = 0;
  sample.alpha = 1.0f;
  do {
    PathVertex pv;
    pv.hasBsdf = false;
    pv.pdf_w = kInf;
    pv.type = V_ENVIRONMENT;
    pv.numLightSamples = 0;
    pv.volInst = -1;
    const float r0 = ray.t0, r1 = ray.t; // the clipping pass's ray interval
    const Hit hit = traceExtension(*c.scene, ray);
    if (hit.hit) {
      pv.type = V_SURFACE;
      ray.t = hit.t;
    }
    pv.wo = neg(ray.dir);
    if (c.scene->hasVolumes) {
      const float freePath = volumeSampleFreePath(*c.scene, ray, hit.hit, r0, r1, ps.rs, pv.volInst, pv.albedo);
      if (freePath < kInf) {
        pv.type = V_VOLUME;
        ps.throughput = ps.throughput * pv.albedo;
      }
    }
    if (ps.depth == 0)
      sample.z = c.projectedDepth ? ray.t * dot(ray.dir, c.cameraDir) : ray.t;
    if (pv.type == V_ENVIRONMENT && ps.straightPath) {
      const vec4f bg = r.bgColor;
      sample.alpha -= luminance(ps.throughput) * (1.0f - bg.w);
      const vec3f bg3 = ps.throughput * v3(bg.x, bg.y, bg.z);
      ps.contribution = ps.contribution + bg3;
    }
    if (pv.type != V_VOLUME)
      ps.contribution = ps.contribution + evaluateVirtualLights(c, ps, last, pv, ray);
    if (pv.type == V_ENVIRONMENT)
      break;
    // with geometry lights the emission at maxDepth is still evaluated
    // (PathSampler.ispc:249-255)
    if (ps.depth >= (uint32_t)r.maxDepth && c.numGeoLights == 0)
      break;
    if (pv.type == V_VOLUME)
      volumePostIntersect(*c.scene, ray, pv);
    else {
      postIntersect(*c.scene, ray, hit, pv.dg, true);
      ps.contribution = ps.contribution + evaluateGeometryLights(c, ps, last, pv, ray);
    }
    if (ps.depth == 0) {
      sample.position = pv.dg.P;
      sample.firstNormal = pv.dg.Ns;
      sample.primID = pv.dg.primID;
      sample.geomID = pv.dg.objID;
      sample.instID = pv.dg.instID;
    }
    if (ps.depth >= (uint32_t)r.maxDepth || ps.scatteringEvents >= (uint32_t)r.maxScatteringEvents)
      break;
    float g = 0.f; // anisotropy of a volume vertex
    if (pv.type == V_SURFACE) {
      pv.hasBsdf = getBSDF(*c.scene, pv.dg, ps.currentMedium, pv.bsdf);
      if (!pv.hasBsdf)
        break;
      pv.albedo = pv.bsdf.albedo;
    } else
      g = c.scene->volumes[c.scene->volInsts[pv.volInst].vol].anisotropy;
    LDSampler_nextGroup(ps.ld);
    // isSmooth (PathTracerUtil.ih:15-28)
    const bool smooth = pv.type == V_SURFACE ? (pv.bsdf.scatteringType & S_SMOOTH) != 0 : hgIsSmooth(g);
    if (smooth) {
      if (ps.scatteringEvents == 0) { // updateAuxilliaryData (PathTracerUtil.ih:48-60)
        sample.albedo = ps.throughput * pv.albedo;
        sample.normal = ps.throughput * (pv.type == V_SURFACE ? pv.bsdf.frame.vz : pv.dg.Ns);
      }
      const vec3f shaded = nextEventEstimation(c, ps, pv);
      ps.contribution = ps.contribution + shaded;
      ps.firstBounceLight = false;
    }
    float ss;
    uint32_t uu3;
    const vec2f s = LDSampler_getNext4Samples(ps.ld, ss, uu3);
    // sampleDirection (PathSampler.ispc:72-107): the surface's same-side
    // check keeps light from leaking
    SampleRes fs;
    if (pv.type == V_SURFACE) {
      fs = bsdfSample(pv.bsdf, pv.wo, s, ss);
      const bool sameSide = dot(fs.wi, pv.dg.Ng) >= 0.f;
      const bool reflection = (fs.type & S_REFLECTION) != 0;
      if (sameSide != reflection)
        fs.weight = v3(0.0f);
    } else
      fs = hgSample(g, pv.wo, s);
    pv.pdf_w = fs.pdf;
    if (reduce_max(fs.weight) <= 0.0f || fs.pdf <= 0.0f)
      break;
    ps.throughput = ps.throughput * fs.weight;
    if (ps.depth >= (uint32_t)r.rouletteDepth) {
      const float rr = LDSampler_finalizeDim3(ps.ld, uu3);
      const float contProb = minf(reduce_max(ps.throughput), kMaxRouletteContProb);
      if (rr > contProb)
        break;
      ps.throughput = ps.throughput * rcp(contProb);
    }
    if (reduce_min(ps.currentMedium.attenuation) < 0.f) {
      const vec3f a = ps.currentMedium.attenuation * ray.t;
      ps.throughput = ps.throughput * v3(fexp(a.x), fexp(a.y), fexp(a.z));
    }
    vec3f rayOrg = pv.dg.P;
    if (pv.type == V_SURFACE && (fs.type & S_TRANSMISSION)) {
      rayOrg = rayOrg - (2.0f * pv.dg.epsilon) * pv.dg.Ng;
      selectNextMedium(*c.scene, pv.dg, ps.currentMedium);
    }
    if ((fs.type & ~SPEC_TRANS) || pv.type == V_VOLUME)
      last = pv;
    if (!eq(ray.dir, fs.wi))
      ps.straightPath = false;
    if (!(fs.type & SPEC_TRANS))
      ps.specularTransmissionPath = false;
    ray.org = rayOrg;
    ray.dir = fs.wi;
    ray.t0 = 0.f;
    ray.t = kInf;
    ps.depth++;
    if (fs.type & S_SMOOTH)
      ps.scatteringEvents++;
  } while (reduce_max(ps.throughput) > r.minContribution);
  sample.rgb = ps.contribution;
  if (std::isnan(ps.contribution.x) || std::isnan(ps.contribution.y) || std::isnan(ps.contribution.z)) {
    sample.rgb = v3(0.f);
    sample.alpha = 1.0f;
  }
}

static inline uint32_t clz32(uint32_t x) { return x ? (uint32_t)__builtin_clz(x) : 32u; }

static ScreenSample renderPixel(const PathCtx &c, const FrameBuffer &fb, const CameraData &cam, uint32_t ix, uint32_t iy) // PathTracer.ispc:75-216
{
  const RendererData &r = *c.r;
  ScreenSample acc;
  acc.rgb = v3(0.f);
  acc.alpha = 0.f;
  acc.z = kInf;
  acc.position = acc.firstNormal = acc.normal = acc.albedo = v3(0.f);
  acc.primID = acc.geomID = acc.instID = kInvalidID;
  const uint32_t accumID = (uint32_t)fb.frameID;
  const uint32_t spp = (uint32_t)r.spp;
  uint32_t seed = 0x49c5f37eu;
  uint32_t sampleOffs, sampleOffsLight, sampleBits, sampleBitsLight;
  if (fb.targetFrames) {
    const uint32_t targetSpp = (uint32_t)fb.targetFrames * spp;
    sampleOffs = OwenScramble4(interleaveZOrder16(ix, iy), seed) * targetSpp;
    sampleOffsLight = sampleOffs * c.numFirst;
    sampleBits = 16 + 32 - clz32(targetSpp - 1);
    sampleBitsLight = 48 - clz32(targetSpp * c.numFirst - 1);
  } else {
    seed ^= (uint32_t)fb.w * iy + ix;
    sampleOffs = 0;
    sampleOffsLight = 0;
    sampleBits = 16;
    sampleBitsLight = 16;
  }
  const float rcpW = 1.f / (float)fb.w, rcpH = 1.f / (float)fb.h;
  for (uint32_t s = 0; s < spp; s++) {
    const uint32_t sampleID = accumID * spp + s;
    PathState ps;
    LDSampler_init(ps.ld, seed, sampleOffs + sampleID, sampleOffsLight + sampleID * c.numFirst, sampleBits, sampleBitsLight);
    RandomSampler_init(ps.rs, seed, sampleOffs + sampleID); // PathTracer.ispc:157
    const vec2f center = v2(((float)ix + 0.5f) * rcpW, ((float)iy + 0.5f) * rcpH);
    vec3ui ss234;
    const vec2f pixelSample = LDSampler_getNext5Samples(ps.ld, ss234);
    const vec2f pf = pixelFilterSample(r, pixelSample);
    const vec2f screen = v2(center.x + pf.x * rcpW, center.y + pf.y * rcpH);
    Ray ray;
    cameraRay(cam, screen, ray);
    ScreenSample smp;
    smp.rgb = v3(0.f);
    smp.alpha = 0.f;
    smp.z = kInf;
    smp.position = smp.firstNormal = smp.normal = smp.albedo = v3(0.f);
    smp.primID = smp.geomID = smp.instID = kInvalidID;
    ps.straightPath = true;
    ps.specularTransmissionPath = true;
    ps.depth = 0;
    ps.scatteringEvents = 0;
    ps.firstBounceLight = true;
    ps.throughput = v3(1.f);
    ps.contribution = v3(0.f);
    ps.currentMedium = {v3(0.f), 1.0f};
    samplePath(c, ps, ray, smp);
    smp.rgb = vmin(smp.rgb, v3(r.maxRadiance));
    // ScreenSample_accumulate (render/ScreenSample.ih:59-77)
    acc.rgb = acc.rgb + smp.rgb;
    acc.alpha = acc.alpha + smp.alpha;
    acc.position = acc.position + smp.position;
    acc.firstNormal = acc.firstNormal + smp.firstNormal;
    acc.normal = acc.normal + smp.normal;
    acc.albedo = acc.albedo + smp.albedo;
    if (std::fabs(smp.z) < std::fabs(acc.z)) {
      acc.z = smp.z;
      acc.primID = smp.primID;
      acc.geomID = smp.geomID;
      acc.instID = smp.instID;
    }
  }
  const float rspp = 1.0f / (float)spp;
  acc.rgb = acc.rgb * rspp;
  acc.alpha *= rspp;
  acc.position = acc.position * rspp;
  acc.firstNormal = acc.firstNormal * rspp;
  acc.normal = acc.normal * rspp;
  acc.albedo = acc.albedo * rspp;
  return acc;
}

template <typename T>
static inline T accum(const T &old, const T &nw, float f)
{
  return f > 0.8f ? nw : old * (1.f - f) + nw * f;
}
static inline vec4f accum4(const vec4f &a, const vec4f &b, float f)
{
  if (f > 0.8f)
    return b;
  const float g = 1.f - f;
  return {g * a.x + f * b.x, g * a.y + f * b.y, g * a.z + f * b.z, g * a.w + f * b.w};
}
static inline vec3f accum3(const vec3f &a, const vec3f &b, float f)
{
  if (f > 0.8f)
    return b;
  return (1.f - f) * a + f * b;
}

// LocalFB_accumulateSample (fb/LocalFB.ispc:14-78)
static void accumulateSample(FrameBuffer &fb, uint32_t ix, uint32_t iy, const ScreenSample &s)
{
  const size_t pid = (size_t)ix + (size_t)fb.w * iy;
  const int accumID = fb.frameID;
  const float accScale = 1.0f / (float)(accumID + 1);
  const vec4f cs = {s.rgb.x, s.rgb.y, s.rgb.z, s.alpha};
  if (!fb.color.empty())
    fb.color[pid] = accum4(fb.color[pid], cs, accScale);
  if (fb.accumulateVariance) {
    const float accVarScale = 1.0f / (float)(accumID / 2 + 1);
    fb.variance[pid] = accum4(fb.variance[pid], cs, accVarScale);
  }
  bool updateIDs = accumID == 0;
  if (!fb.depth.empty()) {
    if (accumID > 0 && std::fabs(s.z) < std::fabs(fb.depth[pid]))
      updateIDs = true;
    if (updateIDs)
      fb.depth[pid] = s.z;
  }
  if (updateIDs) {
    if (!fb.primID.empty())
      fb.primID[pid] = s.primID;
    if (!fb.objID.empty())
      fb.objID[pid] = s.geomID;
    if (!fb.instID.empty())
      fb.instID[pid] = s.instID;
  }
  if (!fb.normal.empty())
    fb.normal[pid] = accum3(fb.normal[pid], s.normal, accScale);
  if (!fb.firstNormal.empty())
    fb.firstNormal[pid] = accum3(fb.firstNormal[pid], s.firstNormal, accScale);
  if (!fb.albedo.empty())
    fb.albedo[pid] = accum3(fb.albedo[pid], s.albedo, accScale);
  if (!fb.position.empty())
    fb.position[pid] = accum3(fb.position[pid], s.position, accScale);
}

// Variance_kernel (fb/frame_ops/Variance.ispc:11-59)
static void varianceKernel(FrameBuffer &fb)
{
  const int tx = (fb.w + 7) / 8, ty = (fb.h + 7) / 8;
  for (int ty_ = 0; ty_ < ty; ty_++)
    for (int tx_ = 0; tx_ < tx; tx_++) {
      float v = 0.f;
      int cnt = 0;
      for (int y = 0; y < 8; y++) {
        const int fy = ty_ * 8 + y;
        if (fy >= fb.h)
          continue;
        for (int x = 0; x < 8; x++) {
          const int fx = tx_ * 8 + x;
          if (fx >= fb.w)
            continue;
          const size_t id = (size_t)fy * fb.w + fx;
          const vec4f c = fb.color[id];
          const vec4f va = fb.variance[id];
          const float den2 = (c.x + c.y + c.z) + (1.f - c.w);
          if (den2 > 0.f) {
            cnt++;
            const float diff = std::fabs(c.x - va.x) + std::fabs(c.y - va.y) + std::fabs(c.z - va.z) + std::fabs(c.w - va.w);
            v += ORC_PLANT == 3 ? diff : diff * rsqrt(den2); // plant 3: no normalisation
          }
        }
      }
      if (cnt)
        v *= rcp((float)cnt);
      fb.taskVariance[(size_t)ty_ * tx + tx_] = v * 64.f;
    }
  fb.varianceValid = true;
}

void varianceHook(FrameBuffer &fb) { varianceKernel(fb); }

// LiveVarianceFrameOp::getAvgError (fb/frame_ops/Variance.cpp:32-46), with the
// sum taken in FP64
float getVariance(const FrameBuffer &fb, const RendererData &r)
{
  if (fb.variance.empty() || !fb.varianceValid)
    return kInf;
  float maxErr = 0.f;
  double sum = 0.0;
  int active = 0;
  for (float e : fb.taskVariance) {
    maxErr = maxf(maxErr, e);
    if (e >= r.errorThreshold) {
      sum += (double)e;
      active++;
    }
  }
  return active ? (float)(sum / (double)active) : maxErr;
}

static inline uint32_t cvtU8(float f) { return (uint32_t)(255.f * clampf(f) + 0.5f); }
static inline float linearToSrgb(float x)
{
  x = clampf(x);
  return x <= 0.0031308f ? 12.92f * x : 1.055f * fexp(flog(x) * (1.f / 2.4f)) - 0.055f;
}

// ColorConversion_kernel (fb/frame_ops/ColorConversion.ispc:11-36)
static void colorConversion(FrameBuffer &fb)
{
  const size_t n = (size_t)fb.w * fb.h;
  if (fb.color.empty() || fb.format == 0)
    return;
  if (fb.format == 3) {
    fb.mapped.resize(n * 16);
    float *o = (float *)fb.mapped.data();
    for (size_t i = 0; i < n; i++) {
      const vec4f v = fb.color[i];
      o[4 * i] = maxf(v.x, 0.f);
      o[4 * i + 1] = maxf(v.y, 0.f);
      o[4 * i + 2] = maxf(v.z, 0.f);
      o[4 * i + 3] = minf(maxf(v.w, 0.f), 1.f);
    }
    return;
  }
  fb.mapped.resize(n * 4);
  uint32_t *o = (uint32_t *)fb.mapped.data();
  for (size_t i = 0; i < n; i++) {
    const vec4f v = fb.color[i];
    if (fb.format == 1)
      o[i] = cvtU8(v.x) | (cvtU8(v.y) << 8) | (cvtU8(v.z) << 16) | (cvtU8(v.w) << 24);
    else
      o[i] = cvtU8(linearToSrgb(v.x)) | (cvtU8(linearToSrgb(v.y)) << 8) | (cvtU8(linearToSrgb(v.z)) << 16)
          | (cvtU8(v.w) << 24);
  }
}

void renderFrame(const Scene &s, FrameBuffer &fb, const CameraData &cam, const RendererData &r, int nthreads, int y0, int y1)
{
  // FrameBuffer::beginFrame (fb/FrameBuffer.cpp:115-137)
  const bool doAccum = (fb.channels & 4u) != 0;
  fb.frameID = doAccum ? fb.frameID + 1 : 0;
  if (!fb.variance.empty()) {
    const bool evenFrame = (fb.frameID & 1) == 0;
    if (evenFrame)
      fb.pickOdd = (fb.mtGen() & 1u) != 0;
    fb.accumulateVariance = evenFrame != fb.pickOdd;
  }
  if (fb.targetFrames && fb.frameID >= fb.targetFrames)
    return; // accumulationFinished (fb/FrameBuffer.h:188-191)

  PathCtx c;
  c.scene = &s;
  c.r = &r;
  // PathTracerData without geometry lights when the renderer disables them
  // (PathTracer.cpp:49, PathTracerData.cpp:37-43)
  c.numGeoLights = r.geometryLights ? s.numGeoLights : 0u;
  c.lights = s.lights.data() + (s.numGeoLights - c.numGeoLights);
  c.numLights = (uint32_t)s.lights.size() - (s.numGeoLights - c.numGeoLights);
  c.lightsCDF = r.geometryLights ? s.lightsCDF.data() : s.worldLightsCDF.data();
  c.numFirst = r.numLightSamples >= 0 && c.numLights > 0 ? (uint32_t)r.numLightSamples
                                                       : std::min(kMaxLightSamples, c.numLights);
  c.numIndirect = r.limitIndirectLightSamples && c.numFirst >= 1 ? 1u : c.numFirst;
  c.cameraDir = cam.direction;
  c.projectedDepth = fb.projectedDepth;

  // LocalFrameBuffer::getRenderTaskIDs (fb/LocalFB.cpp:238-261): once
  // frameID >= minimumAdaptiveFrames (FrameBuffer.h:183-186) and a variance
  // buffer exists, only the 8x8 tasks whose last error exceeds the threshold
  // are rendered; the others keep their buffers
  const int tasksX = (fb.w + 7) / 8;
  const bool adaptive = r.errorThreshold > 0.f && !fb.variance.empty()
      && fb.frameID >= std::max(2, 16 / std::max(1, r.spp));
  auto taskActive = [&](uint32_t x, uint32_t y) {
    return !adaptive || fb.taskVariance[(size_t)(y / 8) * tasksX + x / 8] > r.errorThreshold;
  };

  std::atomic<int> nextRow{y0};
  auto worker = [&]() {
    for (;;) {
      const int y = nextRow.fetch_add(1);
      if (y >= y1)
        break;
      for (int x = 0; x < fb.w; x++) {
        if (!taskActive((uint32_t)x, (uint32_t)y))
          continue;
        const ScreenSample smp = renderPixel(c, fb, cam, (uint32_t)x, (uint32_t)y);
        accumulateSample(fb, (uint32_t)x, (uint32_t)y, smp);
      }
    }
  };
  if (nthreads <= 1) {
    worker();
  } else {
    std::vector<std::thread> th;
    for (int i = 0; i < nthreads; i++)
      th.emplace_back(worker);
    for (auto &t : th)
      t.join();
  }
  // LocalFrameBuffer::postProcess (fb/LocalFB.cpp:459-478)
  if (!fb.variance.empty() && (fb.frameID & 1) == 1)
    varianceKernel(fb);
  colorConversion(fb);
}

} // namespace orc
