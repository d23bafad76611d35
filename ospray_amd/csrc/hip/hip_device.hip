// The "hip" OSPRay device: object types, scene assembly, framebuffer state
// and frame orchestration on one HIP stream.
//
// Mirrors the reference CPU device (modules/cpu/ISPCDevice.cpp:154-658) and
// its object commits (Mesh.cpp:37-204, Spheres.cpp:46-106,
// GeometricModel.cpp:21-87, Group.cpp:69-209, Instance.cpp:20-68,
// World.cpp:70-188, OBJ/Metal/Glass.cpp, QuadLight/DirectionalLight/
// AmbientLight.cpp, PerspectiveCamera.cpp:30-110, Renderer.cpp:38-178,
// PathTracer.cpp:29-107, FrameBuffer.cpp:31-193, LocalFB.cpp:49-528).
// Rendering is the wavefront pipeline of pt_kernels.hip, launched through
// the C-ABI layer of include/ospray_hip/hip_pt.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <cmath>
#include <map>
#include <set>
#include <memory>
#include <random>
#include <thread>
#include <vector>

#include "../core/core.h"
#include "bvh_build.h"
#include "ospray/ospray_hip.h"
#include "ospray_hip/hip_bvh.h"
#include "ospray_hip/hip_nrc.h"
#include "ospray_hip/hip_pt.h"
#include "pt_sampler.h"
#include "pt_shade.h"
#include "pt_volume.h"
#include "pt_types.h"

using namespace ospcore;

namespace hipdev {

using hpt::vec3f;
using hpt::v3;

#define HIP_CHECK(expr)                                                                                            \
  do {                                                                                                             \
    const hipError_t e_ = (expr);                                                                                  \
    if (e_ != hipSuccess)                                                                                          \
      throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " #expr);                 \
  } while (0)

struct DevBuf
{
  void *ptr = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void release()
  {
    if (ptr)
      (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
  void ensure(size_t n)
  {
    if (n <= bytes && ptr)
      return;
    release();
    if (n == 0)
      return;
    HIP_CHECK(hipMalloc(&ptr, n));
    bytes = n;
  }
  template <typename T>
  void upload(const std::vector<T> &v)
  {
    ensure(std::max<size_t>(sizeof(T) * v.size(), 16));
    if (!v.empty())
      HIP_CHECK(hipMemcpy(ptr, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
  }
  template <typename T>
  T *as() const
  {
    return (T *)ptr;
  }
};

class HipDevice;

// The path state of one frame in flight: wavefront buffers, queues and
// counters, the traversal scratch rows, and the stream the frame runs on.
// Frames on different lanes run concurrently; lane 0 runs on the device
// stream.
struct PathLane
{
  hipStream_t stream = nullptr;
  bool ownsStream = false;
  DevBuf pathBuf, recBuf, queueBuf, transBuf, counters, overflow;
  hpt::DPaths paths{};
  size_t pathSlots = 0, recSlots = 0;
  bool auxAllocated = false;
  float *auxBase = nullptr;
  // the shadow-ray stages of a bounce run on `side` beside the next
  // bounce's extension on `stream`, with traversal scratch rows of their
  // own; evShade[b & 1] orders bounce b's shading before its shadow stages,
  // evRes[b & 1] its NEE resolve before the shading of bounce b + 1
  hipStream_t side = nullptr;
  DevBuf sideOverflow;
  hipEvent_t evShade[2] = {nullptr, nullptr}, evRes[2] = {nullptr, nullptr};
  ~PathLane()
  {
    if (side)
      (void)hipStreamSynchronize(side);
    for (int k = 0; k < 2; k++) {
      if (evShade[k])
        (void)hipEventDestroy(evShade[k]);
      if (evRes[k])
        (void)hipEventDestroy(evRes[k]);
    }
    if (side)
      (void)hipStreamDestroy(side);
    if (ownsStream && stream)
      (void)hipStreamDestroy(stream);
  }
};

// ---- data helpers -------------------------------------------------------------------

static Data *getData(ManagedObject *o, const char *name)
{
  return (Data *)o->getObject(name, OSP_DATA);
}

static std::vector<float> dataFloats(Data *d, int comps, const char *what);

// Resamples a metal's (wavelength nm, eta, k) table, ordered by wavelength,
// to the 8 samples at 430, 465, ..., 675 nm by linear interpolation
// (Metal.cpp:34-57).  Before the first and after the last entry the
// nearest entry's values hold.
static void resampleIor(const std::vector<float> &t, float etaS[8], float kS[8])
{
  const size_t n = t.size() / 3;
  if (n == 0)
    throw std::runtime_error("metal material: 'ior' is empty");
  size_t p = 0;
  const float *prev = &t[0];
  float wl = 430.f;
  for (int l = 0; l < 8; wl += 35.f, l++) {
    for (; p < n && t[3 * p] < wl; p++)
      prev = &t[3 * p];
    const float *cur = p < n ? &t[3 * p] : prev;
    if (cur[0] == prev[0]) {
      etaS[l] = prev[1];
      kS[l] = prev[2];
    } else {
      const float f = (wl - prev[0]) / (cur[0] - prev[0]);
      etaS[l] = (1.f - f) * prev[1] + f * cur[1];
      kS[l] = (1.f - f) * prev[2] + f * cur[2];
    }
  }
}

static std::vector<float> dataFloats(Data *d, int comps, const char *what)
{
  const size_t es = sizeOfDataType(d->elemType);
  if (es != sizeof(float) * comps || (d->elemType != OSP_VEC3F && d->elemType != OSP_FLOAT && d->elemType != OSP_VEC4F))
    throw std::runtime_error(std::string("unsupported element type for ") + what);
  std::vector<float> out(d->size() * comps);
  for (uint64_t i = 0; i < d->size(); i++)
    std::memcpy(&out[i * comps], d->item(i), es);
  return out;
}

template <typename T>
static std::vector<T *> dataObjects(Data *d, OSPDataType type)
{
  std::vector<T *> out;
  if (!d)
    return out;
  if (!isObjectType(d->elemType))
    throw std::runtime_error("expected an array of objects");
  for (uint64_t i = 0; i < d->size(); i++) {
    ManagedObject *o = d->get<ManagedObject *>(i);
    if (o && o->managedType != type)
      throw std::runtime_error("array holds objects of the wrong type");
    out.push_back((T *)o);
  }
  return out;
}

// holds a reference for as long as it lives
template <typename T>
struct Ref
{
  T *p = nullptr;
  Ref() = default;
  Ref(const Ref &) = delete;
  Ref &operator=(const Ref &) = delete;
  ~Ref() { reset(nullptr); }
  void reset(T *n)
  {
    if (n)
      n->refInc();
    if (p)
      p->refDec();
    p = n;
  }
  T *operator->() const { return p; }
  explicit operator bool() const { return p != nullptr; }
};

// ---- scene objects ---------------------------------------------------------------------

struct Geometry : public ManagedObject
{
  int geomType = -1; // GeomType
  std::vector<float> vertex, normal, spheres;
  std::vector<uint32_t> index;
  size_t numPrims = 0;
  DevBuf dVertex, dIndex, dNormal, dSpheres;
  uint64_t version = 0;

  Geometry(const std::string &t)
  {
    managedType = OSP_GEOMETRY;
    subtype = t;
  }
  std::string toString() const override { return "ospray::Geometry(" + subtype + ")"; }

  void commit() override
  {
    vertex.clear();
    normal.clear();
    spheres.clear();
    index.clear();
    if (subtype == "mesh") { // Mesh.cpp:37-204
      Data *pos = getData(this, "vertex.position");
      Data *idx = getData(this, "index");
      if (!pos)
        throw std::runtime_error("mesh geometry requires 'vertex.position'");
      vertex = dataFloats(pos, 3, "vertex.position");
      if (Data *nrm = getData(this, "vertex.normal")) {
        normal = dataFloats(nrm, 3, "vertex.normal");
        if (normal.size() != vertex.size())
          throw std::runtime_error("'vertex.normal' must have one normal per vertex");
      }
      if (!idx)
        throw std::runtime_error("mesh geometry requires 'index'");
      int comps;
      if (idx->elemType == OSP_VEC3UI || idx->elemType == OSP_VEC3I)
        comps = 3;
      else if (idx->elemType == OSP_VEC4UI || idx->elemType == OSP_VEC4I)
        comps = 4;
      else
        throw std::runtime_error("mesh 'index' must be vec3ui (triangles) or vec4ui (quads)");
      geomType = comps == 3 ? hpt::GEOM_TRIANGLES : hpt::GEOM_QUADS;
      numPrims = idx->size();
      index.resize(numPrims * comps);
      const size_t nv = vertex.size() / 3;
      for (uint64_t i = 0; i < numPrims; i++) {
        std::memcpy(&index[i * comps], idx->item(i), 4 * comps);
        for (int k = 0; k < comps; k++)
          if (index[i * comps + k] >= nv)
            throw std::runtime_error("mesh index out of range of 'vertex.position'");
      }
      dVertex.upload(vertex);
      dIndex.upload(index);
      if (!normal.empty())
        dNormal.upload(normal);
      else
        dNormal.release();
    } else if (subtype == "sphere") { // Spheres.cpp:46-106
      Data *pos = getData(this, "sphere.position");
      if (!pos)
        throw std::runtime_error("sphere geometry requires 'sphere.position'");
      const std::vector<float> p = dataFloats(pos, 3, "sphere.position");
      const float radius = getFloat("radius", 0.01f);
      std::vector<float> r;
      if (Data *rd = getData(this, "sphere.radius"))
        r = dataFloats(rd, 1, "sphere.radius");
      geomType = hpt::GEOM_SPHERES;
      numPrims = p.size() / 3;
      if (!r.empty() && r.size() != numPrims)
        throw std::runtime_error("'sphere.radius' must have one radius per sphere");
      spheres.resize(numPrims * 4);
      for (size_t i = 0; i < numPrims; i++) {
        spheres[4 * i] = p[3 * i];
        spheres[4 * i + 1] = p[3 * i + 1];
        spheres[4 * i + 2] = p[3 * i + 2];
        spheres[4 * i + 3] = r.empty() ? radius : r[i];
      }
      dSpheres.upload(spheres);
    } else {
      throw std::runtime_error("unsupported geometry type '" + subtype + "'");
    }
    version++;
  }
};

struct Material : public ManagedObject
{
  hpt::DMaterial m{};
  hpt::DMedium inside{}, outside{};
  Material(const std::string &t)
  {
    managedType = OSP_MATERIAL;
    subtype = t;
  }
  std::string toString() const override { return "ospray::Material(" + subtype + ")"; }
  void commit() override
  {
    std::memset(&m, 0, sizeof(m));
    if (subtype == "obj") { // OBJ.cpp:32-57
      m.type = hpt::MAT_OBJ;
      const float kdDef[3] = {0.8f, 0.8f, 0.8f}, zero[3] = {0.f, 0.f, 0.f};
      float kd[3], tf[3], ks[3];
      getVec3f("kd", kd, kdDef);
      getVec3f("tf", tf, zero);
      getVec3f("ks", ks, zero);
      m.d = getFloat("d", 1.f);
      m.Ns = getFloat("ns", 10.f);
      vec3f Kd = v3(kd[0], kd[1], kd[2]), Ks = v3(ks[0], ks[1], ks[2]), Tf = v3(tf[0], tf[1], tf[2]);
      const float total = hpt::reduce_max(Kd + Ks + Tf);
      if (total > 1.f) {
        Kd = Kd / total;
        Ks = Ks / total;
        Tf = Tf / total;
      }
      m.Kd[0] = Kd.x, m.Kd[1] = Kd.y, m.Kd[2] = Kd.z;
      m.Ks[0] = Ks.x, m.Ks[1] = Ks.y, m.Ks[2] = Ks.z;
      m.Tf[0] = Tf.x, m.Tf[1] = Tf.y, m.Tf[2] = Tf.z;
    } else if (subtype == "metal") { // Metal.cpp:30-66
      m.type = hpt::MAT_METAL;
      const float etaDef[3] = {1.47f, 0.984f, 0.602f}, kDef[3] = {7.64f, 6.55f, 5.36f};
      getVec3f("eta", m.eta, etaDef);
      getVec3f("k", m.k, kDef);
      m.roughness = getFloat("roughness", 0.1f);
      m.spectral = 0;
      if (Data *ior = getData(this, "ior")) {
        if (ior->elemType != OSP_VEC3F)
          throw std::runtime_error("metal material: 'ior' must be an array of vec3f (wavelength, eta, k)");
        resampleIor(dataFloats(ior, 3, "metal 'ior'"), m.etaS, m.kS);
        m.spectral = 1;
      }
    } else if (subtype == "glass") { // Glass.cpp:36-62
      m.type = hpt::MAT_GLASS;
      const float etaInside = getFloat("etaInside", getFloat("eta", 1.5f));
      const float etaOutside = getFloat("etaOutside", 1.f);
      const float one[3] = {1.f, 1.f, 1.f};
      float acol[3], acolIn[3], acolOut[3];
      getVec3f("attenuationColor", acol, one);
      getVec3f("attenuationColorInside", acolIn, acol);
      getVec3f("attenuationColorOutside", acolOut, one);
      const float dist = std::max(getFloat("attenuationDistance", 1.0f), 1e-5f);
      inside.ior = etaInside;
      outside.ior = etaOutside;
      for (int k = 0; k < 3; k++) {
        inside.attenuation[k] = std::log(acolIn[k]) / dist;
        outside.attenuation[k] = std::log(acolOut[k]) / dist;
      }
    } else if (subtype == "luminous") { // Luminous.cpp:37-48
      m.type = hpt::MAT_LUMINOUS;
      const float one[3] = {1.f, 1.f, 1.f};
      float color[3];
      getVec3f("color", color, one);
      const float intensity = getFloat("intensity", 1.f);
      for (int k = 0; k < 3; k++)
        m.radiance[k] = color[k] * intensity;
      m.transparency = getFloat("transparency", 0.f);
    } else {
      throw std::runtime_error("unsupported material type '" + subtype + "'");
    }
  }
  // Luminous::isEmissive (Luminous.h:25-28)
  bool emissive() const
  {
    return m.type == hpt::MAT_LUMINOUS && std::max(m.radiance[0], std::max(m.radiance[1], m.radiance[2])) > 0.f;
  }
  bool opaque() const
  {
    if (m.type == hpt::MAT_METAL)
      return true;
    if (m.type == hpt::MAT_GLASS)
      return false;
    if (m.type == hpt::MAT_LUMINOUS)
      return !(m.transparency != 0.f);
    const vec3f T = hpt::ld3(m.Tf) * m.d + v3(1.f - m.d);
    return T.x == 0.f && T.y == 0.f && T.z == 0.f;
  }
};

struct GeometricModel : public ManagedObject
{
  Ref<Geometry> geom;
  Ref<Material> material;
  uint32_t userID = hpt::kInvalidID;
  bool invertNormals = false;
  GeometricModel() { managedType = OSP_GEOMETRIC_MODEL; }
  std::string toString() const override { return "ospray::GeometricModel"; }
  void commit() override // GeometricModel.cpp:21-87
  {
    geom.reset((Geometry *)getObject("geometry", OSP_GEOMETRY));
    if (!geom)
      throw std::runtime_error("geometric model requires a 'geometry'");
    Material *mat = (Material *)getObject("material", OSP_MATERIAL);
    if (!mat) {
      if (Data *md = getData(this, "material")) {
        if (md->elemType == OSP_MATERIAL && md->size() > 0) {
          if (md->size() > 1)
            throw std::runtime_error("per-primitive material arrays are not supported by the hip device");
          mat = (Material *)md->get<ManagedObject *>(0);
        } else if (md->elemType == OSP_UINT || md->elemType == OSP_UCHAR) {
          throw std::runtime_error("renderer material lists are not supported by the hip device");
        }
      }
      if (findParamAny("material") && !mat && getUInt("material", 0xffffffffu) != 0xffffffffu)
        throw std::runtime_error("renderer material indices are not supported by the hip device");
    }
    material.reset(mat);
    userID = getUInt("id", hpt::kInvalidID);
    invertNormals = getBool("invertNormals", false);
  }
};

// ---- volumes ---------------------------------------------------------------------------

// "structuredRegular" volume (OpenVKL's structured-regular grid behind
// volume/Volume.cpp:27-160): vertex-centred voxels of a 3D "data" array
// (x fastest) at gridOrigin + i * gridSpacing, trilinear filter.  The
// voxels go to HBM as float; the macrocell (min, max) grid of kMacro^3
// cells is built at commit.
struct Volume : public ManagedObject
{
  int32_t dims[3] = {0, 0, 0}, mdims[3] = {0, 0, 0};
  float origin[3] = {0, 0, 0}, spacing[3] = {1, 1, 1};
  DevBuf dVoxels;
  std::vector<float> macro; // (min, max) per macrocell
  uint64_t version = 0;
  Volume(const std::string &t)
  {
    managedType = OSP_VOLUME;
    subtype = t;
  }
  std::string toString() const override { return "ospray::Volume(" + subtype + ")"; }
  void commit() override
  {
    Data *d = getData(this, "data");
    if (!d)
      throw std::runtime_error("structuredRegular volume requires a 3D 'data' array");
    for (int a = 0; a < 3; a++) {
      if (d->numItems[a] < 2 || d->numItems[a] > (1u << 20))
        throw std::runtime_error("structuredRegular volume: every dimension of 'data' must be in [2, 2^20]");
      dims[a] = (int32_t)d->numItems[a];
    }
    const float zero[3] = {0.f, 0.f, 0.f}, one[3] = {1.f, 1.f, 1.f};
    getVec3f("gridOrigin", origin, zero);
    getVec3f("gridSpacing", spacing, one);
    if (getBool("cellCentered", false))
      throw std::runtime_error("cell-centred structuredRegular volumes are not supported by the hip device");
    if (getInt("filter", OSP_VOLUME_FILTER_LINEAR) != OSP_VOLUME_FILTER_LINEAR)
      throw std::runtime_error("the hip device samples volumes with the trilinear filter only");
    (void)getFloat("background", 0.f);
    const size_t n = d->size();
    std::vector<float> vox(n);
    for (size_t i = 0; i < n; i++) {
      switch (d->elemType) {
      case OSP_FLOAT: vox[i] = d->get<float>(i); break;
      case OSP_DOUBLE: vox[i] = (float)d->get<double>(i); break;
      case OSP_UCHAR: vox[i] = (float)d->get<uint8_t>(i); break;
      case OSP_USHORT: vox[i] = (float)d->get<uint16_t>(i); break;
      case OSP_SHORT: vox[i] = (float)d->get<int16_t>(i); break;
      default: throw std::runtime_error("structuredRegular 'data' must be float, double, uchar, short or ushort");
      }
    }
    // macrocell m covers voxels [m * kMacro, (m + 1) * kMacro] per axis
    for (int a = 0; a < 3; a++)
      mdims[a] = (dims[a] - 1 + hpt::kMacro - 1) / hpt::kMacro;
    macro.assign(2 * (size_t)mdims[0] * mdims[1] * mdims[2], 0.f);
    for (int mz = 0; mz < mdims[2]; mz++)
      for (int my = 0; my < mdims[1]; my++)
        for (int mx = 0; mx < mdims[0]; mx++) {
          float lo = hpt::kInf, hi = -hpt::kInf;
          for (int z = mz * hpt::kMacro; z <= std::min((mz + 1) * hpt::kMacro, dims[2] - 1); z++)
            for (int y = my * hpt::kMacro; y <= std::min((my + 1) * hpt::kMacro, dims[1] - 1); y++) {
              const float *row = vox.data() + ((size_t)z * dims[1] + y) * dims[0];
              for (int x = mx * hpt::kMacro; x <= std::min((mx + 1) * hpt::kMacro, dims[0] - 1); x++) {
                lo = std::min(lo, row[x]);
                hi = std::max(hi, row[x]);
              }
            }
          const size_t m = ((size_t)mz * mdims[1] + my) * mdims[0] + mx;
          macro[2 * m] = lo;
          macro[2 * m + 1] = hi;
        }
    dVoxels.upload(vox);
    version++;
  }
};

// "piecewiseLinear" transfer function (LinearTransferFunction.cpp:18-160,
// TransferFunction.cpp:19): colour and opacity control points spread
// evenly over the value range, and the precomputed maximum opacity of
// every pair of kTfSubranges value sub-ranges
struct TransferFunction : public ManagedObject
{
  std::vector<float> color, opacity;
  float valueRange[2] = {0.f, 1.f};
  float maxOpacity[hpt::kTfSubranges * hpt::kTfSubranges];
  DevBuf dColor, dOpacity;
  TransferFunction(const std::string &t)
  {
    managedType = OSP_TRANSFER_FUNCTION;
    subtype = t;
  }
  std::string toString() const override { return "ospray::TransferFunction(" + subtype + ")"; }
  void commit() override
  {
    Data *c = getData(this, "color");
    Data *o = getData(this, "opacity");
    if (!c || !o || c->size() == 0 || o->size() == 0)
      throw std::runtime_error("piecewiseLinear transfer function requires 'color' and 'opacity' arrays");
    color = dataFloats(c, 3, "transfer function 'color'");
    opacity = dataFloats(o, 1, "transfer function 'opacity'");
    valueRange[0] = 0.f;
    valueRange[1] = 1.f;
    if (const Param *pr = findParam("value", OSP_BOX1F))
      std::memcpy(valueRange, pr->bytes, 8);
    // precomputeMaxOpacityRanges (LinearTransferFunction.cpp:108-160)
    const int maxOpacityDim = (int)opacity.size() - 1;
    const int maxPre = hpt::kTfSubranges - 1;
    auto at = [&](int i, int j) -> float & { return maxOpacity[i * hpt::kTfSubranges + j]; };
    for (int i = 0; i < hpt::kTfSubranges; i++) {
      const int lo = (int)std::floor(maxOpacityDim * (float)i / maxPre);
      const int hi = (int)std::ceil(maxOpacityDim * (float)i / maxPre);
      float m = opacity[lo];
      for (int k = lo; k <= hi; k++)
        m = std::max(m, opacity[k]);
      at(i, i) = m;
    }
    for (int i = 0; i < hpt::kTfSubranges; i++)
      for (int j = i + 1; j < hpt::kTfSubranges; j++) {
        const int lo = (int)std::floor(maxOpacityDim * (float)i / maxPre);
        const int hi = (int)std::ceil(maxOpacityDim * (float)j / maxPre);
        float m = at(i, i);
        for (int k = lo; k <= hi; k++)
          m = std::max(m, opacity[k]);
        at(i, j) = m;
      }
    for (int i = 0; i < hpt::kTfSubranges; i++) // inverted ranges: zero-filled
      for (int j = 0; j < i; j++)
        at(i, j) = 0.f;
    dColor.upload(color);
    dOpacity.upload(opacity);
  }
};

// VolumetricModel (volume/VolumetricModel.cpp:31-82)
struct VolumetricModel : public ManagedObject
{
  Ref<Volume> volume;
  Ref<TransferFunction> tf;
  float densityScale = 1.f, anisotropy = 0.f;
  uint32_t userID = hpt::kInvalidID;
  DevBuf dMajorant; // densityScale * max opacity per macrocell
  VolumetricModel() { managedType = OSP_VOLUMETRIC_MODEL; }
  std::string toString() const override { return "ospray::VolumetricModel"; }
  void commit() override
  {
    volume.reset((Volume *)getObject("volume", OSP_VOLUME));
    if (!volume)
      throw std::runtime_error("volumetric model requires a 'volume'");
    tf.reset((TransferFunction *)getObject("transferFunction", OSP_TRANSFER_FUNCTION));
    if (!tf)
      throw std::runtime_error("volumetric model requires a 'transferFunction'");
    densityScale = getFloat("densityScale", 1.f);
    anisotropy = getFloat("anisotropy", 0.f);
    (void)getFloat("gradientShadingScale", 0.f);
    userID = getUInt("id", hpt::kInvalidID);
  }
  // the majorant grid of the current volume, transfer function and
  // densityScale (built when the world is committed)
  void buildMajorant()
  {
    const Volume &V = *volume.p;
    const TransferFunction &T = *tf.p;
    const size_t n = V.macro.size() / 2;
    std::vector<float> mj(n);
    for (size_t m = 0; m < n; m++)
      mj[m] = densityScale * hpt::tfMaxOpacity(T.maxOpacity, T.valueRange[0], T.valueRange[1], V.macro[2 * m], V.macro[2 * m + 1]);
    dMajorant.upload(mj);
  }
  hpt::DVolume device()
  {
    buildMajorant();
    hpt::DVolume v;
    std::memset(&v, 0, sizeof(v));
    const Volume &V = *volume.p;
    v.voxels = V.dVoxels.as<const float>();
    v.majorant = dMajorant.as<const float>();
    v.tfColor = tf->dColor.as<const float>();
    v.tfOpacity = tf->dOpacity.as<const float>();
    for (int a = 0; a < 3; a++) {
      v.dims[a] = V.dims[a];
      v.mdims[a] = V.mdims[a];
      v.origin[a] = V.origin[a];
      v.spacing[a] = V.spacing[a];
      v.upper[a] = V.origin[a] + (float)(V.dims[a] - 1) * V.spacing[a];
    }
    v.numColors = (int32_t)(tf->color.size() / 3);
    v.numOpacities = (int32_t)tf->opacity.size();
    v.tfLower = tf->valueRange[0];
    v.tfUpper = tf->valueRange[1];
    v.densityScale = densityScale;
    v.anisotropy = anisotropy;
    v.userID = userID;
    return v;
  }
};

struct Group : public ManagedObject
{
  std::vector<GeometricModel *> models;
  std::vector<std::unique_ptr<Ref<GeometricModel>>> refs;
  std::vector<VolumetricModel *> volumes;
  std::vector<std::unique_ptr<Ref<VolumetricModel>>> volumeRefs;
  Group() { managedType = OSP_GROUP; }
  std::string toString() const override { return "ospray::Group"; }
  void commit() override // Group.cpp:69-209
  {
    models = dataObjects<GeometricModel>(getData(this, "geometry"), OSP_GEOMETRIC_MODEL);
    refs.clear();
    for (auto *m : models) {
      if (!m)
        throw std::runtime_error("group 'geometry' holds a null model");
      refs.emplace_back(new Ref<GeometricModel>());
      refs.back()->reset(m);
    }
    volumes = dataObjects<VolumetricModel>(getData(this, "volume"), OSP_VOLUMETRIC_MODEL);
    volumeRefs.clear();
    for (auto *v : volumes) {
      if (!v || !v->volume)
        throw std::runtime_error("group 'volume' holds a null or uncommitted volumetric model");
      volumeRefs.emplace_back(new Ref<VolumetricModel>());
      volumeRefs.back()->reset(v);
    }
    if (getData(this, "light"))
      throw std::runtime_error("group (instance) lights are not supported by the hip device");
    (void)getData(this, "clippingGeometry");
  }
};

struct Instance : public ManagedObject
{
  Ref<Group> group;
  float xfm[12] = {1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0};
  uint32_t userID = hpt::kInvalidID;
  Instance() { managedType = OSP_INSTANCE; }
  std::string toString() const override { return "ospray::Instance"; }
  void commit() override // Instance.cpp:20-34
  {
    group.reset((Group *)getObject("group", OSP_GROUP));
    if (!group)
      throw std::runtime_error("instance requires a 'group'");
    const float id[12] = {1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0};
    std::memcpy(xfm, id, sizeof(xfm));
    if (!getAffine3f("transform", xfm))
      getAffine3f("xfm", xfm);
    userID = getUInt("id", hpt::kInvalidID);
  }
};

struct Light : public ManagedObject
{
  hpt::DLight l{};
  Light(const std::string &t)
  {
    managedType = OSP_LIGHT;
    subtype = t;
  }
  std::string toString() const override { return "ospray::Light(" + subtype + ")"; }
  void commit() override
  {
    std::memset(&l, 0, sizeof(l));
    l.visible = getBool("visible", true) ? 1 : 0; // Light.cpp:20-24
    const float one[3] = {1.f, 1.f, 1.f};
    float col[3];
    getVec3f("color", col, one);
    const float intensity = getFloat("intensity", 1.f);
    const vec3f colored = v3(col[0], col[1], col[2]) * intensity;
    auto put = [](float *dst, vec3f v) {
      dst[0] = v.x;
      dst[1] = v.y;
      dst[2] = v.z;
    };
    if (subtype == "quad") { // QuadLight.cpp:68-105, QuadLight.ispc:15-33
      l.type = hpt::LIGHT_QUAD;
      const float p0[3] = {0.f, 0.f, 0.f}, e1d[3] = {1.f, 0.f, 0.f}, e2d[3] = {0.f, 1.f, 0.f};
      getVec3f("position", l.position, p0);
      getVec3f("edge1", l.edge1, e1d);
      getVec3f("edge2", l.edge2, e2d);
      const uint32_t q = getUInt("intensityQuantity", OSP_INTENSITY_QUANTITY_RADIANCE);
      const vec3f e1 = hpt::ld3(l.edge1), e2 = hpt::ld3(l.edge2);
      const float area = hpt::length(hpt::cross(e1, e2));
      vec3f rad = v3(0.f);
      if (q == OSP_INTENSITY_QUANTITY_INTENSITY)
        rad = colored / area;
      else if (q == OSP_INTENSITY_QUANTITY_POWER)
        rad = colored / (hpt::kPi * area);
      else if (q == OSP_INTENSITY_QUANTITY_RADIANCE)
        rad = colored;
      put(l.radiance, rad);
      const vec3f nd = hpt::cross(e2, e1);
      l.ppdf = hpt::rcp(hpt::length(nd));
      put(l.nnormal, nd * l.ppdf);
    } else if (subtype == "distant") { // DirectionalLight.cpp:20-106
      l.type = hpt::LIGHT_DISTANT;
      const float dDef[3] = {0.f, 0.f, 1.f};
      float d[3];
      getVec3f("direction", d, dDef);
      const vec3f dir = hpt::neg(hpt::normalize(v3(d[0], d[1], d[2])));
      const float ad = hpt::clampf(getFloat("angularDiameter", 0.f), 0.f, 180.f);
      l.cosAngle = std::cos(ad * 0.5f * 0.01745329251994329577f);
      const uint32_t q = getUInt("intensityQuantity", OSP_INTENSITY_QUANTITY_IRRADIANCE);
      vec3f irr = v3(0.f);
      if (q == OSP_INTENSITY_QUANTITY_RADIANCE)
        irr = (2.0f * hpt::kPi * (1.0f - l.cosAngle)) * colored;
      else if (q == OSP_INTENSITY_QUANTITY_IRRADIANCE)
        irr = colored;
      put(l.irradiance, irr);
      const hpt::linear3f fr = hpt::frame(dir);
      put(l.frame, fr.vx);
      put(l.frame + 3, fr.vy);
      put(l.frame + 6, fr.vz);
      l.pdf = l.cosAngle < 0.99999988f ? hpt::rcp(hpt::kTwoPi * (1.0f - l.cosAngle)) : hpt::kInf;
    } else if (subtype == "ambient") { // AmbientLight.cpp:41-60
      l.type = hpt::LIGHT_AMBIENT;
      const uint32_t q = getUInt("intensityQuantity", OSP_INTENSITY_QUANTITY_RADIANCE);
      vec3f rad = v3(0.f);
      if (q == OSP_INTENSITY_QUANTITY_IRRADIANCE)
        rad = colored / hpt::kPi;
      else if (q == OSP_INTENSITY_QUANTITY_RADIANCE)
        rad = colored;
      put(l.radiance, rad);
    } else {
      throw std::runtime_error("unsupported light type '" + subtype + "'");
    }
  }
};

// BLAS of one list of geometries (shared by groups with the same list):
// built on the device (nodes and records in dNodes/dPrims) or, as the
// fallback, on the host (nodes/prims)
struct Blas
{
  std::vector<hpt::DNode4> nodes;
  std::vector<hpt::DPrim> prims;
  DevBuf dNodes, dPrims;
  bool onDevice = false;
  size_t numNodes = 0, numPrims = 0;
  hpt::BuildBox bounds;
  int maxDepth = 0;
  int stackNeed = 0;
  double buildMs = 0.0;
};

struct World : public ManagedObject
{
  HipDevice *dev;
  std::vector<std::unique_ptr<Ref<Instance>>> instRefs;
  std::vector<std::unique_ptr<Ref<Light>>> lightRefs;
  std::vector<Instance *> instances;
  std::vector<Material *> materials; // index order of the material table
  DevBuf dNodes, dPrims, dInstances, dModels, dMaterials, dMedia, dLights, dLightsCDF, dWorldLightsCDF, dGeoCDF;
  DevBuf dVolumes, dVolInsts;
  hpt::DScene scene{};
  bool built = false;
  int64_t totalNodes = 0, totalPrims = 0;
  int maxDepth = 0, stackNeed = 0;
  double blasBuildMs = 0.0; // device BLAS builds of the last commit (0 for cached or host builds)
  int deviceBlas = 0; // BLASes of this world built on the device
  float bounds[6] = {0, 0, 0, 0, 0, 0};
  World(HipDevice *d) : dev(d) { managedType = OSP_WORLD; }
  std::string toString() const override { return "ospray::World"; }
  void commit() override;
};

struct Camera : public ManagedObject
{
  float pos[3], dir[3], up[3], fovy, aspect, nearClip;
  Camera(const std::string &t)
  {
    managedType = OSP_CAMERA;
    subtype = t;
  }
  std::string toString() const override { return "ospray::Camera(" + subtype + ")"; }
  void commit() override // Camera.cpp:38-52, PerspectiveCamera.cpp:30-44
  {
    const float z[3] = {0.f, 0.f, 0.f}, zd[3] = {0.f, 0.f, 1.f}, yu[3] = {0.f, 1.f, 0.f};
    getVec3f("position", pos, z);
    getVec3f("direction", dir, zd);
    getVec3f("up", up, yu);
    nearClip = std::max(getFloat("nearClip", 1e-6f), 0.f);
    fovy = getFloat("fovy", 60.f);
    aspect = getFloat("aspect", 1.f);
    if (getFloat("apertureRadius", 0.f) != 0.f)
      throw std::runtime_error("perspective camera: depth of field is not supported by the hip device");
    if (getUInt("stereoMode", OSP_STEREO_NONE) != OSP_STEREO_NONE)
      throw std::runtime_error("perspective camera: stereo is not supported by the hip device");
  }
  // PerspectiveCamera::commit set-up (PerspectiveCamera.cpp:44-110)
  void setup(hpt::DFrame &F) const
  {
    const float imgY = 2.f * std::tan(0.5f * fovy * 0.01745329251994329577f);
    const float imgX = imgY * aspect;
    vec3f d00 = hpt::normalize(v3(dir[0], dir[1], dir[2]));
    vec3f du = hpt::normalize(hpt::cross(d00, v3(up[0], up[1], up[2])));
    vec3f dv = hpt::cross(du, d00);
    du = du * imgX;
    dv = dv * imgY;
    d00 = d00 - 0.5f * du - 0.5f * dv;
    auto put = [](float *dst, vec3f v) {
      dst[0] = v.x;
      dst[1] = v.y;
      dst[2] = v.z;
    };
    put(F.camOrg, v3(pos[0], pos[1], pos[2]));
    put(F.camDir00, d00);
    put(F.camDu, du);
    put(F.camDv, dv);
    put(F.camDirection, v3(dir[0], dir[1], dir[2]));
    F.nearClip = nearClip;
  }
};

struct Renderer : public ManagedObject
{
  int spp = 1, maxDepth = 20, rouletteDepth = 5, maxScatteringEvents = 20, numLightSamples = -1;
  bool limitIndirect = true, geometryLights = true;
  float minContribution = 0.001f, maxRadiance = hpt::kInf, errorThreshold = 0.f;
  float bg[4] = {0, 0, 0, 0};
  int filterType = 2;
  float filterWidth = 3.f;
  int lutSize = 0;
  DevBuf dLutX, dLutY;
  // neural radiance cache (hip-only parameters, no reference counterpart):
  // "nrc" on/off, the depth of the cache vertex, training paths and Adam
  // steps per frame, learning rate, initial-weight seed; the network and
  // its optimiser state persist across frames (online training) until
  // "nrcReset" is committed or the seed changes
  bool nrc = false, nrcSelfTrain = true;
  int nrcBounce = 1, nrcTrainPaths = 8192, nrcTrainSteps = 4;
  float nrcLr = 1e-2f;
  uint32_t nrcSeed = 1;
  bool nrcInit = false;
  int nrcStep = 0;
  int64_t nrcQueries = 0, nrcRecords = 0; // of the last frame
  // fraction of training slots whose path reaches the cache vertex (running
  // estimate): the training stride aims at nrcTrainPaths records per pass
  double nrcSurvival = 1.0;
  DevBuf nW, nWT, nMaster, nM, nV, nPartial, nTileLoss, nLoss;
  HptNrc nrcNet{};
  Renderer(const std::string &t)
  {
    managedType = OSP_RENDERER;
    subtype = t;
  }
  std::string toString() const override { return "ospray::PathTracer"; }
  void commit() override // Renderer.cpp:38-92, PathTracer.cpp:29-51
  {
    spp = std::max(1, getInt("pixelSamples", 1));
    maxDepth = std::max(0, getInt("maxPathLength", 20));
    minContribution = getFloat("minContribution", 0.001f);
    errorThreshold = getFloat("varianceThreshold", 0.f);
    rouletteDepth = getInt("roulettePathLength", 5);
    maxScatteringEvents = getInt("maxScatteringEvents", 20);
    maxRadiance = getFloat("maxContribution", hpt::kInf);
    numLightSamples = getInt("lightSamples", -1);
    limitIndirect = getBool("limitIndirectLightSamples", true);
    const float bgF = getFloat("backgroundColor", 0.f);
    const float bg3d[3] = {bgF, bgF, bgF};
    float bg3[3];
    getVec3f("backgroundColor", bg3, bg3d);
    const float bg4d[4] = {bg3[0], bg3[1], bg3[2], 0.f};
    getVec4f("backgroundColor", bg, bg4d);
    if (getBool("backgroundRefraction", false))
      throw std::runtime_error("pathtracer: 'backgroundRefraction' is not supported by the hip device");
    float scp[4];
    const float scpd[4] = {0, 0, 0, 0};
    getVec4f("shadowCatcherPlane", scp, scpd);
    if (scp[0] != 0.f || scp[1] != 0.f || scp[2] != 0.f)
      throw std::runtime_error("pathtracer: shadow catcher is not supported by the hip device");
    geometryLights = getBool("geometryLights", true);
    (void)getFloat("mipMapBias", 0.f);
    nrc = getBool("nrc", false);
    nrcBounce = std::max(1, getInt("nrcBounce", 1));
    nrcTrainPaths = std::max(64, getInt("nrcTrainPaths", 8192));
    nrcTrainSteps = std::max(1, getInt("nrcTrainSteps", 4));
    nrcSelfTrain = getBool("nrcSelfTrain", true);
    nrcLr = getFloat("nrcLearningRate", 1e-2f);
    const uint32_t seed = (uint32_t)getInt("nrcSeed", 1);
    if (getBool("nrcReset", false) || seed != nrcSeed)
      nrcInit = false;
    nrcSeed = seed;
    if (findParamAny("nrcReset"))
      removeParam("nrcReset");
    const uint32_t pf = getUInt("pixelFilter", OSP_PIXELFILTER_GAUSS);
    if (pf == OSP_PIXELFILTER_POINT) {
      filterType = 0;
      filterWidth = 0.f;
    } else if (pf == OSP_PIXELFILTER_BOX) {
      filterType = 1;
      filterWidth = 1.f;
    } else if (pf == OSP_PIXELFILTER_GAUSS) {
      // Gaussian LUT of width 3, 5 bins per pixel (PixelFilter.cpp:19-56,
      // LUTPixelFilter.ispc:17-120)
      filterType = 2;
      filterWidth = 3.f;
      const int n = (int)std::ceil(filterWidth) * 5;
      lutSize = n;
      std::vector<float> cx(n * n), cy(n);
      const float sigma = (0.5f * filterWidth) / 3.0f;
      const float sigmaSqr = sigma * sigma;
      for (int y = 0; y < n; y++) {
        const float fy = (y + 0.5f) * (1.0f / (float)n);
        for (int x = 0; x < n; x++) {
          const float fx = (x + 0.5f) * (1.0f / (float)n);
          const float ux = (fx - 0.5f) * filterWidth, uy = (fy - 0.5f) * filterWidth;
          const float distSqr = ux * ux + uy * uy;
          cx[y * n + x] = hpt::maxf(0.0f, hpt::fexp(-0.5f * (distSqr / sigmaSqr)));
        }
        cy[y] = hpt::distribution1DCreate(n, cx.data() + y * n);
      }
      hpt::distribution1DCreate(n, cy.data());
      dLutX.upload(cx);
      dLutY.upload(cy);
    } else {
      throw std::runtime_error("pixel filter type not supported by the hip device");
    }
  }
};

struct FrameBuffer : public ManagedObject
{
  HipDevice *dev;
  int w, h;
  OSPFrameBufferFormat format;
  uint32_t channels;
  int frameID = -1;
  std::mt19937 mtGen{43};
  bool pickOdd = false;
  bool accumulateVariance = false;
  int targetFrames = 0;
  bool projectedDepth = false;
  bool varianceValid = false;
  float errorThreshold = 0.f;
  DevBuf color, variance, depth, normal, albedo, firstNormal, position, primID, objID, instID, taskVar;
  DevBuf accRGBA, accAux, accZ, accIDs, converted, varianceOut;
  DevBuf cancelLatch; // set on the device when a frame of this framebuffer is cancelled
  std::map<int, std::vector<unsigned char>> staging;
  float *hostVariance = nullptr; // pinned
  FrameBuffer(HipDevice *d, int w_, int h_, OSPFrameBufferFormat f, uint32_t ch);
  ~FrameBuffer() override;
  std::string toString() const override { return "ospray::LocalFrameBuffer"; }
  void commit() override // FrameBuffer.cpp:85-93
  {
    targetFrames = std::max(0, getInt("targetFrames", 0));
    projectedDepth = getBool("projectedDepth", false);
    if (!(channels & OSP_FB_ACCUM))
      targetFrames = 1;
  }
  void clear() // FrameBuffer.cpp:95-103
  {
    frameID = -1;
    mtGen.seed(43);
    varianceValid = false;
  }
};

struct Future : public ManagedObject
{
  HipDevice *dev;
  hipEvent_t start = nullptr, rendered = nullptr, finished = nullptr;
  // one event after each pass's accumulation (ospGetProgress)
  std::vector<hipEvent_t> passDone;
  // the frame's generation: ospCancel writes it into the device's
  // cancellation ring (0: a captured frame, not cancellable)
  uint32_t gen = 0;
  Ref<ManagedObject> fb, renderer, camera, world;
  bool empty = false;
  Future(HipDevice *d) : dev(d) { managedType = OSP_FUTURE; }
  ~Future() override;
  std::string toString() const override { return "ospray::Future"; }
};

// ---- the device --------------------------------------------------------------------------

struct KernelTimer
{
  std::string name;
  int64_t launches = 0;
  double totalMs = 0;
};

class HipDevice : public Device
{
 public:
  hipStream_t stream = nullptr;
  int numCUs = 0;
  std::string gpuName;
  // path state lanes (frames in flight), shared by all framebuffers of
  // this device; `lane` is the one frames are enqueued on
  std::vector<std::unique_ptr<PathLane>> lanes;
  PathLane *lane = nullptr;
  int batchStreams = 4; // lanes a batch of views spreads over
  int slotTileWidth = 8; // pixel tile of 64 path slots: 4x16, 8x8, 16x4 or 32x2
  int frameLanes = 1; // lanes the passes of one frame alternate over (1: one stream)
  bool inBatch = false; // enqueuing the views of a batch
  int batchView = -1; // index of the view being enqueued in a batch
  // captured batches: one frame-index word per view (written before each
  // graph launch) and the variance-frame coin table of mt19937(43)
  DevBuf batchState, varianceCoins;
  uint32_t *batchStateHost = nullptr;
  size_t batchStateViews = 0;
  hipEvent_t passEvents[2] = {nullptr, nullptr};
  DevBuf stats;
  // radiance cache per-pass buffers: vertex records, queues, raw inputs,
  // bf16 features, predictions, training targets
  DevBuf nrcRec, nrcQueues, nrcRaw, nrcFeat, nrcOut, nrcTarget;
  size_t nrcSlots = 0;
  int residentTraceBlocks = 0, shadeBlocks = 0;
  size_t pathBudgetBytes = (size_t)16 << 30;
  // profileBounces: the stage timers of the bounce loop are kept per bounce
  // ("extend@2", ...).  overlapShadow: the shadow-ray stages of a bounce run
  // on a second stream beside the next bounce's extension (eager frames of
  // one lane).
  bool profile = false, travStats = false, profileBounces = false, overlapShadow = true;
  // BLAS build: on the host (binned SAH, the default: about 10 % fewer node
  // visits per ray than the device build on the 4.2M-triangle terrain) or on
  // the device (Morton sort + PLOC, about 8x faster to commit); device
  // parameter "hostBvh".  Pinned words for the device build's readbacks.
  bool hostBvh = true;
  int bvhRadius = 16; // the device build's clustering window ("bvhRadius", 1..32)
  uint64_t *bvhWords = nullptr;
  uint64_t *bvhHostWords()
  {
    if (!bvhWords)
      HIP_CHECK(hipHostMalloc((void **)&bvhWords, 4 * sizeof(uint64_t), hipHostMallocDefault));
    return bvhWords;
  }
  bool capturing = false; // enqueuing a batch of views into a HIP graph
  // ospCancel: pinned host words, one per frame generation modulo
  // kCancelRing; cancelling frame g writes g into word g % kCancelRing,
  // which the frame's per-bounce counter resets read
  static constexpr uint32_t kCancelRing = 256;
  uint32_t *cancelRing = nullptr, *cancelRingDev = nullptr;
  uint32_t frameGen = 0;
  std::vector<KernelTimer> timers;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pendingEvents;
  std::vector<hipEvent_t> timerEvents; // recorded and read timer events, for reuse
  // profileStages: the stages the kernel timers cover (comma-separated
  // names as in the timer list, without the "@bounce" suffix; empty: all)
  std::vector<std::string> profileStages;
  std::map<std::vector<std::pair<const Geometry *, uint64_t>>, std::shared_ptr<Blas>> blasCache;

  HipDevice() {}
  ~HipDevice() override
  {
    collectTimers();
    for (auto e : timerEvents)
      (void)hipEventDestroy(e);
    releaseGraph();
    for (auto e : viewEvents)
      (void)hipEventDestroy(e);
    if (batchStart)
      (void)hipEventDestroy(batchStart);
    if (bvhWords)
      (void)hipHostFree(bvhWords);
    if (batchStateHost)
      (void)hipHostFree(batchStateHost);
    if (cancelRing) {
      if (stream)
        (void)hipStreamSynchronize(stream);
      (void)hipHostFree(cancelRing);
    }
    for (auto e : laneEvents)
      (void)hipEventDestroy(e);
    for (auto e : passEvents)
      if (e)
        (void)hipEventDestroy(e);
    lanes.clear();
    if (stream)
      (void)hipStreamDestroy(stream);
  }

  void commitDevice() override
  {
    deviceID = getInt("deviceID", 0);
    profile = getBool("profileKernels", false);
    travStats = getBool("traversalStats", false);
    profileBounces = getBool("profileBounces", false);
    profileStages.clear();
    {
      const std::string list = getString("profileStages", "");
      size_t a = 0;
      while (a <= list.size()) {
        const size_t e = std::min(list.find(',', a), list.size());
        if (e > a)
          profileStages.push_back(list.substr(a, e - a));
        a = e + 1;
      }
    }
    overlapShadow = getBool("overlapShadow", overlapShadow);
    traceBlocksPerCU = std::max(1, std::min(64, getInt("traceBlocksPerCU", traceBlocksPerCU)));
    shadeBlocksPerCU = std::max(1, std::min(64, getInt("shadeBlocksPerCU", shadeBlocksPerCU)));
    batchStreams = std::max(1, std::min(8, getInt("batchStreams", batchStreams)));
    frameLanes = std::max(1, std::min(8, getInt("frameLanes", frameLanes)));
    hostBvh = getBool("hostBvh", hostBvh);
    bvhRadius = std::max(1, std::min(32, getInt("bvhRadius", bvhRadius)));
    if (stream) {
      HIP_CHECK(hipStreamSynchronize(stream));
      sizeGrids();
    }
    // defaults to the current budget, so a recommit keeps the tuned value
    const int budgetMB = getInt("pathBudgetMB", (int)(pathBudgetBytes >> 20));
    pathBudgetBytes = (size_t)std::max(64, budgetMB) << 20;
    // last, so that an invalid value leaves the other parameters committed
    const int tw = getInt("tileWidth", slotTileWidth);
    if (tw != 4 && tw != 8 && tw != 16 && tw != 32)
      throw std::runtime_error("device parameter tileWidth must be 4, 8, 16 or 32");
    slotTileWidth = tw;
  }

  // GPU state is created on first use, so the object layer works (and
  // reports errors) on a host without a GPU
  int deviceID = 0;
  void gpu()
  {
    if (!stream) {
      const int dev = deviceID;
      HIP_CHECK(hipSetDevice(dev));
      HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
      hipDeviceProp_t prop;
      HIP_CHECK(hipGetDeviceProperties(&prop, dev));
      numCUs = prop.multiProcessorCount;
      gpuName = prop.gcnArchName;
      stats.ensure(128);
      HIP_CHECK(hipMemset(stats.ptr, 0, 128));
      HIP_CHECK(hipHostMalloc((void **)&cancelRing, kCancelRing * sizeof(uint32_t),
          hipHostMallocMapped | hipHostMallocCoherent));
      std::memset(cancelRing, 0, kCancelRing * sizeof(uint32_t));
      HIP_CHECK(hipHostGetDevicePointer((void **)&cancelRingDev, cancelRing, 0));
      ensureLane(0);
      lane = lanes[0].get();
      sizeGrids();
    }
  }

  // lane i (created with its stream, counters and scratch rows on first use)
  PathLane &ensureLane(int i)
  {
    while ((int)lanes.size() <= i) {
      auto L = std::make_unique<PathLane>();
      if (lanes.empty()) {
        L->stream = stream;
      } else {
        HIP_CHECK(hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking));
        L->ownsStream = true;
      }
      L->counters.ensure(64);
      if (residentTraceBlocks > 0)
        ensureScratch(*L);
      L->paths.counters = L->counters.as<int32_t>();
      L->paths.overflow = L->overflow.as<uint32_t>();
      L->paths.stats = (unsigned long long *)stats.ptr;
      lanes.push_back(std::move(L));
    }
    return *lanes[i];
  }

  // resident grids of the bounce kernels (grid-stride loops) and the
  // traversal-stack overflow rows they need
  int traceBlocksPerCU = 12, shadeBlocksPerCU = 8;
  void sizeGrids()
  {
    releaseGraph(); // a captured batch refers to the grid sizes and the overflow rows
    residentTraceBlocks = std::max(1, numCUs) * traceBlocksPerCU;
    shadeBlocks = std::max(1, numCUs) * shadeBlocksPerCU;
    for (auto &L : lanes) {
      HIP_CHECK(hipStreamSynchronize(L->stream));
      ensureScratch(*L);
      L->paths.overflow = L->overflow.as<uint32_t>();
    }
  }

  // the traversal scratch rows of a lane, cleared when (re)allocated: the
  // occlusion pass keeps each row's recorded-hit count at 0 between rays
  void ensureScratch(PathLane &L)
  {
    const size_t bytes = (size_t)residentTraceBlocks * hpt_trace_block_size() * hpt_overflow_depth() * 4;
    if (L.overflow.ptr && L.overflow.bytes >= bytes)
      return;
    if (capturing)
      throw std::runtime_error("batched views: traversal scratch would grow while capturing");
    L.overflow.ensure(bytes);
    HIP_CHECK(hipMemset(L.overflow.ptr, 0, bytes));
    if (L.side) { // the side stream's rows follow the grid size too
      L.sideOverflow.ensure(bytes);
      HIP_CHECK(hipMemset(L.sideOverflow.ptr, 0, bytes));
    }
  }

  // the side stream of a lane, its scratch rows and events (first use)
  hipStream_t ensureSide(PathLane &L)
  {
    if (!L.side) {
      HIP_CHECK(hipStreamCreateWithFlags(&L.side, hipStreamNonBlocking));
      for (int k = 0; k < 2; k++) {
        HIP_CHECK(hipEventCreateWithFlags(&L.evShade[k], hipEventDisableTiming));
        HIP_CHECK(hipEventCreateWithFlags(&L.evRes[k], hipEventDisableTiming));
      }
      const size_t bytes = (size_t)residentTraceBlocks * hpt_trace_block_size() * hpt_overflow_depth() * 4;
      L.sideOverflow.ensure(bytes);
      HIP_CHECK(hipMemset(L.sideOverflow.ptr, 0, bytes));
    }
    return L.side;
  }

  void commitObject(ManagedObject *obj) override
  {
    // objects that own device memory need the GPU
    const OSPDataType t = obj->managedType;
    if (t == OSP_GEOMETRY || t == OSP_WORLD || t == OSP_RENDERER)
      gpu();
    Device::commitObject(obj);
  }

  // ---- timing ----
  int timerIndex(const char *name)
  {
    for (size_t i = 0; i < timers.size(); i++)
      if (timers[i].name == name)
        return (int)i;
    timers.push_back({name, 0, 0.0});
    return (int)timers.size() - 1;
  }
  // grid of the near-tie pass: near-ties are rare, so a small resident grid
  // (at most the traversal grid, whose scratch rows it uses)
  int tieBlocks() const { return std::max(1, std::min(residentTraceBlocks, 2 * numCUs)); }
  // whether the kernel timers cover stage `name` ("extend" or "extend@3")
  bool timed(const char *name) const
  {
    if (profileStages.empty())
      return true;
    const char *at = std::strchr(name, '@');
    const size_t len = at ? (size_t)(at - name) : std::strlen(name);
    for (const auto &st : profileStages)
      if (st.size() == len && st.compare(0, len, name, len) == 0)
        return true;
    return false;
  }
  hipEvent_t timerEvent()
  {
    if (timerEvents.empty()) {
      hipEvent_t e;
      HIP_CHECK(hipEventCreate(&e));
      return e;
    }
    const hipEvent_t e = timerEvents.back();
    timerEvents.pop_back();
    return e;
  }
  // enqueues fn (a launch on `on`, the lane's stream by default), timed on
  // that stream with events when profiling (events are reused once read)
  template <typename F>
  void launch(const char *name, F &&fn, hipStream_t on = nullptr)
  {
    if (!profile || capturing || !timed(name)) {
      HIP_CHECK(fn());
      return;
    }
    const hipStream_t st = on ? on : lane->stream;
    const hipEvent_t a = timerEvent(), b = timerEvent();
    HIP_CHECK(hipEventRecord(a, st));
    HIP_CHECK(fn());
    HIP_CHECK(hipEventRecord(b, st));
    pendingEvents.push_back({timerIndex(name), {a, b}});
  }
  void collectTimers()
  {
    for (auto &pe : pendingEvents) {
      float ms = 0.f;
      if (hipEventSynchronize(pe.second.second) == hipSuccess
          && hipEventElapsedTime(&ms, pe.second.first, pe.second.second) == hipSuccess) {
        timers[pe.first].launches++;
        timers[pe.first].totalMs += ms;
      }
      timerEvents.push_back(pe.second.first);
      timerEvents.push_back(pe.second.second);
    }
    pendingEvents.clear();
  }

  // ---- objects ----
  ManagedObject *newObject(OSPDataType kind, const char *type) override
  {
    const std::string t(type ? type : "");
    switch (kind) {
    case OSP_CAMERA:
      if (t == "perspective")
        return new Camera(t);
      break;
    case OSP_GEOMETRY:
      if (t == "mesh" || t == "sphere")
        return new Geometry(t);
      break;
    case OSP_MATERIAL:
      if (t == "obj" || t == "metal" || t == "glass" || t == "luminous")
        return new Material(t);
      break;
    case OSP_LIGHT:
      if (t == "quad" || t == "distant" || t == "ambient")
        return new Light(t);
      break;
    case OSP_RENDERER:
      if (t == "pathtracer")
        return new Renderer(t);
      break;
    case OSP_VOLUME:
      if (t == "structuredRegular")
        return new Volume(t);
      break;
    case OSP_TRANSFER_FUNCTION:
      if (t == "piecewiseLinear")
        return new TransferFunction(t);
      break;
    case OSP_VOLUMETRIC_MODEL:
      return new VolumetricModel();
    case OSP_GEOMETRIC_MODEL:
      return new GeometricModel();
    case OSP_GROUP:
      return new Group();
    case OSP_INSTANCE:
      return new Instance();
    case OSP_WORLD:
      return new World(this);
    default:
      break;
    }
    static const std::map<OSPDataType, std::string> names = {{OSP_CAMERA, "camera"}, {OSP_GEOMETRY, "geometry"},
        {OSP_MATERIAL, "material"}, {OSP_LIGHT, "light"}, {OSP_RENDERER, "renderer"}, {OSP_VOLUME, "volume"},
        {OSP_TEXTURE, "texture"}, {OSP_TRANSFER_FUNCTION, "transfer function"},
        {OSP_IMAGE_OPERATION, "image operation"}, {OSP_VOLUMETRIC_MODEL, "volumetric model"}};
    auto it = names.find(kind);
    throw std::runtime_error("Could not find " + (it != names.end() ? it->second : std::string("object")) + " of type: " + t
        + " (not provided by the hip device)");
  }

  OSPBounds getBounds(ManagedObject *obj) override
  {
    OSPBounds b;
    for (int i = 0; i < 3; i++) {
      b.lower[i] = hpt::kInf;
      b.upper[i] = -hpt::kInf;
    }
    if (obj->managedType == OSP_WORLD) {
      World *w = (World *)obj;
      if (w->built)
        for (int i = 0; i < 3; i++) {
          b.lower[i] = w->bounds[i];
          b.upper[i] = w->bounds[3 + i];
        }
    } else if (obj->managedType == OSP_GEOMETRY) {
      Geometry *g = (Geometry *)obj;
      const std::vector<float> &v = g->geomType == hpt::GEOM_SPHERES ? g->spheres : g->vertex;
      const int stride = g->geomType == hpt::GEOM_SPHERES ? 4 : 3;
      for (size_t i = 0; i + 2 < v.size(); i += stride)
        for (int k = 0; k < 3; k++) {
          const float r = stride == 4 ? v[i + 3] : 0.f;
          b.lower[k] = std::min(b.lower[k], v[i + k] - r);
          b.upper[k] = std::max(b.upper[k], v[i + k] + r);
        }
    }
    return b;
  }

  // ---- framebuffer ----
  ManagedObject *frameBufferCreate(int w, int h, OSPFrameBufferFormat f, uint32_t ch) override
  {
    if (w <= 0 || h <= 0)
      throw std::runtime_error("framebuffer has invalid size. Dimensions must be greater than 0");
    gpu();
    return new FrameBuffer(this, w, h, f, ch);
  }

  const void *frameBufferMap(ManagedObject *o, OSPFrameBufferChannel ch) override // LocalFB.cpp:489-522
  {
    FrameBuffer *fb = (FrameBuffer *)o;
    HIP_CHECK(hipStreamSynchronize(stream));
    const size_t npix = (size_t)fb->w * fb->h;
    const DevBuf *src = nullptr;
    size_t bytes = 0;
    switch (ch) {
    case OSP_FB_COLOR:
      if (fb->format == OSP_FB_NONE || !(fb->channels & OSP_FB_COLOR))
        return nullptr;
      src = &fb->converted;
      bytes = npix * (fb->format == OSP_FB_RGBA32F ? 16 : 4);
      break;
    case OSP_FB_DEPTH: src = &fb->depth; bytes = npix * 4; break;
    case OSP_FB_NORMAL: src = &fb->normal; bytes = npix * 12; break;
    case OSP_FB_ALBEDO: src = &fb->albedo; bytes = npix * 12; break;
    case OSP_FB_FIRST_NORMAL: src = &fb->firstNormal; bytes = npix * 12; break;
    case OSP_FB_POSITION: src = &fb->position; bytes = npix * 12; break;
    case OSP_FB_ID_PRIMITIVE: src = &fb->primID; bytes = npix * 4; break;
    case OSP_FB_ID_OBJECT: src = &fb->objID; bytes = npix * 4; break;
    case OSP_FB_ID_INSTANCE: src = &fb->instID; bytes = npix * 4; break;
    default: return nullptr;
    }
    if (!src->ptr)
      return nullptr;
    auto &st = fb->staging[(int)ch];
    st.resize(bytes);
    HIP_CHECK(hipMemcpy(st.data(), src->ptr, bytes, hipMemcpyDeviceToHost));
    fb->refInc();
    return st.data();
  }

  void frameBufferUnmap(const void *mapped, ManagedObject *o) override
  {
    if (mapped)
      o->refDec();
  }

  float getVariance(ManagedObject *o) override // LocalFB.cpp:263-272
  {
    FrameBuffer *fb = (FrameBuffer *)o;
    if (!fb->variance.ptr || !fb->varianceValid)
      return hpt::kInf;
    HIP_CHECK(hipStreamSynchronize(stream));
    return *fb->hostVariance;
  }

  void resetAccumulation(ManagedObject *o) override { ((FrameBuffer *)o)->clear(); }

  // ---- rendering ----
  ManagedObject *renderFrame(ManagedObject *fbo, ManagedObject *ro, ManagedObject *co, ManagedObject *wo) override;
  int isReady(ManagedObject *fo, OSPSyncEvent e) override
  {
    Future *f = (Future *)fo;
    if (f->empty)
      return 1;
    hipEvent_t ev = e <= OSP_WORLD_RENDERED ? f->rendered : f->finished;
    return hipEventQuery(ev) == hipSuccess ? 1 : 0;
  }
  void wait(ManagedObject *fo, OSPSyncEvent e) override
  {
    Future *f = (Future *)fo;
    if (f->empty)
      return;
    HIP_CHECK(hipEventSynchronize(e <= OSP_WORLD_RENDERED ? f->rendered : f->finished));
    if (e >= OSP_FRAME_FINISHED)
      collectTimers();
  }
  // FrameBuffer::cancelFrame (FrameBuffer.cpp:180-187): the frame's
  // remaining passes and bounces do no work (the reference's render tasks
  // return early, PathTracer.ispc:231) and a cancelled frame leaves the
  // framebuffer's image as it was; the frame still advances the frame index
  // and runs its frame operations, as in the reference.
  void cancel(ManagedObject *fo) override
  {
    Future *f = (Future *)fo;
    if (f->empty || !f->gen || !cancelRing)
      return;
    __atomic_store_n(&cancelRing[f->gen % kCancelRing], f->gen, __ATOMIC_SEQ_CST);
  }
  // FrameBuffer::getCurrentProgress (FrameBuffer.cpp:169-178): the fraction
  // of the frame's passes whose samples have been accumulated (every pass
  // covers every pixel), 1 once the frame has finished.
  float getProgress(ManagedObject *fo) override
  {
    Future *f = (Future *)fo;
    if (isReady(fo, OSP_TASK_FINISHED))
      return 1.f;
    if (f->passDone.empty())
      return 0.f;
    size_t done = 0;
    for (auto e : f->passDone)
      done += hipEventQuery(e) == hipSuccess ? 1 : 0;
    return (float)done / (float)f->passDone.size();
  }
  float getTaskDuration(ManagedObject *fo) override
  {
    Future *f = (Future *)fo;
    if (f->empty)
      return 0.f;
    HIP_CHECK(hipEventSynchronize(f->finished));
    float ms = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms, f->start, f->finished));
    return ms * 1e-3f;
  }
  OSPPickResult pick(ManagedObject *fbo, ManagedObject *ro, ManagedObject *co, ManagedObject *wo, float x, float y) override;

  void ensurePaths(size_t slots, int maxRec, bool aux);

  // ---- batched views (no reference counterpart) ----
  // Many views (framebuffer, renderer, camera) of one world in one
  // enqueue: eagerly (mode 0) or as one captured HIP graph (mode 1), built
  // once per batch and relaunched; every call renders the next frame of
  // each view.  outMs[i] = GPU time from the batch start to the end of
  // view i.
  struct BatchView
  {
    ManagedObject *fb, *renderer, *camera;
  };
  std::vector<void *> graphKey;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graphExec = nullptr;
  std::vector<hipEvent_t> viewEvents;
  std::vector<hipEvent_t> laneEvents; // join of lane k into the device stream
  hipEvent_t batchStart = nullptr;
  void releaseGraph()
  {
    if (graphExec)
      (void)hipGraphExecDestroy(graphExec);
    if (graph)
      (void)hipGraphDestroy(graph);
    graphExec = nullptr;
    graph = nullptr;
    graphKey.clear();
  }
  // Lanes a batch runs on: views go round-robin to min(batchStreams,
  // views) lanes, which fork from and join into the device stream.  One
  // lane when two views share a framebuffer (their accumulation is
  // ordered), when a renderer uses the radiance cache (one cache state
  // per renderer, trained in view order) or with traversal statistics.
  int batchLanes(const std::vector<BatchView> &v) const
  {
    if (travStats)
      return 1;
    std::set<ManagedObject *> fbs;
    for (auto &x : v) {
      if (!fbs.insert(x.fb).second || ((Renderer *)x.renderer)->nrc)
        return 1;
    }
    return (int)std::max<size_t>(1, std::min<size_t>((size_t)batchStreams, v.size()));
  }
  void enqueueViews(const std::vector<BatchView> &v, ManagedObject *world, std::vector<ManagedObject *> &futures)
  {
    const int K = batchLanes(v);
    for (int k = 0; k < K; k++)
      ensureLane(k);
    while ((int)laneEvents.size() < K) {
      hipEvent_t e;
      HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      laneEvents.push_back(e);
    }
    HIP_CHECK(hipEventRecord(batchStart, stream));
    for (int k = 1; k < K; k++) // fork: the lanes start after everything queued so far
      HIP_CHECK(hipStreamWaitEvent(lanes[k]->stream, batchStart, 0));
    inBatch = true;
    try {
      for (size_t i = 0; i < v.size(); i++) {
        lane = lanes[i % K].get();
        batchView = (int)i;
        futures.push_back(renderFrame(v[i].fb, v[i].renderer, v[i].camera, world));
        HIP_CHECK(hipEventRecord(viewEvents[i], lane->stream));
      }
    } catch (...) {
      lane = lanes[0].get();
      inBatch = false;
      batchView = -1;
      throw;
    }
    lane = lanes[0].get();
    inBatch = false;
    batchView = -1;
    for (int k = 1; k < K; k++) { // join
      HIP_CHECK(hipEventRecord(laneEvents[k], lanes[k]->stream));
      HIP_CHECK(hipStreamWaitEvent(stream, laneEvents[k], 0));
    }
  }
  void renderBatch(const std::vector<BatchView> &v, ManagedObject *world, int mode, float *outMs)
  {
    gpu();
    while (viewEvents.size() < v.size()) {
      hipEvent_t e;
      HIP_CHECK(hipEventCreate(&e));
      viewEvents.push_back(e);
    }
    if (!batchStart)
      HIP_CHECK(hipEventCreate(&batchStart));
    std::vector<ManagedObject *> futures;
    bool timed = false;
    auto releaseFutures = [&]() {
      for (auto *f : futures)
        f->refDec();
      futures.clear();
    };
    try {
      if (mode == 0) {
        enqueueViews(v, world, futures);
      } else {
        // the graph holds the captured parameters: rebuilt when any object
        // of the batch is another one or has been committed since
        std::vector<void *> key{world, (void *)(uintptr_t)world->commitCounter};
        for (auto &x : v)
          for (ManagedObject *o : {x.fb, x.renderer, x.camera}) {
            key.push_back(o);
            key.push_back((void *)(uintptr_t)o->commitCounter);
          }
        if (key != graphKey || !graphExec) {
          releaseGraph();
          // this call's frame is rendered eagerly: it sizes every buffer,
          // so nothing is allocated (or synchronised) while capturing
          enqueueViews(v, world, futures);
          HIP_CHECK(hipStreamSynchronize(stream));
          releaseFutures();
          viewTimes(v.size(), outMs);
          timed = true;
          ensureBatchState(v.size());
          // the captured renderFrame calls begin a frame on the host; the
          // framebuffers' host frame state is restored after the capture,
          // so that every call renders exactly one frame of each view
          std::map<FrameBuffer *, FrameHostState> saved;
          for (auto &x : v)
            saved.emplace((FrameBuffer *)x.fb, FrameHostState((FrameBuffer *)x.fb));
          capturing = true;
          HIP_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
          try {
            enqueueViews(v, world, futures);
          } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(stream, &g);
            if (g)
              (void)hipGraphDestroy(g);
            capturing = false;
            for (auto &kv : saved)
              kv.second.restore(kv.first);
            throw;
          }
          HIP_CHECK(hipStreamEndCapture(stream, &graph));
          capturing = false;
          for (auto &kv : saved)
            kv.second.restore(kv.first);
          releaseFutures();
          HIP_CHECK(hipGraphInstantiate(&graphExec, graph, nullptr, nullptr, 0));
          graphKey = key;
        } else {
          // each launch begins the next frame of every view as
          // ospRenderFrame would (FrameBuffer::beginFrame, FrameBuffer.cpp:
          // 115-137), the device reading the index written here
          for (size_t i = 0; i < v.size(); i++) {
            FrameBuffer *fb = (FrameBuffer *)v[i].fb;
            if (fb->channels & OSP_FB_ACCUM)
              beginFrameHost(fb);
            batchStateHost[i] = (uint32_t)std::max(0, fb->frameID);
          }
          HIP_CHECK(hipMemcpyAsync(batchState.ptr, batchStateHost, v.size() * sizeof(uint32_t),
              hipMemcpyHostToDevice, stream));
          HIP_CHECK(hipGraphLaunch(graphExec, stream));
        }
      }
      HIP_CHECK(hipStreamSynchronize(stream));
      if (mode != 0) // the error estimates the captured frames left on the device
        for (auto &x : v) {
          FrameBuffer *fb = (FrameBuffer *)x.fb;
          if (fb->variance.ptr && fb->frameID >= 1) {
            HIP_CHECK(hipMemcpy(fb->hostVariance, fb->varianceOut.ptr, sizeof(float), hipMemcpyDeviceToHost));
            fb->varianceValid = true;
          }
        }
    } catch (...) {
      capturing = false;
      releaseFutures();
      throw;
    }
    releaseFutures();
    if (!timed)
      viewTimes(v.size(), outMs);
  }
  // GPU milliseconds from the batch start to the end of each view
  void viewTimes(size_t n, float *outMs)
  {
    for (size_t i = 0; outMs && i < n; i++)
      if (hipEventElapsedTime(&outMs[i], batchStart, viewEvents[i]) != hipSuccess)
        outMs[i] = -1.f;
  }
  // the host half of a framebuffer's frame state (what beginFrame changes)
  struct FrameHostState
  {
    int frameID;
    std::mt19937 mtGen;
    bool pickOdd, accumulateVariance, varianceValid;
    explicit FrameHostState(const FrameBuffer *fb)
        : frameID(fb->frameID), mtGen(fb->mtGen), pickOdd(fb->pickOdd), accumulateVariance(fb->accumulateVariance),
          varianceValid(fb->varianceValid)
    {
    }
    void restore(FrameBuffer *fb) const
    {
      fb->frameID = frameID;
      fb->mtGen = mtGen;
      fb->pickOdd = pickOdd;
      fb->accumulateVariance = accumulateVariance;
      fb->varianceValid = varianceValid;
    }
  };

  // device words of a captured batch's frame indices (one per view) and the
  // coin table of the variance frames: bit k is the k-th draw of
  // mt19937(43) (FrameBuffer.h:134-137), which a framebuffer takes on its
  // frame 2k.  The graph refers to batchState: growing it drops the graph.
  void ensureBatchState(size_t views)
  {
    if (views > batchStateViews) {
      releaseGraph();
      batchState.ensure(views * sizeof(uint32_t));
      if (batchStateHost)
        HIP_CHECK(hipHostFree(batchStateHost));
      batchStateHost = nullptr;
      HIP_CHECK(hipHostMalloc((void **)&batchStateHost, views * sizeof(uint32_t), hipHostMallocDefault));
      batchStateViews = views;
    }
    if (!varianceCoins.ptr) {
      std::vector<uint32_t> bits(hpt::kVarianceCoinBits / 32, 0u);
      std::mt19937 g(43);
      for (uint32_t k = 0; k < hpt::kVarianceCoinBits; k++)
        if (g() & 1u)
          bits[k >> 5] |= 1u << (k & 31u);
      varianceCoins.upload(bits);
    }
  }
  // FrameBuffer::beginFrame's host half (FrameBuffer.cpp:115-137)
  static void beginFrameHost(FrameBuffer *fb)
  {
    fb->frameID = (fb->channels & OSP_FB_ACCUM) ? fb->frameID + 1 : 0;
    if (fb->variance.ptr) {
      const bool evenFrame = (fb->frameID & 1) == 0;
      if (evenFrame)
        fb->pickOdd = (fb->mtGen() & 1u) != 0;
      fb->accumulateVariance = evenFrame != fb->pickOdd;
    }
  }
  // The batch ends when its last lane joins the device stream: the latest
  // of the views' end events (each measured from the batch start).
  static float batchMs(const std::vector<float> &ms)
  {
    float t = 0.f;
    for (float x : ms)
      t = std::max(t, x < 0.f ? 1e30f : x);
    return t;
  }
  // Autotuner of the batch: times the eager batch for each candidate of
  // the resident traversal grid, the shading grid and the path-state
  // budget (which sets the samples per pass, i.e. the wavefront-queue
  // sizes), then of the lane count, then of the slot tile's width; keeps
  // the fastest.  Returns its milliseconds; best[5] = traceBlocksPerCU,
  // shadeBlocksPerCU, pathBudgetMB, batchStreams, tileWidth.
  float tuneBatch(const std::vector<BatchView> &v, ManagedObject *world, int reps, int *best)
  {
    gpu();
    const int tb0 = traceBlocksPerCU, sb0 = shadeBlocksPerCU;
    const size_t pb0 = pathBudgetBytes;
    const int traceCand[] = {8, 12, 16}, shadeCand[] = {4, 8, 16};
    const size_t budgetCand[] = {(size_t)2048 << 20, (size_t)16384 << 20};
    float bestMs = 1e30f;
    int bt = tb0, bs = sb0;
    size_t bb = pb0;
    std::vector<float> ms(v.size());
    for (size_t pb : budgetCand)
      for (int tb : traceCand)
        for (int sbk : shadeCand) {
          traceBlocksPerCU = tb;
          shadeBlocksPerCU = sbk;
          pathBudgetBytes = pb;
          sizeGrids();
          renderBatch(v, world, 0, ms.data()); // warm-up (allocations)
          float t = 1e30f;
          for (int k = 0; k < std::max(1, reps); k++) {
            renderBatch(v, world, 0, ms.data());
            t = std::min(t, batchMs(ms));
          }
          if (t < bestMs) {
            bestMs = t;
            bt = tb;
            bs = sbk;
            bb = pb;
          }
        }
    traceBlocksPerCU = bt;
    shadeBlocksPerCU = bs;
    pathBudgetBytes = bb;
    sizeGrids();
    // second stage at the chosen grids and budget: the lanes the views
    // spread over (1 when a batch cannot run concurrently, see batchLanes)
    const int ls0 = batchStreams;
    int bl = ls0;
    bestMs = 1e30f;
    for (int ls : {1, 2, 4, 8}) {
      batchStreams = ls;
      renderBatch(v, world, 0, ms.data()); // warm-up (lane allocations)
      float t = 1e30f;
      for (int k = 0; k < std::max(1, reps); k++) {
        renderBatch(v, world, 0, ms.data());
        t = std::min(t, batchMs(ms));
      }
      if (t < bestMs) {
        bestMs = t;
        bl = ls;
      }
    }
    batchStreams = bl;
    // third stage: the pixel tile of the path slots (the shape of the pixel
    // block a wave's primary rays cover)
    const int tw0 = slotTileWidth;
    int bw = tw0;
    bestMs = 1e30f;
    for (int tw : {4, 8, 16, 32}) {
      slotTileWidth = tw;
      renderBatch(v, world, 0, ms.data()); // warm-up (queue sizes)
      float t = 1e30f;
      for (int k = 0; k < std::max(1, reps); k++) {
        renderBatch(v, world, 0, ms.data());
        t = std::min(t, batchMs(ms));
      }
      if (t < bestMs) {
        bestMs = t;
        bw = tw;
      }
    }
    slotTileWidth = bw;
    releaseGraph();
    if (best) {
      best[0] = bt;
      best[1] = bs;
      best[2] = (int)(bb >> 20);
      best[3] = bl;
      best[4] = bw;
    }
    return bestMs;
  }
  void ensureNrc(size_t slots);
  void nrcPass(Renderer *r, World *w);
  int traceRays(World *w, const float *rays8, int n, float *outTUV, uint32_t *outIDs, uint64_t *stats2);
};

FrameBuffer::FrameBuffer(HipDevice *d, int w_, int h_, OSPFrameBufferFormat f, uint32_t ch)
    : dev(d), w(w_), h(h_), format(f), channels(ch)
{
  managedType = OSP_FRAMEBUFFER;
  targetFrames = (ch & OSP_FB_ACCUM) ? 0 : 1;
  const size_t n = (size_t)w * h;
  auto zero = [&](DevBuf &b, size_t bytes) {
    b.ensure(bytes);
    HIP_CHECK(hipMemset(b.ptr, 0, bytes));
  };
  if (ch & OSP_FB_COLOR)
    zero(color, n * 16);
  if (ch & OSP_FB_VARIANCE)
    zero(variance, n * 16);
  if (ch & OSP_FB_DEPTH)
    zero(depth, n * 4);
  if (ch & OSP_FB_NORMAL)
    zero(normal, n * 12);
  if (ch & OSP_FB_ALBEDO)
    zero(albedo, n * 12);
  if (ch & OSP_FB_FIRST_NORMAL)
    zero(firstNormal, n * 12);
  if (ch & OSP_FB_POSITION)
    zero(position, n * 12);
  if (ch & OSP_FB_ID_PRIMITIVE)
    zero(primID, n * 4);
  if (ch & OSP_FB_ID_OBJECT)
    zero(objID, n * 4);
  if (ch & OSP_FB_ID_INSTANCE)
    zero(instID, n * 4);
  const size_t ntask = (size_t)((w + 7) / 8) * ((h + 7) / 8);
  zero(taskVar, ntask * 4);
  zero(accRGBA, n * 16);
  if (ch & (OSP_FB_NORMAL | OSP_FB_ALBEDO | OSP_FB_FIRST_NORMAL | OSP_FB_POSITION))
    zero(accAux, n * 48);
  zero(accZ, n * 4);
  zero(accIDs, n * 12);
  if ((ch & OSP_FB_COLOR) && f != OSP_FB_NONE)
    zero(converted, n * (f == OSP_FB_RGBA32F ? 16 : 4));
  zero(varianceOut, 16);
  zero(cancelLatch, 16);
  HIP_CHECK(hipHostMalloc((void **)&hostVariance, sizeof(float), hipHostMallocDefault));
  *hostVariance = hpt::kInf;
}

FrameBuffer::~FrameBuffer()
{
  if (dev && dev->stream)
    (void)hipStreamSynchronize(dev->stream);
  if (hostVariance)
    (void)hipHostFree(hostVariance);
}

Future::~Future()
{
  if (!empty && finished)
    (void)hipEventSynchronize(finished);
  if (start)
    (void)hipEventDestroy(start);
  if (rendered)
    (void)hipEventDestroy(rendered);
  if (finished)
    (void)hipEventDestroy(finished);
  for (auto e : passDone)
    (void)hipEventDestroy(e);
}

// ---- world build -------------------------------------------------------------------------

static hpt::BuildBox primBox(const Geometry &g, size_t p, int half)
{
  hpt::BuildBox b;
  auto vtx = [&](uint32_t i) { return &g.vertex[3 * (size_t)i]; };
  if (g.geomType == hpt::GEOM_SPHERES) {
    const float *s = &g.spheres[4 * p];
    for (int k = 0; k < 3; k++) {
      b.lo[k] = s[k] - s[3];
      b.hi[k] = s[k] + s[3];
    }
    return b;
  }
  const float *v[3];
  if (g.geomType == hpt::GEOM_TRIANGLES) {
    const uint32_t *ix = &g.index[3 * p];
    v[0] = vtx(ix[0]), v[1] = vtx(ix[1]), v[2] = vtx(ix[2]);
  } else {
    const uint32_t *ix = &g.index[4 * p];
    if (half == 0)
      v[0] = vtx(ix[0]), v[1] = vtx(ix[1]), v[2] = vtx(ix[3]);
    else
      v[0] = vtx(ix[2]), v[1] = vtx(ix[3]), v[2] = vtx(ix[1]);
  }
  for (int k = 0; k < 3; k++) {
    b.lo[k] = std::min(v[0][k], std::min(v[1][k], v[2][k]));
    b.hi[k] = std::max(v[0][k], std::max(v[1][k], v[2][k]));
  }
  return b;
}

// Device build (hip_bvh.h) of the BLAS of `geoms`; null when the list is
// too small for it or the tree needs more traversal stack than the device
// build may use (the host build takes over).
constexpr int kDeviceBlasStackLimit = 96;
static std::shared_ptr<Blas> buildBlasDevice(HipDevice &D, const std::vector<Geometry *> &geoms)
{
  std::vector<HptBvhGeometry> hg;
  for (auto *g : geoms)
    hg.push_back({g->dVertex.as<const float>(), g->dIndex.as<const uint32_t>(), g->dSpheres.as<const float>(),
        g->geomType, (uint32_t)g->numPrims});
  const uint64_t n = hpt_bvh_items(hg.data(), (int)hg.size());
  if (n < 3 || n >= (1ull << 28) || geoms.size() > 4096)
    return nullptr;
  D.gpu();
  const auto t0 = std::chrono::steady_clock::now();
  auto B = std::make_shared<Blas>();
  DevBuf ws, nodes;
  ws.ensure(hpt_bvh_workspace_bytes(n));
  nodes.ensure(n * sizeof(hpt::DNode4));
  B->dPrims.ensure(n * sizeof(hpt::DPrim));
  HptBvhResult r;
  HIP_CHECK(hpt_bvh_build(hg.data(), (int)hg.size(), D.bvhRadius, ws.ptr, ws.bytes, nodes.ptr, B->dPrims.ptr,
      D.bvhHostWords(), &r, D.stream));
  if (r.stackNeed > kDeviceBlasStackLimit)
    return nullptr;
  B->dNodes.ensure((size_t)r.numNodes * sizeof(hpt::DNode4));
  HIP_CHECK(hipMemcpyAsync(B->dNodes.ptr, nodes.ptr, (size_t)r.numNodes * sizeof(hpt::DNode4), hipMemcpyDeviceToDevice,
      D.stream));
  HIP_CHECK(hipStreamSynchronize(D.stream));
  B->onDevice = true;
  B->numNodes = r.numNodes;
  B->numPrims = n;
  for (int k = 0; k < 3; k++) {
    B->bounds.lo[k] = r.lo[k];
    B->bounds.hi[k] = r.hi[k];
  }
  B->maxDepth = r.levels;
  B->stackNeed = r.stackNeed;
  B->buildMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return B;
}

static std::shared_ptr<Blas> buildBlasHost(const std::vector<Geometry *> &geoms)
{
  const auto t0 = std::chrono::steady_clock::now();
  struct Item
  {
    uint32_t model, prim, half;
  };
  std::vector<Item> items;
  std::vector<hpt::BuildBox> boxes;
  for (uint32_t m = 0; m < geoms.size(); m++) {
    const Geometry &g = *geoms[m];
    const int halves = g.geomType == hpt::GEOM_QUADS ? 2 : 1;
    for (size_t p = 0; p < g.numPrims; p++)
      for (int h = 0; h < halves; h++) {
        items.push_back({m, (uint32_t)p, (uint32_t)h});
        boxes.push_back(primBox(g, p, h));
      }
  }
  const int nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
  // leaves of at most 2 primitives: measured fastest on the 1080p scene
  // (leaf size 1, 2, 4, 6, 8: 260, 275, 266, 259, 251 M samples/s)
  hpt::BuildResult R = hpt::buildBVH(boxes, 2, std::min(nthreads, 16));
  auto B = std::make_shared<Blas>();
  B->nodes = std::move(R.nodes);
  B->numNodes = B->nodes.size();
  B->numPrims = R.order.size();
  B->bounds = R.bounds;
  B->maxDepth = R.maxDepth;
  B->stackNeed = R.stackNeed;
  B->prims.resize(R.order.size());
  for (size_t k = 0; k < R.order.size(); k++) {
    const Item &it = items[R.order[k]];
    const Geometry &g = *geoms[it.model];
    hpt::DPrim &P = B->prims[k];
    std::memset(&P, 0, sizeof(P));
    P.primID = it.prim;
    if (g.geomType == hpt::GEOM_SPHERES) {
      const float *s = &g.spheres[4 * (size_t)it.prim];
      P.a[0] = s[0], P.a[1] = s[1], P.a[2] = s[2];
      P.b[0] = s[3];
      P.modelKind = it.model | ((uint32_t)hpt::PRIM_SPHERE << 28);
      continue;
    }
    uint32_t vi[3];
    uint32_t kind;
    if (g.geomType == hpt::GEOM_TRIANGLES) {
      const uint32_t *ix = &g.index[3 * (size_t)it.prim];
      vi[0] = ix[0], vi[1] = ix[1], vi[2] = ix[2];
      kind = hpt::PRIM_TRIANGLE;
    } else {
      const uint32_t *ix = &g.index[4 * (size_t)it.prim];
      if (it.half == 0)
        vi[0] = ix[0], vi[1] = ix[1], vi[2] = ix[3], kind = hpt::PRIM_QUAD0;
      else
        vi[0] = ix[2], vi[1] = ix[3], vi[2] = ix[1], kind = hpt::PRIM_QUAD1;
    }
    for (int k = 0; k < 3; k++) {
      P.a[k] = g.vertex[3 * (size_t)vi[0] + k];
      P.b[k] = g.vertex[3 * (size_t)vi[1] + k];
      P.c[k] = g.vertex[3 * (size_t)vi[2] + k];
    }
    P.modelKind = it.model | (kind << 28);
  }
  if (geoms.size() >= (1u << 28))
    throw std::runtime_error("too many geometric models in one group");
  B->buildMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return B;
}

void World::commit() // World.cpp:70-188 plus the Group BLAS builds (Group.cpp:93-143)
{
  HipDevice &D = *dev;
  instances = dataObjects<Instance>(getData(this, "instance"), OSP_INSTANCE);
  std::vector<Light *> lights = dataObjects<Light>(getData(this, "light"), OSP_LIGHT);
  instRefs.clear();
  lightRefs.clear();
  for (auto *i : instances) {
    if (!i || !i->group)
      throw std::runtime_error("world 'instance' holds an uncommitted or null instance");
    instRefs.emplace_back(new Ref<Instance>());
    instRefs.back()->reset(i);
  }
  for (auto *l : lights) {
    if (!l)
      throw std::runtime_error("world 'light' holds a null light");
    lightRefs.emplace_back(new Ref<Light>());
    lightRefs.back()->reset(l);
  }
  if (getData(this, "volume"))
    throw std::runtime_error("volumes are not supported by the hip device");

  // BLAS per distinct geometry list, cached across commits
  std::vector<Group *> groups;
  std::map<Group *, int> groupIndex;
  for (auto *i : instances)
    if (!groupIndex.count(i->group.p)) {
      groupIndex[i->group.p] = (int)groups.size();
      groups.push_back(i->group.p);
    }
  std::vector<std::shared_ptr<Blas>> groupBlas(groups.size());
  std::map<const Blas *, int> blasSlot;
  std::vector<std::shared_ptr<Blas>> uniqueBlas;
  for (size_t gi = 0; gi < groups.size(); gi++) {
    std::vector<Geometry *> geoms;
    std::vector<std::pair<const Geometry *, uint64_t>> key;
    for (auto *m : groups[gi]->models) {
      if (!m->geom)
        throw std::runtime_error("group holds an uncommitted geometric model");
      geoms.push_back(m->geom.p);
      key.push_back({m->geom.p, m->geom->version});
    }
    auto it = D.blasCache.find(key);
    if (it == D.blasCache.end()) {
      std::shared_ptr<Blas> b = D.hostBvh ? nullptr : buildBlasDevice(D, geoms);
      it = D.blasCache.emplace(key, b ? b : buildBlasHost(geoms)).first;
    }
    groupBlas[gi] = it->second;
    if (!blasSlot.count(it->second.get())) {
      blasSlot[it->second.get()] = (int)uniqueBlas.size();
      uniqueBlas.push_back(it->second);
    }
  }
  // drop cache entries no longer referenced by this world
  for (auto it = D.blasCache.begin(); it != D.blasCache.end();) {
    if (!blasSlot.count(it->second.get()))
      it = D.blasCache.erase(it);
    else
      ++it;
  }

  // materials, media and models
  materials.clear();
  std::map<Material *, int> matIndex;
  std::vector<hpt::DMaterial> dmats;
  std::vector<hpt::DMedium> media(1);
  media[0] = {{0.f, 0.f, 0.f}, 1.0f}; // vacuum (Medium.ih make_Medium_vacuum)
  auto materialOf = [&](Material *m) -> int {
    if (!m)
      return -1;
    auto it = matIndex.find(m);
    if (it != matIndex.end())
      return it->second;
    hpt::DMaterial dm = m->m;
    if (dm.type == hpt::MAT_GLASS) {
      dm.inside = (int)media.size();
      media.push_back(m->inside);
      dm.outside = (int)media.size();
      media.push_back(m->outside);
    }
    const int idx = (int)dmats.size();
    dmats.push_back(dm);
    materials.push_back(m);
    matIndex[m] = idx;
    return idx;
  };
  std::vector<hpt::DModel> models;
  std::vector<int> groupModelOffset(groups.size());
  for (size_t gi = 0; gi < groups.size(); gi++) {
    groupModelOffset[gi] = (int)models.size();
    for (auto *m : groups[gi]->models) {
      const Geometry &g = *m->geom.p;
      hpt::DModel dm;
      std::memset(&dm, 0, sizeof(dm));
      dm.vertex = g.dVertex.as<const float>();
      dm.index = g.dIndex.as<const uint32_t>();
      dm.normal = g.normal.empty() ? nullptr : g.dNormal.as<const float>();
      dm.spheres = g.dSpheres.as<const float>();
      dm.geomType = g.geomType;
      dm.material = materialOf(m->material.p);
      dm.userID = m->userID;
      dm.invertNormals = m->invertNormals ? 1 : 0;
      dm.opaque = m->material ? (m->material->opaque() ? 1 : 0) : 1;
      models.push_back(dm);
    }
  }
  if (media.size() >= (1u << 12))
    throw std::runtime_error("too many glass materials in one world");

  // concatenate BLAS nodes and primitives
  std::vector<int> blasNodeOff(uniqueBlas.size()), blasPrimOff(uniqueBlas.size());
  size_t nNodes = 0, nPrims = 0;
  maxDepth = 0;
  int blasStackNeed = 0;
  for (size_t k = 0; k < uniqueBlas.size(); k++) {
    blasNodeOff[k] = (int)nNodes;
    blasPrimOff[k] = (int)nPrims;
    nNodes += uniqueBlas[k]->numNodes;
    nPrims += uniqueBlas[k]->numPrims;
    maxDepth = std::max(maxDepth, uniqueBlas[k]->maxDepth);
    blasStackNeed = std::max(blasStackNeed, uniqueBlas[k]->stackNeed);
  }
  // instances and TLAS
  std::vector<hpt::DInstance> dinst(instances.size());
  std::vector<hpt::BuildBox> instBoxes(instances.size());
  for (size_t i = 0; i < instances.size(); i++) {
    const Instance &I = *instances[i];
    const int gi = groupIndex[I.group.p];
    const Blas &B = *groupBlas[gi];
    hpt::DInstance &di = dinst[i];
    std::memset(&di, 0, sizeof(di));
    const float *x = I.xfm;
    const hpt::affine3f a = {{{x[0], x[1], x[2]}, {x[3], x[4], x[5]}, {x[6], x[7], x[8]}}, {x[9], x[10], x[11]}};
    const hpt::affine3f r = hpt::inverse(a);
    const float rv[12] = {r.l.vx.x, r.l.vx.y, r.l.vx.z, r.l.vy.x, r.l.vy.y, r.l.vy.z, r.l.vz.x, r.l.vz.y, r.l.vz.z, r.p.x, r.p.y, r.p.z};
    std::memcpy(di.rcp, rv, sizeof(rv));
    std::memcpy(di.xfm, x, 48);
    di.blasRoot = blasNodeOff[blasSlot[&B]];
    di.modelOffset = groupModelOffset[gi];
    di.userID = I.userID;
    di.epsScale = hpt::maxf(std::fabs(x[0]), hpt::maxf(std::fabs(x[4]), std::fabs(x[8])));
    hpt::BuildBox wb;
    for (int k = 0; k < 3; k++) {
      wb.lo[k] = hpt::kInf;
      wb.hi[k] = -hpt::kInf;
    }
    if (B.bounds.lo[0] <= B.bounds.hi[0]) {
      for (int c = 0; c < 8; c++) {
        const vec3f p = v3(c & 1 ? B.bounds.hi[0] : B.bounds.lo[0], c & 2 ? B.bounds.hi[1] : B.bounds.lo[1],
            c & 4 ? B.bounds.hi[2] : B.bounds.lo[2]);
        const vec3f q = hpt::mul(a.l, p) + a.p;
        const float qq[3] = {q.x, q.y, q.z};
        for (int k = 0; k < 3; k++) {
          wb.lo[k] = std::min(wb.lo[k], qq[k]);
          wb.hi[k] = std::max(wb.hi[k], qq[k]);
        }
      }
      for (int k = 0; k < 3; k++) {
        const float pad = (std::fabs(wb.lo[k]) + std::fabs(wb.hi[k])) * 1e-6f;
        wb.lo[k] -= pad;
        wb.hi[k] += pad;
      }
    }
    instBoxes[i] = wb;
  }
  std::vector<uint32_t> tlasInst; // instances with geometry
  std::vector<hpt::BuildBox> tlasBoxes;
  for (size_t i = 0; i < instBoxes.size(); i++)
    if (instBoxes[i].lo[0] <= instBoxes[i].hi[0]) {
      tlasInst.push_back((uint32_t)i);
      tlasBoxes.push_back(instBoxes[i]);
    }
  hpt::BuildResult T = hpt::buildBVH(tlasBoxes, 1, 1);
  std::vector<uint32_t> tlasOrder(T.order.size());
  for (size_t k = 0; k < T.order.size(); k++)
    tlasOrder[k] = tlasInst[T.order[k]];
  const int tlasOff = (int)nNodes;
  auto remap = [](hpt::DNode4 n, int nodeOff, int primOff, const std::vector<uint32_t> *order) {
    for (int c = 0; c < 4; c++) {
      // an unused slot (box at +inf) is never entered: leave it alone
      if (n.lox[c] == hpt::kInf)
        continue;
      int32_t ch = n.child[c];
      if (ch >= 0) {
        n.child[c] = ch + nodeOff;
      } else {
        const uint32_t enc = ~(uint32_t)ch;
        uint32_t first = enc >> 3;
        const uint32_t cnt = enc & 7u;
        first = order ? (*order)[first] : first + (uint32_t)primOff;
        n.child[c] = (int32_t) ~((first << 3) | cnt);
      }
    }
    return n;
  };
  // the world's node array: every BLAS (relocated on the device for device
  // builds, on the host otherwise), then the TLAS; primitive records of
  // every BLAS in the same order
  const size_t totalNodeCount = nNodes + T.nodes.size();
  D.gpu();
  HIP_CHECK(hipStreamSynchronize(D.stream)); // frames in flight read the old arrays
  dNodes.ensure(std::max<size_t>(sizeof(hpt::DNode4) * totalNodeCount, 16));
  dPrims.ensure(std::max<size_t>(sizeof(hpt::DPrim) * nPrims, 16));
  hpt::DNode4 *gNodes = dNodes.as<hpt::DNode4>();
  hpt::DPrim *gPrims = dPrims.as<hpt::DPrim>();
  for (size_t k = 0; k < uniqueBlas.size(); k++) {
    const Blas &B = *uniqueBlas[k];
    if (B.onDevice) {
      HIP_CHECK(hpt_bvh_relocate(gNodes + blasNodeOff[k], B.dNodes.ptr, B.numNodes, blasNodeOff[k],
          (uint32_t)blasPrimOff[k], D.stream));
      HIP_CHECK(hipMemcpyAsync(gPrims + blasPrimOff[k], B.dPrims.ptr, sizeof(hpt::DPrim) * B.numPrims,
          hipMemcpyDeviceToDevice, D.stream));
    } else {
      std::vector<hpt::DNode4> moved(B.nodes.size());
      for (size_t i = 0; i < B.nodes.size(); i++)
        moved[i] = remap(B.nodes[i], blasNodeOff[k], blasPrimOff[k], nullptr);
      HIP_CHECK(hipMemcpy(gNodes + blasNodeOff[k], moved.data(), sizeof(hpt::DNode4) * moved.size(),
          hipMemcpyHostToDevice));
      if (!B.prims.empty())
        HIP_CHECK(hipMemcpy(gPrims + blasPrimOff[k], B.prims.data(), sizeof(hpt::DPrim) * B.prims.size(),
            hipMemcpyHostToDevice));
    }
  }
  {
    std::vector<hpt::DNode4> tl(T.nodes.size());
    for (size_t i = 0; i < T.nodes.size(); i++)
      tl[i] = remap(T.nodes[i], tlasOff, 0, &tlasOrder);
    HIP_CHECK(hipMemcpy(gNodes + tlasOff, tl.data(), sizeof(hpt::DNode4) * tl.size(), hipMemcpyHostToDevice));
  }
  HIP_CHECK(hipStreamSynchronize(D.stream));

  // lights (PathTracerData.cpp:32-82): [geometry lights, world lights],
  // uniform selection CDFs with and without the geometry lights.  Geometry
  // lights (PathTracerData.cpp:121-162): every emissive model of every
  // instance, primitives chosen by world-space area; the model's areaPDF is
  // the one of the last instance that gathers it (GeometricModel.ispc:14-55)
  std::vector<hpt::DLight> dl;
  std::vector<float> geoCDF;
  std::map<const GeometricModel *, float> modelAreaPDF;
  for (size_t i = 0; i < instances.size(); i++) {
    const int gi = groupIndex[instances[i]->group.p];
    const float *x = instances[i]->xfm;
    const hpt::linear3f xl = {{x[0], x[1], x[2]}, {x[3], x[4], x[5]}, {x[6], x[7], x[8]}};
    const std::vector<GeometricModel *> &gm = groups[gi]->models;
    for (size_t k = 0; k < gm.size(); k++) {
      const GeometricModel *m = gm[k];
      const Geometry &g = *m->geom.p;
      if (!m->material || !m->material->emissive() || g.numPrims == 0)
        continue;
      hpt::DLight L;
      std::memset(&L, 0, sizeof(L));
      L.type = hpt::LIGHT_GEOMETRY;
      L.visible = 1;
      L.geoInst = (int32_t)i;
      L.geoModel = groupModelOffset[gi] + (int32_t)k;
      L.geoNumPrims = (int32_t)g.numPrims;
      L.geoCdfOffset = (int32_t)geoCDF.size();
      const size_t off = geoCDF.size();
      geoCDF.resize(off + g.numPrims);
      for (size_t p = 0; p < g.numPrims; p++)
        geoCDF[off + p] = hpt::primArea(g.vertex.data(), g.index.data(), g.spheres.data(), g.geomType, (uint32_t)p, xl);
      L.geoPdf = 1.f / hpt::distribution1DCreate((int)g.numPrims, geoCDF.data() + off);
      modelAreaPDF[m] = L.geoPdf;
      dl.push_back(L);
    }
  }
  scene.numGeoLights = (int)dl.size();
  for (size_t gi = 0; gi < groups.size(); gi++)
    for (size_t k = 0; k < groups[gi]->models.size(); k++) {
      auto it = modelAreaPDF.find(groups[gi]->models[k]);
      if (it != modelAreaPDF.end())
        models[groupModelOffset[gi] + k].areaPDF = it->second;
    }
  for (auto *l : lights)
    dl.push_back(l->l);
  std::vector<float> cdf(dl.size(), 1.f);
  if (!cdf.empty())
    hpt::distribution1DCreate((int)cdf.size(), cdf.data());
  std::vector<float> worldCdf(lights.size(), 1.f);
  if (!worldCdf.empty())
    hpt::distribution1DCreate((int)worldCdf.size(), worldCdf.data());

  // volumes: one DVolume per volumetric model, and every (instance, model)
  // pair in instance order, then group order
  std::vector<hpt::DVolume> dvols;
  std::vector<hpt::DVolInst> volInsts;
  std::map<const VolumetricModel *, int> volIndex;
  for (size_t i = 0; i < instances.size(); i++) {
    const std::vector<VolumetricModel *> &gv = groups[groupIndex[instances[i]->group.p]]->volumes;
    for (size_t k = 0; k < gv.size(); k++) {
      auto it = volIndex.find(gv[k]);
      if (it == volIndex.end()) {
        it = volIndex.emplace(gv[k], (int)dvols.size()).first;
        dvols.push_back(gv[k]->device());
      }
      volInsts.push_back({(int32_t)i, it->second, (int32_t)k, 0});
    }
  }
  if (instances.size() >= hpt::kVolumeHit || volInsts.size() >= hpt::kVolumeHit)
    throw std::runtime_error("too many instances for the hip device");

  dInstances.upload(dinst);
  dModels.upload(models);
  dMaterials.upload(dmats);
  dMedia.upload(media);
  dLights.upload(dl);
  dLightsCDF.upload(cdf);
  dWorldLightsCDF.upload(worldCdf);
  dGeoCDF.upload(geoCDF);
  scene.nodes = dNodes.as<const hpt::DNode4>();
  scene.prims = dPrims.as<const hpt::DPrim>();
  scene.instances = dInstances.as<const hpt::DInstance>();
  scene.models = dModels.as<const hpt::DModel>();
  scene.materials = dMaterials.as<const hpt::DMaterial>();
  scene.media = dMedia.as<const hpt::DMedium>();
  scene.lights = dLights.as<const hpt::DLight>();
  scene.lightsCDF = dLightsCDF.as<const float>();
  scene.geoCDF = dGeoCDF.as<const float>();
  scene.tlasRoot = tlasOff;
  scene.numInstances = (int)instances.size();
  scene.numLights = (int)dl.size();
  scene.numMaterials = (int)dmats.size();
  dVolumes.upload(dvols);
  dVolInsts.upload(volInsts);
  scene.volumes = dVolumes.as<const hpt::DVolume>();
  scene.volInsts = dVolInsts.as<const hpt::DVolInst>();
  scene.numVolInsts = (int32_t)volInsts.size();
  totalNodes = (int64_t)totalNodeCount;
  totalPrims = (int64_t)nPrims;
  blasBuildMs = 0.0;
  deviceBlas = 0;
  for (auto &b : uniqueBlas) {
    blasBuildMs += b->buildMs;
    deviceBlas += b->onDevice ? 1 : 0;
  }
  for (int k = 0; k < 3; k++) {
    bounds[k] = T.bounds.lo[k];
    bounds[3 + k] = T.bounds.hi[k];
  }
  // the traversal stack holds the TLAS entries of one root-to-leaf path,
  // the next TLAS node (a held instance leaf is entered after the lane went
  // on in the TLAS), the instance marker and the BLAS entries
  if (T.stackNeed + 2 + blasStackNeed > hpt_stack_capacity())
    throw std::runtime_error("BVH too deep for the traversal stack");
  stackNeed = T.stackNeed + 2 + blasStackNeed;
  maxDepth = std::max(maxDepth, T.maxDepth);
  built = true;
}

// ---- frame orchestration -------------------------------------------------------------------

void HipDevice::ensurePaths(size_t slots, int maxRec, bool aux)
{
  PathLane &L = *lane;
  size_t &pathSlots = L.pathSlots, &recSlots = L.recSlots;
  bool &auxAllocated = L.auxAllocated;
  hpt::DPaths &paths = L.paths;
  maxRec = std::max(1, maxRec);
  const bool grow = slots > pathSlots || (size_t)maxRec * slots > recSlots || (aux && !auxAllocated);
  if (!grow)
    return;
  if (capturing)
    throw std::runtime_error("batched views: path buffers would grow while capturing");
  releaseGraph(); // a captured batch refers to the old buffers
  HIP_CHECK(hipStreamSynchronize(L.stream));
  pathSlots = std::max(slots, pathSlots);
  recSlots = std::max((size_t)maxRec * pathSlots, recSlots);
  auxAllocated = auxAllocated || aux;
  const size_t S = pathSlots, R = recSlots;
  // per-slot arrays: 9 x float4 + hitPrim + ids(uint4) + neeN + volume
  // sampler (uint4) + NEE vertex record (6 x float4) (+ aux 16 floats)
  const size_t perSlot = 9 * 16 + 4 + 16 + 4 + 16 + 6 * 16 + (auxAllocated ? 64 : 0);
  L.pathBuf.ensure(S * perSlot);
  L.recBuf.ensure(R * 80);
  L.queueBuf.ensure((3 * S + 3 * R) * 4);
  // recorded transparent hits of up to 4M transparency-queue records (the
  // rest are traversed again); 4 words per hit + a count
  const size_t transCap = std::min<size_t>(R, (size_t)1 << 22);
  const size_t transHitBytes = transCap * 16 * hpt_trans_record_hits();
  L.transBuf.ensure(transHitBytes + transCap * 4);
  char *b = (char *)L.pathBuf.ptr;
  auto take = [&](size_t bytes) {
    char *p = b;
    b += bytes;
    return p;
  };
  paths.org = (float *)take(S * 16);
  paths.dir = (float *)take(S * 16);
  paths.thr = (float *)take(S * 16);
  paths.contrib = (float *)take(S * 16);
  paths.ld = (uint32_t *)take(S * 16);
  paths.lastP = (float *)take(S * 16);
  paths.lastNs = (float *)take(S * 16);
  paths.hitA = (float *)take(S * 16);
  paths.hitB = (float *)take(S * 16);
  paths.ids = (uint32_t *)take(S * 16);
  paths.hitPrim = (uint32_t *)take(S * 4);
  paths.neeN = (uint32_t *)take(S * 4);
  paths.rng = (uint32_t *)take(S * 16);
  paths.vtxP = (float *)take(S * 16);
  paths.vtxNg = (float *)take(S * 16);
  paths.vtxNs = (float *)take(S * 16);
  paths.vtxThr = (float *)take(S * 16);
  paths.vtxWo = (float *)take(S * 16);
  paths.vtxLd = (uint32_t *)take(S * 16);
  L.auxBase = auxAllocated ? (float *)take(S * 64) : nullptr;
  paths.aux = L.auxBase;
  char *r = (char *)L.recBuf.ptr;
  paths.recOrg = (float *)r;
  paths.recDir = (float *)(r + R * 16);
  paths.recContrib = (float *)(r + R * 32);
  paths.recResult = (float *)(r + R * 48);
  paths.recVol = (float *)(r + R * 64);
  uint32_t *q = (uint32_t *)L.queueBuf.ptr;
  paths.queue[0] = q;
  paths.queue[1] = q + S;
  paths.shadowQueue = q + 2 * S;
  paths.transQueue = q + 2 * S + R;
  paths.neeSampleQueue = q + 2 * S + 2 * R;
  paths.tieQueue = q + 2 * S + 3 * R;
  paths.transHits = (float *)L.transBuf.ptr;
  paths.transCount = (uint32_t *)((char *)L.transBuf.ptr + transHitBytes);
  paths.transCap = (int32_t)transCap;
  paths.counters = L.counters.as<int32_t>();
  paths.overflow = L.overflow.as<uint32_t>();
  paths.stats = (unsigned long long *)stats.ptr;
}

void HipDevice::ensureNrc(size_t slots)
{
  if (slots > nrcSlots) {
    releaseGraph();
    HIP_CHECK(hipStreamSynchronize(lane->stream));
    nrcSlots = slots;
    // 16 floats query record + 16 floats training record + 4 floats
    // contribution before the training vertex
    nrcRec.ensure(slots * 144);
    nrcQueues.ensure(slots * 8);
    nrcRaw.ensure(slots * 48);
    nrcFeat.ensure(slots * HPT_NRC_INPUTS * 2);
    nrcOut.ensure(slots * 12);
    nrcTarget.ensure(slots * 12);
  }
  hpt::DPaths &paths = lane->paths;
  paths.nrcRec = nrcRec.as<float>();
  paths.nrcTrainRec = nrcRec.as<float>() + 16 * nrcSlots; // aliased to nrcRec without self-training
  paths.nrcBefore = nrcRec.as<float>() + 32 * nrcSlots;
  paths.nrcQueryQueue = nrcQueues.as<uint32_t>();
  paths.nrcTrainQueue = nrcQueues.as<uint32_t>() + nrcSlots;
}

// The radiance cache after a pass's bounces: predictions for the paths that
// ended at the cache, then Adam steps on the training records' targets.
// The record counts are read back (one synchronisation per pass).
void HipDevice::nrcPass(Renderer *r, World *w)
{
  const hpt::DPaths &paths = lane->paths;
  const hipStream_t stream = lane->stream;
  const HptPaths *P = reinterpret_cast<const HptPaths *>(&paths);
  int32_t cnt[2] = {0, 0};
  HIP_CHECK(hipMemcpyAsync(cnt, paths.counters + hpt::kCntNrcQuery, sizeof(cnt), hipMemcpyDeviceToHost, stream));
  HIP_CHECK(hipStreamSynchronize(stream));
  const int nQ = cnt[0], nT = cnt[1];
  r->nrcQueries += nQ;
  r->nrcRecords += nT;
  // scene box for the position encoding, padded so it is never empty
  float lo[3], hi[3];
  for (int k = 0; k < 3; k++) {
    const float pad = 1e-3f * (1.f + std::fabs(w->bounds[k]) + std::fabs(w->bounds[3 + k]));
    lo[k] = w->bounds[k] - pad;
    hi[k] = w->bounds[3 + k] + pad;
  }
  uint16_t *feat = nrcFeat.as<uint16_t>();
  // a pass yields at most about 2 x nrcTrainPaths training records; the
  // per-tile gradient rows of one Adam step are sized for that
  const int maxTiles = (2 * r->nrcTrainPaths + 63) / 64 + 1;
  if (!r->nrcInit || maxTiles > r->nrcNet.maxTiles) {
    const size_t P4 = (size_t)HPT_NRC_PARAMS * 4;
    r->nW.ensure(HPT_NRC_PARAMS * 2);
    r->nWT.ensure(HPT_NRC_TPARAMS * 2);
    r->nMaster.ensure(P4);
    r->nM.ensure(P4);
    r->nV.ensure(P4);
    r->nPartial.ensure((size_t)maxTiles * P4);
    r->nTileLoss.ensure((size_t)maxTiles * 4);
    r->nLoss.ensure(16);
    r->nrcNet = {r->nW.as<uint16_t>(), r->nWT.as<uint16_t>(), r->nMaster.as<float>(), r->nM.as<float>(),
        r->nV.as<float>(), r->nPartial.as<float>(), r->nTileLoss.as<float>(), r->nLoss.as<float>(), maxTiles};
  }
  if (!r->nrcInit) {
    launch("nrc_init", [&]() { return hpt_nrc_init(&r->nrcNet, r->nrcSeed, stream); });
    r->nrcInit = true;
    r->nrcStep = 0;
  }
  if (nQ > 0) // gather, encode, infer and apply in one kernel
    launch("nrc_infer", [&]() {
      return hpt_nrc_query(&r->nrcNet, paths.nrcRec, paths.nrcQueryQueue, nQ, lo, hi, paths.contrib, stream);
    });
  if (nT > 0) {
    int steps = std::min(r->nrcTrainSteps, (nT + 63) / 64);
    int batch = (nT + steps - 1) / steps;
    if ((batch + 63) / 64 > r->nrcNet.maxTiles) { // more, smaller steps
      batch = r->nrcNet.maxTiles * 64;
      steps = (nT + batch - 1) / batch;
    }
    launch("nrc_train", [&]() {
      hipError_t e = hpt_nrc_targets(P, nT, nrcTarget.as<float>(), stream);
      if (e == hipSuccess)
        e = hpt_nrc_gather(P, 1, nT, nrcRaw.as<float>(), stream);
      if (e == hipSuccess)
        e = hpt_nrc_encode(nrcRaw.as<float>(), nT, lo, hi, feat, stream);
      for (int k = 0; k < steps && e == hipSuccess; k++) {
        const int b0 = k * batch, nb = std::min(batch, nT - b0);
        e = hpt_nrc_train(&r->nrcNet, feat + (size_t)b0 * HPT_NRC_INPUTS, nrcTarget.as<float>() + 3 * (size_t)b0, nb,
            stream);
        if (e == hipSuccess)
          e = hpt_nrc_adam(&r->nrcNet, nb, r->nrcLr, ++r->nrcStep, stream);
      }
      return e;
    });
  }
}

ManagedObject *HipDevice::renderFrame(ManagedObject *fbo, ManagedObject *ro, ManagedObject *co, ManagedObject *wo)
{
  if (fbo->managedType != OSP_FRAMEBUFFER || ro->managedType != OSP_RENDERER || co->managedType != OSP_CAMERA
      || wo->managedType != OSP_WORLD)
    throw std::runtime_error("ospRenderFrame: wrong object types");
  FrameBuffer *fb = (FrameBuffer *)fbo;
  Renderer *r = (Renderer *)ro;
  Camera *cam = (Camera *)co;
  World *w = (World *)wo;
  if (!w->built)
    throw std::runtime_error("ospRenderFrame: the world has not been committed");
  gpu();
  if (r->commitCounter == 0 || cam->commitCounter == 0)
    throw std::runtime_error("ospRenderFrame: renderer and camera must be committed");

  // this frame's lane: its stream and path state
  const hipStream_t stream = lane->stream;
  hpt::DPaths &paths = lane->paths;
  auto *fut = new Future(this);
  fut->fb.reset(fbo);
  fut->renderer.reset(ro);
  fut->camera.reset(co);
  fut->world.reset(wo);
  HIP_CHECK(hipEventCreate(&fut->start));
  HIP_CHECK(hipEventCreate(&fut->rendered));
  HIP_CHECK(hipEventCreate(&fut->finished));
  if (capturing) {
    fut->empty = true; // a captured frame: its completion is the graph's
  } else {
    HIP_CHECK(hipEventRecord(fut->start, stream));
    if (++frameGen == 0)
      ++frameGen;
    fut->gen = frameGen;
  }

  // FrameBuffer::beginFrame (FrameBuffer.cpp:115-137)
  const bool doAccum = (fb->channels & OSP_FB_ACCUM) != 0;
  beginFrameHost(fb);
  fb->errorThreshold = r->errorThreshold;
  const bool finishedAccum = fb->targetFrames && fb->frameID >= fb->targetFrames;

  hpt::DFrame F;
  std::memset(&F, 0, sizeof(F));
  F.width = fb->w;
  F.height = fb->h;
  F.tilesX = (fb->w + 7) / 8;
  F.tilesY = (fb->h + 7) / 8;
  // the path slots of a sample: 64-slot tiles of slotTileWidth x (64 /
  // slotTileWidth) pixels, so a wave's primary rays cover one tile
  F.slotShift = slotTileWidth == 4 ? 2 : (slotTileWidth == 16 ? 4 : (slotTileWidth == 32 ? 5 : 3));
  F.slotTilesX = (fb->w + (1 << F.slotShift) - 1) >> F.slotShift;
  F.numPixPadded = F.slotTilesX * ((fb->h + (64 >> F.slotShift) - 1) / (64 >> F.slotShift)) * 64;
  F.spp = r->spp;
  F.accumID = (uint32_t)fb->frameID;
  F.targetFrames = fb->targetFrames;
  F.projectedDepth = fb->projectedDepth ? 1 : 0;
  F.channels = fb->channels;
  F.accumulateVariance = fb->accumulateVariance ? 1 : 0;
  cam->setup(F);
  F.maxDepth = r->maxDepth;
  F.rouletteDepth = r->rouletteDepth;
  F.maxScatteringEvents = r->maxScatteringEvents;
  // the world's scene as this renderer sees it: without the geometry lights
  // when "geometryLights" is off (PathTracer.cpp:49, PathTracerData.cpp:37-43)
  hpt::DScene sceneF = w->scene;
  if (!r->geometryLights && sceneF.numGeoLights > 0) {
    sceneF.lights += sceneF.numGeoLights;
    sceneF.numLights -= sceneF.numGeoLights;
    sceneF.numGeoLights = 0;
    sceneF.lightsCDF = w->dWorldLightsCDF.as<const float>();
  }
  // light sample counts (PathTracer.ispc:95-110)
  const uint32_t numLights = (uint32_t)sceneF.numLights;
  F.numFirstLightSamples = r->numLightSamples >= 0 && numLights > 0 ? (uint32_t)r->numLightSamples
                                                                   : std::min<uint32_t>(1024u, numLights);
  F.numIndirectLightSamples = r->limitIndirect && F.numFirstLightSamples >= 1 ? 1u : F.numFirstLightSamples;
  F.minContribution = r->minContribution;
  F.maxRadiance = r->maxRadiance;
  std::memcpy(F.bgColor, r->bg, 16);
  F.filterType = r->filterType;
  F.filterWidth = r->filterWidth;
  F.lutSize = r->lutSize;
  F.lutCdfX = r->dLutX.as<const float>();
  F.lutCdfY = r->dLutY.as<const float>();
  // LDSampler bits (PathTracer.ispc:128-147)
  uint32_t bits = 16, bitsLight = 16;
  if (F.targetFrames) {
    const uint32_t targetSpp = (uint32_t)F.targetFrames * (uint32_t)F.spp;
    auto clz = [](uint32_t x) { return x ? (uint32_t)__builtin_clz(x) : 32u; };
    bits = 16 + 32 - clz(targetSpp - 1);
    bitsLight = 48 - clz(targetSpp * F.numFirstLightSamples - 1);
  }
  F.idxMask = bits > 31 ? 0xffffffffu : ~(0xffffffffu >> bits);
  F.idxMaskLight = bitsLight > 31 ? 0xffffffffu : ~(0xffffffffu >> bitsLight);
  F.accRGBA = fb->accRGBA.as<float>();
  F.accAux = fb->accAux.as<float>();
  F.accZ = fb->accZ.as<float>();
  F.accIDs = fb->accIDs.as<uint32_t>();
  F.fbColor = fb->color.as<float>();
  F.fbVariance = fb->variance.as<float>();
  F.fbDepth = fb->depth.as<float>();
  F.fbNormal = fb->normal.as<float>();
  F.fbAlbedo = fb->albedo.as<float>();
  F.fbFirstNormal = fb->firstNormal.as<float>();
  F.fbPosition = fb->position.as<float>();
  F.fbPrimID = fb->primID.as<uint32_t>();
  F.fbObjID = fb->objID.as<uint32_t>();
  F.fbInstID = fb->instID.as<uint32_t>();
  F.taskVariance = fb->taskVar.as<float>();
  // LocalFrameBuffer::getRenderTaskIDs (fb/LocalFB.cpp:238-261) with
  // minimumAdaptiveFrames = max(2, 16 / spp) (fb/FrameBuffer.h:183-186)
  if (r->errorThreshold > 0.f && fb->variance.ptr && fb->frameID >= std::max(2, 16 / std::max(1, r->spp)))
    F.adaptiveThreshold = r->errorThreshold;
  // a captured frame on an accumulating framebuffer reads its index from
  // the device, so each launch of the graph renders the next frame
  if (capturing && doAccum) {
    if (fb->targetFrames > 0)
      throw std::runtime_error("batched views: a captured batch cannot stop at targetFrames");
    if (batchView < 0 || (size_t)batchView >= batchStateViews)
      throw std::runtime_error("batched views: no frame state for this view");
    F.frameState = batchState.as<const uint32_t>() + batchView;
    F.varianceCoins = varianceCoins.as<const uint32_t>();
    F.errorThreshold = fb->variance.ptr ? r->errorThreshold : 0.f;
    F.minAdaptiveFrames = std::max(2, 16 / std::max(1, r->spp));
  }
  F.maxRec = (int)std::max<uint32_t>(1u, F.numFirstLightSamples);
  F.hasVolumes = sceneF.numVolInsts > 0 ? 1 : 0;
  F.stats = travStats ? 1 : 0;
  F.nrcBounce = r->nrc && r->nrcBounce <= F.maxDepth ? r->nrcBounce : -1;
  if (capturing && F.nrcBounce >= 0)
    throw std::runtime_error("batched views: the radiance cache needs a host synchronisation per pass");
  r->nrcQueries = r->nrcRecords = 0;

  if (!finishedAccum) {
    const bool aux = F.accAux != nullptr;
    const size_t perSlotBytes = 9 * 16 + 56 + 96 + (aux ? 64 : 0) + (size_t)F.maxRec * 92 + 8
        + (F.nrcBounce >= 0 ? 80 + 8 + 48 + 2 * HPT_NRC_INPUTS + 24 : 0);
    const size_t budgetSlots = std::max<size_t>(F.numPixPadded, pathBudgetBytes / perSlotBytes);
    int sppPerPass = (int)std::max<size_t>(1, std::min<size_t>((size_t)F.spp, budgetSlots / F.numPixPadded));
    // Pass pipelining: a frame of several samples per pixel (outside a
    // batch, without the radiance cache) runs its passes alternately on
    // frameLanes lanes, so that one pass's sparse late bounces overlap the
    // next pass's dense early ones.  Each pass's accumulation waits for the
    // previous pass's: the samples are summed in index order as in one pass.
    int nl = 1;
    if (!inBatch && F.nrcBounce < 0 && frameLanes > 1 && F.spp > 1) {
      nl = std::min(frameLanes, F.spp);
      sppPerPass = std::min(sppPerPass, (F.spp + nl - 1) / nl);
      nl = std::min(nl, (F.spp + sppPerPass - 1) / sppPerPass);
    }
    PathLane *const frameLane = lane;
    for (int k = nl - 1; k >= 0; k--) {
      lane = k == 0 ? frameLane : &ensureLane(k);
      ensurePaths((size_t)sppPerPass * F.numPixPadded, F.maxRec, aux);
      lane->paths.aux = aux ? lane->auxBase : nullptr;
      lane->paths.overflow = lane->overflow.as<uint32_t>();
    }
    if (F.nrcBounce >= 0) {
      ensureNrc(lane->pathSlots);
      F.nrcSelfTrain = r->nrcSelfTrain ? 1 : 0;
      if (!F.nrcSelfTrain) // training records in the query records: the path runs to its end
        lane->paths.nrcTrainRec = lane->paths.nrcRec;
    }
    for (int k = 0; k < 2; k++)
      if (!passEvents[k])
        HIP_CHECK(hipEventCreateWithFlags(&passEvents[k], hipEventDisableTiming));
    if ((int)laneEvents.size() < nl)
      for (int k = (int)laneEvents.size(); k < nl; k++) {
        hipEvent_t e;
        HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        laneEvents.push_back(e);
      }
    F.residentThreads = residentTraceBlocks * hpt_trace_block_size();
    const HptScene *S = reinterpret_cast<const HptScene *>(&sceneF);
    if (travStats)
      HIP_CHECK(hpt_fill_u32((uint32_t *)stats.ptr, 32, 0u, stream));
    uint32_t *const latch = capturing ? nullptr : fb->cancelLatch.as<uint32_t>();
    if (latch)
      HIP_CHECK(hpt_fill_u32(latch, 1, 0u, stream));
    F.cancelLatch = latch;
    if (nl > 1) { // fork: the other lanes start after the frame's earlier work
      HIP_CHECK(hipEventRecord(laneEvents[0], stream));
      for (int k = 1; k < nl; k++)
        HIP_CHECK(hipStreamWaitEvent(lanes[k]->stream, laneEvents[0], 0));
    }
    for (int s0 = 0, pass = 0; s0 < F.spp; s0 += sppPerPass, pass++) {
      lane = pass % nl == 0 ? frameLane : lanes[pass % nl].get();
      const hipStream_t stream = lane->stream;
      lane->paths.cancelFlag = latch ? cancelRingDev + fut->gen % kCancelRing : nullptr;
      lane->paths.cancelGen = fut->gen;
      lane->paths.cancelLatch = latch;
      const HptPaths *P = reinterpret_cast<const HptPaths *>(&lane->paths);
      F.passS0 = s0;
      F.passCount = std::min(sppPerPass, F.spp - s0);
      if (F.nrcBounce >= 0) { // every stride-th slot of the pass trains the cache
        const size_t slotsPass = (size_t)F.passCount * F.numPixPadded;
        F.nrcTrainStride = (uint32_t)std::max<double>(1.0, (double)slotsPass * r->nrcSurvival / r->nrcTrainPaths);
        F.nrcTrainOffset = (uint32_t)((size_t)fb->frameID * 7919u + (size_t)s0) % F.nrcTrainStride;
      }
      // the delta-tracking kernels take one ray (path) per thread: their
      // walks vary widely in length, and with every ray on a thread of its
      // own the hardware starts a new wave as each one ends, so the kernel
      // waits on its longest walk rather than on a thread's sum of walks
      const int volBlocks = std::max<int>(shadeBlocks, (int)(((size_t)F.passCount * F.numPixPadded + 255) / 256));
      hpt::DFrame Fno = F; // the frame as the bounces without a cache vertex see it
      Fno.nrcBounce = -1;
      const HptFrame *Fp = reinterpret_cast<const HptFrame *>(&F);
      const HptFrame *FpNo = reinterpret_cast<const HptFrame *>(&Fno);
      launch("reset", [&]() { return hpt_reset_counters(P, -1, stream); });
      launch("generate", [&]() { return hpt_generate(Fp, P, stream); });
      // with self-training no path goes beyond the vertex after the cache's
      // (in worlds without volumes: training paths from volume vertices run
      // to their end)
      int lastBounce = F.maxDepth;
      if (F.nrcBounce >= 0 && F.nrcSelfTrain && !F.hasVolumes)
        lastBounce = std::min(F.maxDepth, F.nrcBounce + 1);
      // Shadow stages of bounce b on the lane's side stream, beside the
      // extension of bounce b + 1 (they read and write disjoint path state;
      // the side stream has its own traversal scratch rows and resets its
      // own counters).  The shading of bounce b + 1 (and the free-path
      // sampling, which shares the paths' volume sampler with the
      // transmittance estimates) waits for bounce b's resolve, so every
      // contribution is added in the same order as on one stream.
      const bool overlap = overlapShadow && nl == 1 && !capturing && !inBatch && F.numFirstLightSamples > 0;
      const hipStream_t side = overlap ? ensureSide(*lane) : nullptr;
      hpt::DPaths sidePaths = lane->paths;
      if (overlap)
        sidePaths.overflow = lane->sideOverflow.as<uint32_t>();
      const HptPaths *PS = overlap ? reinterpret_cast<const HptPaths *>(&sidePaths) : P;
      const hipStream_t shadowStream = overlap ? side : stream;
      int pendingRes = -1; // bounce whose resolve the next shading must wait for
      int q = 0;
      for (int b = 0; b <= lastBounce; b++) {
        const std::string at = profile && profileBounces ? "@" + std::to_string(b) : std::string();
        auto nm = [&](const char *stage) { return at.empty() ? std::string(stage) : stage + at; };
        if (overlap)
          launch(nm("reset").c_str(), [&]() { return hpt_reset_extension_counters(P, q, stream); });
        else
          launch(nm("reset").c_str(), [&]() { return hpt_reset_counters(P, q, stream); });
        launch(nm("extend").c_str(),
            [&]() { return hpt_extend(S, P, q, residentTraceBlocks, travStats ? 1 : 0, stream); });
        launch(nm("ties").c_str(), [&]() { return hpt_ties(S, P, tieBlocks(), travStats ? 1 : 0, stream); });
        if (pendingRes >= 0) {
          HIP_CHECK(hipStreamWaitEvent(stream, lane->evRes[pendingRes & 1], 0));
          pendingRes = -1;
        }
        if (overlap)
          launch(nm("reset").c_str(), [&]() { return hpt_reset_append_counters(P, q, stream); });
        if (F.hasVolumes)
          launch(nm("volume").c_str(), [&]() { return hpt_volume(S, P, q, volBlocks, travStats ? 1 : 0, stream); });
        launch(nm("shade").c_str(),
            [&]() {
              const bool cacheVertex = b == F.nrcBounce || (F.nrcSelfTrain && b == F.nrcBounce + 1);
              return hpt_shade(S, cacheVertex ? Fp : FpNo, P, q, shadeBlocks, stream);
            });
        // the last bounce of a self-training radiance cache only ends paths
        // at cache queries (no NEE there)
        const bool suffixOnly = F.nrcBounce >= 0 && F.nrcSelfTrain && !F.hasVolumes && b == F.nrcBounce + 1;
        if (F.numFirstLightSamples > 0 && !suffixOnly) {
          if (overlap) {
            HIP_CHECK(hipEventRecord(lane->evShade[b & 1], stream));
            HIP_CHECK(hipStreamWaitEvent(side, lane->evShade[b & 1], 0));
            launch(nm("reset").c_str(), [&]() { return hpt_reset_shadow_counters(PS, side); }, side);
          }
          launch(nm("nee").c_str(), [&]() { return hpt_nee(S, Fp, PS, q, shadeBlocks, shadowStream); }, shadowStream);
          launch(nm("shadow").c_str(),
              [&]() { return hpt_shadow(S, Fp, PS, residentTraceBlocks, travStats ? 1 : 0, shadowStream); },
              shadowStream);
          launch(nm("shadow_trans").c_str(),
              [&]() { return hpt_shadow_trans(S, Fp, PS, residentTraceBlocks, travStats ? 1 : 0, shadowStream); },
              shadowStream);
          launch(nm("resolve").c_str(), [&]() { return hpt_resolve(S, Fp, PS, q, F.hasVolumes ? volBlocks : shadeBlocks, shadowStream); },
              shadowStream);
          if (overlap) {
            HIP_CHECK(hipEventRecord(lane->evRes[b & 1], side));
            pendingRes = b;
          }
        }
        q ^= 1;
      }
      if (pendingRes >= 0) // join: the pass's last resolve
        HIP_CHECK(hipStreamWaitEvent(stream, lane->evRes[pendingRes & 1], 0));
      if (F.nrcBounce >= 0) {
        const int64_t before = r->nrcRecords;
        nrcPass(r, w);
        const double slots = std::ceil((double)F.passCount * F.numPixPadded / F.nrcTrainStride);
        const double survival = std::min(1.0, std::max(1e-3, (double)(r->nrcRecords - before) / slots));
        r->nrcSurvival = 0.5 * r->nrcSurvival + 0.5 * survival;
      }
      if (nl > 1 && pass > 0) // accumulate after the previous pass
        HIP_CHECK(hipStreamWaitEvent(stream, passEvents[(pass - 1) & 1], 0));
      launch("finish_pass", [&]() { return hpt_finish_pass(Fp, P, stream); });
      if (nl > 1)
        HIP_CHECK(hipEventRecord(passEvents[pass & 1], stream));
      if (!capturing) {
        hipEvent_t e;
        HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        fut->passDone.push_back(e);
        HIP_CHECK(hipEventRecord(e, stream));
      }
    }
    lane = frameLane;
    for (int k = 1; k < nl; k++) { // join
      HIP_CHECK(hipEventRecord(laneEvents[k], lanes[k]->stream));
      HIP_CHECK(hipStreamWaitEvent(stream, laneEvents[k], 0));
    }
    if (!capturing)
      HIP_CHECK(hipEventRecord(fut->rendered, stream));
    const HptFrame *Fp = reinterpret_cast<const HptFrame *>(&F);
    launch("fb_accumulate", [&]() { return hpt_fb_accumulate(Fp, stream); });
    // LocalFrameBuffer::postProcess (LocalFB.cpp:459-478)
    // (a captured frame runs them on every launch; they check the device-
    // resident frame index and do their work after odd frames)
    if (fb->variance.ptr && ((fb->frameID & 1) == 1 || F.frameState)) {
      const int ntask = F.tilesX * F.tilesY;
      launch("variance", [&]() { return hpt_variance(Fp, stream); });
      launch("variance_reduce", [&]() {
        return hpt_variance_reduce(F.taskVariance, ntask, fb->errorThreshold, fb->varianceOut.as<float>(),
            F.frameState, stream);
      });
      if (!capturing) // a captured batch leaves the error estimate on the device
        HIP_CHECK(hipMemcpyAsync(fb->hostVariance, fb->varianceOut.ptr, sizeof(float), hipMemcpyDeviceToHost, stream));
      fb->varianceValid = true;
    }
    if (fb->color.ptr && fb->format != OSP_FB_NONE) {
      launch("color_convert", [&]() {
        return hpt_color_convert(F.fbColor, F.width * F.height, (int)fb->format, fb->converted.ptr, stream);
      });
    }
  } else {
    if (!capturing)
      HIP_CHECK(hipEventRecord(fut->rendered, stream));
  }
  if (!capturing)
    HIP_CHECK(hipEventRecord(fut->finished, stream));
  return fut;
}

int HipDevice::traceRays(World *w, const float *rays8, int n, float *outTUV, uint32_t *outIDs, uint64_t *stats2)
{
  if (!w->built)
    throw std::runtime_error("world not committed");
  gpu();
  DevBuf dr, dt, di;
  dr.ensure((size_t)std::max(n, 1) * 32);
  dt.ensure((size_t)std::max(n, 1) * 12);
  di.ensure((size_t)std::max(n, 1) * 12);
  HIP_CHECK(hipMemcpy(dr.ptr, rays8, (size_t)n * 32, hipMemcpyHostToDevice));
  HIP_CHECK(hpt_fill_u32((uint32_t *)stats.ptr, 32, 0u, stream));
  HIP_CHECK(hpt_trace_rays(reinterpret_cast<const HptScene *>(&w->scene), dr.as<float>(), n, dt.as<float>(),
      di.as<uint32_t>(), lanes[0]->overflow.as<uint32_t>(), residentTraceBlocks, (unsigned long long *)stats.ptr,
      stream));
  HIP_CHECK(hipStreamSynchronize(stream));
  HIP_CHECK(hipMemcpy(outTUV, dt.ptr, (size_t)n * 12, hipMemcpyDeviceToHost));
  HIP_CHECK(hipMemcpy(outIDs, di.ptr, (size_t)n * 12, hipMemcpyDeviceToHost));
  if (stats2) {
    unsigned long long s[2];
    HIP_CHECK(hipMemcpy(s, stats.ptr, 16, hipMemcpyDeviceToHost));
    stats2[0] = s[0];
    stats2[1] = s[1];
  }
  return n;
}

OSPPickResult HipDevice::pick(ManagedObject *fbo, ManagedObject *, ManagedObject *co, ManagedObject *wo, float x, float y)
{
  // Renderer_pick (render/Renderer.ispc:43-95): camera ray through the
  // screen position, closest hit
  OSPPickResult res;
  std::memset(&res, 0, sizeof(res));
  res.primID = hpt::kInvalidID;
  World *w = (World *)wo;
  Camera *cam = (Camera *)co;
  (void)fbo;
  hpt::DFrame F;
  std::memset(&F, 0, sizeof(F));
  cam->setup(F);
  const vec3f dir = hpt::normalize(hpt::ld3(F.camDir00) + x * hpt::ld3(F.camDu) + y * hpt::ld3(F.camDv));
  const float ray[8] = {F.camOrg[0], F.camOrg[1], F.camOrg[2], dir.x, dir.y, dir.z, F.nearClip, hpt::kInf};
  float tuv[3];
  uint32_t ids[3];
  traceRays(w, ray, 1, tuv, ids, nullptr);
  if (ids[0] == hpt::kInvalidID)
    return res;
  Instance *inst = w->instances[ids[0]];
  GeometricModel *model = inst->group->models[ids[1]];
  inst->refInc();
  model->refInc();
  res.hasHit = 1;
  res.instance = (OSPInstance)inst;
  res.model = (OSPGeometricModel)model;
  res.primID = ids[2];
  for (int k = 0; k < 3; k++)
    res.worldPosition[k] = F.camOrg[k] + tuv[0] * (&dir.x)[k];
  return res;
}

static HipDevice *current()
{
  Device *d = ospray_current_device();
  return dynamic_cast<HipDevice *>(d);
}

} // namespace hipdev

using namespace hipdev;

extern "C" {

// module entry point (modules/cpu/ModuleInit.cpp:24-54 pattern)
OSPRAY_INTERFACE OSPError ospray_module_init_hip(int16_t major, int16_t minor, int16_t patch)
{
  (void)patch;
  if (major != OSPRAY_VERSION_MAJOR || minor != OSPRAY_VERSION_MINOR)
    return OSP_VERSION_MISMATCH;
  ospray_register_device_type("hip", []() -> Device * { return new HipDevice(); });
  return OSP_NO_ERROR;
}

OSPRAY_INTERFACE int ospHipGetKernelStats(OSPHipKernelStat *out, int maxEntries)
{
  HipDevice *d = current();
  if (!d)
    return -1;
  d->collectTimers();
  int n = 0;
  for (auto &t : d->timers) {
    if (n >= maxEntries)
      break;
    std::memset(out[n].name, 0, sizeof(out[n].name));
    std::strncpy(out[n].name, t.name.c_str(), sizeof(out[n].name) - 1);
    out[n].launches = t.launches;
    out[n].totalMs = t.totalMs;
    n++;
  }
  return n;
}

OSPRAY_INTERFACE int ospHipResetKernelStats(void)
{
  HipDevice *d = current();
  if (!d)
    return -1;
  d->collectTimers();
  d->timers.clear();
  return 0;
}

OSPRAY_INTERFACE int ospHipGetTraversalStats(uint64_t out9[9])
{
  HipDevice *d = current();
  if (!d)
    return -1;
  if (!d->stream)
    return -2;
  try {
    unsigned long long s[9];
    HIP_CHECK(hipStreamSynchronize(d->stream));
    HIP_CHECK(hipMemcpy(s, d->stats.ptr, sizeof(s), hipMemcpyDeviceToHost));
    for (int k = 0; k < 9; k++)
      out9[k] = s[k];
  } catch (const std::exception &) {
    return -2;
  }
  return 0;
}

static void extensionError(HipDevice *d, const char *msg);

static bool batchViews(int n, const OSPFrameBuffer *fbs, const OSPRenderer *rs, const OSPCamera *cams,
    std::vector<HipDevice::BatchView> &v)
{
  if (n <= 0 || !fbs || !rs || !cams)
    return false;
  for (int i = 0; i < n; i++) {
    ManagedObject *f = (ManagedObject *)fbs[i], *r = (ManagedObject *)rs[i], *c = (ManagedObject *)cams[i];
    if (!f || !r || !c || f->managedType != OSP_FRAMEBUFFER || r->managedType != OSP_RENDERER
        || c->managedType != OSP_CAMERA)
      return false;
    v.push_back({f, r, c});
  }
  return true;
}

OSPRAY_INTERFACE int ospHipRenderBatch(int n, const OSPFrameBuffer *fbs, const OSPRenderer *renderers,
    const OSPCamera *cameras, OSPWorld world, int mode, float *viewMs)
{
  HipDevice *d = current();
  std::vector<HipDevice::BatchView> v;
  if (!d || !world || !batchViews(n, fbs, renderers, cameras, v) || (mode != 0 && mode != 1))
    return -1;
  try {
    d->renderBatch(v, (ManagedObject *)world, mode, viewMs);
  } catch (const std::exception &e) {
    extensionError(d, e.what());
    return -2;
  }
  return 0;
}

OSPRAY_INTERFACE float ospHipTuneBatch(int n, const OSPFrameBuffer *fbs, const OSPRenderer *renderers,
    const OSPCamera *cameras, OSPWorld world, int reps, int *best5)
{
  HipDevice *d = current();
  std::vector<HipDevice::BatchView> v;
  if (!d || !world || !batchViews(n, fbs, renderers, cameras, v))
    return -1.f;
  try {
    return d->tuneBatch(v, (ManagedObject *)world, reps, best5);
  } catch (const std::exception &e) {
    extensionError(d, e.what());
    return -2.f;
  }
}

OSPRAY_INTERFACE int ospHipGetNrcState(OSPRenderer renderer, float *loss, int *steps, int64_t *queries,
    int64_t *trainingRecords)
{
  HipDevice *d = current();
  if (!d || !renderer || !loss || !steps || !queries || !trainingRecords)
    return -1;
  ManagedObject *o = (ManagedObject *)renderer;
  if (o->managedType != OSP_RENDERER)
    return -1;
  Renderer *r = (Renderer *)o;
  *loss = 0.f;
  *steps = r->nrcInit ? r->nrcStep : 0;
  *queries = r->nrcQueries;
  *trainingRecords = r->nrcRecords;
  if (!d->stream || !r->nrcInit || r->nrcStep == 0)
    return 0;
  try {
    HIP_CHECK(hipStreamSynchronize(d->stream));
    HIP_CHECK(hipMemcpy(loss, r->nLoss.ptr, sizeof(float), hipMemcpyDeviceToHost));
  } catch (const std::exception &) {
    return -2;
  }
  return 0;
}

OSPRAY_INTERFACE int ospHipGetVolumeStats(uint64_t out6[6])
{
  HipDevice *d = current();
  if (!d)
    return -1;
  if (!d->stream)
    return -2;
  try {
    unsigned long long s[15];
    HIP_CHECK(hipStreamSynchronize(d->stream));
    HIP_CHECK(hipMemcpy(s, d->stats.ptr, sizeof(s), hipMemcpyDeviceToHost));
    for (int k = 0; k < 6; k++)
      out6[k] = s[9 + k];
  } catch (const std::exception &) {
    return -2;
  }
  return 0;
}

// errors of the extension entry points reach the application like those of
// the C API: last error code and message, and the error callback
static void extensionError(HipDevice *d, const char *msg)
{
  d->lastErrorCode = OSP_UNKNOWN_ERROR;
  d->lastErrorMsg = msg;
  if (d->errorFcn)
    d->errorFcn(d->errorUserData, OSP_UNKNOWN_ERROR, msg);
}

OSPRAY_INTERFACE int ospHipTraceRays(OSPWorld world, const float *rays8, int n, float *outTUV, uint32_t *outIDs, uint64_t *stats2)
{
  HipDevice *d = current();
  if (!d || !world)
    return -1;
  try {
    return d->traceRays((World *)world, rays8, n, outTUV, outIDs, stats2);
  } catch (const std::exception &e) {
    extensionError(d, e.what());
    return -2;
  }
}

OSPRAY_INTERFACE const void *ospHipGetWorldLaunchScene(OSPWorld world)
{
  HipDevice *d = current();
  if (!d || !world)
    return nullptr;
  World *w = (World *)world;
  if (w->managedType != OSP_WORLD || !w->built)
    return nullptr;
  return &w->scene;
}

OSPRAY_INTERFACE int ospHipGetDeviceInfo(char *name, int len, int *numCUs)
{
  HipDevice *d = current();
  if (!d)
    return -1;
  try {
    d->gpu();
  } catch (const std::exception &) {
    return -2;
  }
  if (name && len > 0) {
    std::strncpy(name, d->gpuName.c_str(), len - 1);
    name[len - 1] = 0;
  }
  if (numCUs)
    *numCUs = d->numCUs;
  return 0;
}

OSPRAY_INTERFACE int ospHipCopyBandwidth(size_t bytes, int reps, float *msPerCopy)
{
  HipDevice *d = current();
  if (!d || reps < 1 || bytes < 16 || !msPerCopy)
    return -1;
  void *src = nullptr, *dst = nullptr;
  hipEvent_t a = nullptr, b = nullptr;
  int rc = 0;
  try {
    d->gpu();
    bytes &= ~(size_t)15;
    HIP_CHECK(hipMalloc(&src, bytes));
    HIP_CHECK(hipMalloc(&dst, bytes));
    HIP_CHECK(hpt_fill_u32((uint32_t *)src, bytes / 4, 0x3f800000u, d->stream));
    HIP_CHECK(hpt_copy(dst, src, bytes, d->stream)); // untimed first touch
    HIP_CHECK(hipEventCreate(&a));
    HIP_CHECK(hipEventCreate(&b));
    HIP_CHECK(hipEventRecord(a, d->stream));
    for (int i = 0; i < reps; i++)
      HIP_CHECK(hpt_copy(dst, src, bytes, d->stream));
    HIP_CHECK(hipEventRecord(b, d->stream));
    HIP_CHECK(hipEventSynchronize(b));
    float ms = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms, a, b));
    *msPerCopy = ms / reps;
  } catch (const std::exception &e) {
    extensionError(d, e.what());
    rc = -2;
  }
  if (a)
    (void)hipEventDestroy(a);
  if (b)
    (void)hipEventDestroy(b);
  if (src)
    (void)hipFree(src);
  if (dst)
    (void)hipFree(dst);
  return rc;
}

OSPRAY_INTERFACE int ospHipGetWorldBuild(OSPWorld world, double *blasBuildMs, int *deviceBlas, int *stackNeed)
{
  HipDevice *d = current();
  if (!d || !world)
    return -1;
  World *w = (World *)world;
  if (blasBuildMs)
    *blasBuildMs = w->blasBuildMs;
  if (deviceBlas)
    *deviceBlas = w->deviceBlas;
  if (stackNeed)
    *stackNeed = w->stackNeed;
  return 0;
}

OSPRAY_INTERFACE int ospHipGetWorldInfo(OSPWorld world, int64_t out4[4])
{
  HipDevice *d = current();
  if (!d || !world)
    return -1;
  World *w = (World *)world;
  out4[0] = w->totalNodes;
  out4[1] = w->totalPrims;
  out4[2] = w->maxDepth;
  out4[3] = w->scene.numInstances;
  return 0;
}

} // extern "C"
