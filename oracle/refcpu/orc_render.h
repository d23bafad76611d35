// ORACLE — test infrastructure only.
//
// Restatement of the reference `pathtracer` renderer on the CPU:
//   per pixel      modules/cpu/render/pathtracer/PathTracer.ispc:75-216
//   bounce loop    modules/cpu/render/pathtracer/PathSampler.ispc:109-415
//   NEE            modules/cpu/render/pathtracer/NextEventEstimation.ispc:25-160
//   shadows        modules/cpu/render/pathtracer/TransparentShadow.ispc:18-85
//   virtual lights modules/cpu/render/pathtracer/VirtualLight.ispc:17-107
//   accumulation   modules/cpu/fb/LocalFB.ispc:14-78, fb/FrameBuffer.cpp:115-137
//   variance       modules/cpu/fb/frame_ops/Variance.ispc:11-59, Variance.cpp:32-46
//   colour formats modules/cpu/fb/frame_ops/ColorConversion.ispc:11-36
#pragma once

#include <cstdint>
#include <random>
#include <vector>

#include "orc_geom.h"
#include "orc_sampler.h"
#include "orc_volume.h"

namespace orc {

static constexpr float kUlpEpsilon = 0.000000504636764526367187500f; // 0x1.0fp-21f

struct DG
{
  vec3f P, Ng, Ns, dPds, dPdt;
  vec4f color;
  float epsilon;
  float areaPDF; // of the model as a geometry light (GeometricModel.ih:58)
  int material; // -1 = none
  uint32_t primID, objID, instID;
};

struct Medium
{
  vec3f attenuation;
  float ior;
};
inline bool eq(const Medium &a, const Medium &b) { return a.ior == b.ior && eq(a.attenuation, b.attenuation); }

struct MaterialData
{
  int type;
  // obj
  vec3f Kd, Tf, Ks;
  float d, Ns;
  // metal
  vec3f eta, k;
  float roughness;
  bool spectral; // eta and k sampled at 430, 465, ..., 675 nm
  float etaS[8], kS[8];
  // glass
  Medium inside, outside;
  // luminous (Luminous.cpp:37-48)
  vec3f radiance;
  float transparency;
  bool emissive;
};

struct LightData
{
  int type;
  bool visible;
  // quad
  vec3f radiance, position, edge1, edge2, nnormal;
  float ppdf;
  // distant
  linear3f frame;
  vec3f irradiance;
  float cosAngle, pdf;
  // geometry light (GeometryLightShared.h): an emissive model of one
  // instance, primitives chosen by area
  int inst, model; // instance index, OrcModel index
  std::vector<float> primCDF;
  float geoPdf; // 1 / total world-space area
};

struct CameraData
{
  vec3f org, dir_00, du_size, dv_up, direction;
  float nearClip;
};

struct RendererData
{
  int spp, maxDepth, rouletteDepth, maxScatteringEvents, numLightSamples;
  bool limitIndirectLightSamples;
  float minContribution, maxRadiance;
  vec4f bgColor;
  int filterType; // 0 point, 1 box, 2 LUT
  float filterWidth;
  int lutSize;
  std::vector<float> lutCdfX, lutCdfY;
  float errorThreshold;
  bool geometryLights;
};

struct Scene
{
  World world;
  std::vector<MaterialData> materials;
  // [geometry lights, world lights] (PathTracerData.cpp:32-82) and the
  // uniform selection CDFs with and without the geometry lights
  std::vector<LightData> lights;
  uint32_t numGeoLights = 0;
  std::vector<float> lightsCDF, worldLightsCDF;
  std::vector<float> modelAreaPDF; // per OrcModel
  std::vector<VolumeData> volumes; // per OrcVolume
  // every (instance, volume) pair in instance order, then group order: the
  // order in which a ray meets the volume intervals
  struct VolInst
  {
    int inst, vol, local; // instance, OrcVolume, index in the group
  };
  std::vector<VolInst> volInsts;
  bool hasVolumes = false; // some instance holds a volume (FFO_VOLUME_IN_SCENE)
};

struct FrameBuffer
{
  int w, h;
  uint32_t channels;
  int format;
  int frameID = -1;
  bool accumulateVariance = false;
  std::mt19937 mtGen{43};
  bool pickOdd = false;
  int targetFrames = 0;
  bool projectedDepth = false;
  std::vector<vec4f> color, variance;
  std::vector<float> depth;
  std::vector<vec3f> normal, firstNormal, albedo, position;
  std::vector<uint32_t> primID, objID, instID;
  std::vector<float> taskVariance;
  bool varianceValid = false;
  std::vector<uint8_t> mapped; // converted colour for OSP_FB_RGBA8 / SRGBA / RGBA32F
  void clear();
};

void setupScene(Scene &s, const OrcScene &desc);
void setupCamera(CameraData &c, const OrcCamera &desc);
void setupRenderer(RendererData &r, const OrcRenderer &desc);
void initFrameBuffer(FrameBuffer &fb, int w, int h, int format, uint32_t channels);

// one ospRenderFrame: beginFrame, render rows [y0, y1) (all rows: 0, h),
// then postProcess (variance, colour conversion)
void renderFrame(const Scene &s, FrameBuffer &fb, const CameraData &cam, const RendererData &r, int nthreads, int y0, int y1);
float getVariance(const FrameBuffer &fb, const RendererData &r);

// pixel filter sample and camera ray of one sample (test hooks)
vec2f pixelFilterSample(const RendererData &r, vec2f s);
void cameraRay(const CameraData &c, vec2f screen, Ray &ray);

// portable transcendental float functions
float hostExp(float x);

// volume sampling test hooks: transmittance (0/1) and free path of a world
// ray, Henyey-Greenstein sampling and evaluation
float volumeTransmittanceHook(const Scene &s, const Ray &ray, float r0, float r1, RandomSampler &rs);
float volumeFreePathHook(const Scene &s, Ray &ray, float r0, float r1, RandomSampler &rs, int &volInst, vec3f &albedo);
void hgSampleHook(float g, vec3f wo, vec2f s, vec3f &wi, float &pdf, uint32_t &type);
// the OBJ specular lobe of reflectivity ks and Phong exponent ns with the
// shading normal +z: eval(wo, wi) -> out[0..3] (value, pdf); sample(wo, s)
// -> out[4..10] (wi, pdf, weight)
void specularHook(vec3f ks, float ns, vec3f wo, vec3f wi, vec2f s, float out[11]);
// the BSDF of one material (shading frame = world axes, entered from the
// outside medium): eval(wo, wi) value and pdf, then sample(wo, s, ss) wi,
// pdf, weight, type, and the BSDF's scattering type
void bsdfHook(const OrcMaterial &m, vec3f wo, vec3f wi, vec2f s, float ss, float out[13]);
float hgEvalHook(float g, vec3f wo, vec3f wi);
// Variance_kernel over the frame buffer's colour and variance buffers as
// they stand (fills taskVariance, marks the estimate valid)
void varianceHook(FrameBuffer &fb);
// lightSample of world light `light` (index into the scene's light list)
// at a surface point P with normals Ng, Ns: out = dir (3), dist, weight (3),
// pdf
void lightSampleHook(const Scene &s, int light, vec3f P, vec3f Ng, vec3f Ns, vec2f smp, float out[8]);

} // namespace orc
