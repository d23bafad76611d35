// ORACLE — test infrastructure only.
//
// C ABI of liborc.so, loaded with ctypes by tests/ and by bench.py's
// cpu_baseline leg.  Nothing in the product library links or loads it.
#include <cstring>
#include <memory>

#include "orc_render.h"

using namespace orc;

namespace {
struct OrcSceneHandle
{
  Scene scene;
};
struct OrcFbHandle
{
  FrameBuffer fb;
};
} // namespace

extern "C" {

void *orc_scene_create(const OrcScene *desc)
{
  auto *h = new OrcSceneHandle;
  setupScene(h->scene, *desc);
  return h;
}

void orc_scene_destroy(void *h) { delete (OrcSceneHandle *)h; }

void *orc_fb_create(int w, int h, int format, uint32_t channels)
{
  auto *f = new OrcFbHandle;
  initFrameBuffer(f->fb, w, h, format, channels);
  return f;
}

void orc_fb_destroy(void *h) { delete (OrcFbHandle *)h; }

void orc_fb_reset(void *h) { ((OrcFbHandle *)h)->fb.clear(); }

void orc_fb_set_projected_depth(void *h, int on) { ((OrcFbHandle *)h)->fb.projectedDepth = on != 0; }

int orc_fb_frame_id(void *h) { return ((OrcFbHandle *)h)->fb.frameID; }

void orc_render_frame(void *scene, void *fb, const OrcCamera *cam, const OrcRenderer *ren, int nthreads, int y0, int y1)
{
  CameraData c;
  setupCamera(c, *cam);
  RendererData r;
  setupRenderer(r, *ren);
  FrameBuffer &f = ((OrcFbHandle *)fb)->fb;
  if (y1 < 0 || y1 > f.h)
    y1 = f.h;
  renderFrame(((OrcSceneHandle *)scene)->scene, f, c, r, nthreads, y0, y1);
}

float orc_fb_variance(void *fb, const OrcRenderer *ren)
{
  RendererData r;
  setupRenderer(r, *ren);
  return getVariance(((OrcFbHandle *)fb)->fb, r);
}

// pointer to a channel buffer (OSPFrameBufferChannel value); colour returns
// the converted buffer in the framebuffer format
const void *orc_fb_map(void *fb, uint32_t channel)
{
  FrameBuffer &f = ((OrcFbHandle *)fb)->fb;
  switch (channel) {
  case 1:
    return f.mapped.empty() ? nullptr : (const void *)f.mapped.data();
  case 2:
    return f.depth.empty() ? nullptr : (const void *)f.depth.data();
  case 16:
    return f.normal.empty() ? nullptr : (const void *)f.normal.data();
  case 32:
    return f.albedo.empty() ? nullptr : (const void *)f.albedo.data();
  case 64:
    return f.primID.empty() ? nullptr : (const void *)f.primID.data();
  case 128:
    return f.objID.empty() ? nullptr : (const void *)f.objID.data();
  case 256:
    return f.instID.empty() ? nullptr : (const void *)f.instID.data();
  case 512:
    return f.firstNormal.empty() ? nullptr : (const void *)f.firstNormal.data();
  case 1024:
    return f.position.empty() ? nullptr : (const void *)f.position.data();
  default:
    return nullptr;
  }
}

// raw accumulated RGBA32F colour (before format conversion)
const float *orc_fb_accum_color(void *fb)
{
  FrameBuffer &f = ((OrcFbHandle *)fb)->fb;
  return f.color.empty() ? nullptr : &f.color[0].x;
}

const float *orc_fb_task_variance(void *fb, int *n)
{
  FrameBuffer &f = ((OrcFbHandle *)fb)->fb;
  *n = (int)f.taskVariance.size();
  return f.taskVariance.data();
}

// ---- unit-level hooks used by the golden / property tests ------------------

void orc_sobol_rev_table(uint32_t *out128) { std::memcpy(out128, sobolTable().rev, sizeof(uint32_t) * 128); }

// kind: 0 MurmurHash3_mix(x, seed), 1 MurmurHash3_finalize(x), 2 tripleHash,
// 3 OwenHash2(x, seed), 4 OwenScramble4(x, seed), 5 reverseBits,
// 6 interleaveZOrder16(x, seed)
void orc_hash(int kind, const uint32_t *x, const uint32_t *seed, uint32_t *out, int64_t n)
{
  for (int64_t i = 0; i < n; i++) {
    const uint32_t a = x[i], s = seed ? seed[i] : 0;
    switch (kind) {
    case 0: out[i] = MurmurHash3_mix(a, s); break;
    case 1: out[i] = MurmurHash3_finalize(a); break;
    case 2: out[i] = tripleHash(a); break;
    case 3: out[i] = OwenHash2(a, s); break;
    case 4: out[i] = OwenScramble4(a, s); break;
    case 5: out[i] = reverseBits(a); break;
    default: out[i] = interleaveZOrder16(a, s); break;
    }
  }
}

// raw PCG32 steps from (state, increment-stream) without the hashing init
void orc_pcg_raw(uint64_t state, uint32_t stream, uint32_t *out, int n)
{
  RandomSampler r{state, stream};
  for (int i = 0; i < n; i++)
    out[i] = pcg32(r);
}

void orc_pcg_init_sequence(uint32_t seed, uint32_t stream, uint32_t *out, int n)
{
  RandomSampler r;
  RandomSampler_init(r, seed, stream);
  for (int i = 0; i < n; i++)
    out[i] = pcg32(r);
}

// first five LD dimensions of a pixel sample: pixel (2), lens (2), time (1)
void orc_ld_samples5(uint32_t seed, uint32_t sampleIndex, uint32_t bits, float *out5)
{
  LDSampler ld;
  LDSampler_init(ld, seed, sampleIndex, 0, bits, bits);
  vec3ui ss;
  const vec2f p = LDSampler_getNext5Samples(ld, ss);
  const vec2f lens = LDSampler_finalizeDim23(ld, ss);
  out5[0] = p.x;
  out5[1] = p.y;
  out5[2] = lens.x;
  out5[3] = lens.y;
  out5[4] = LDSampler_finalizeDim4(ld, ss);
}

// kind: 0 sin, 1 cos, 2 acos, 3 exp, 4 log
void orc_fmath(int kind, const float *in, float *out, int64_t n)
{
  for (int64_t i = 0; i < n; i++) {
    float s, c;
    switch (kind) {
    case 0: fsincos(in[i], s, c); out[i] = s; break;
    case 1: fsincos(in[i], s, c); out[i] = c; break;
    case 2: out[i] = facos(in[i]); break;
    case 3: out[i] = fexp(in[i]); break;
    default: out[i] = flog(in[i]); break;
    }
  }
}

void orc_specular(const float *ks3, float ns, const float *wo3, const float *wi3, const float *s2, float *out11,
    int64_t n)
{
  for (int64_t i = 0; i < n; i++)
    specularHook(v3(ks3[0], ks3[1], ks3[2]), ns, v3(wo3[3 * i], wo3[3 * i + 1], wo3[3 * i + 2]),
        v3(wi3[3 * i], wi3[3 * i + 1], wi3[3 * i + 2]), v2(s2[2 * i], s2[2 * i + 1]), out11 + 11 * i);
}

void orc_bsdf(const OrcMaterial *m, const float *wo3, const float *wi3, const float *s2, const float *ss,
    float *out13, int64_t n)
{
  for (int64_t i = 0; i < n; i++)
    bsdfHook(*m, v3(wo3[3 * i], wo3[3 * i + 1], wo3[3 * i + 2]), v3(wi3[3 * i], wi3[3 * i + 1], wi3[3 * i + 2]),
        v2(s2[2 * i], s2[2 * i + 1]), ss[i], out13 + 13 * i);
}

void orc_fpow(const float *x, const float *y, float *out, int64_t n)
{
  for (int64_t i = 0; i < n; i++)
    out[i] = fpow(x[i], y[i]);
}

void orc_pixel_filter(const OrcRenderer *ren, const float *s2, float *out2, int64_t n)
{
  RendererData r;
  setupRenderer(r, *ren);
  for (int64_t i = 0; i < n; i++) {
    const vec2f o = pixelFilterSample(r, v2(s2[2 * i], s2[2 * i + 1]));
    out2[2 * i] = o.x;
    out2[2 * i + 1] = o.y;
  }
}

void orc_camera_rays(const OrcCamera *cam, const float *screen2, float *out6, int64_t n)
{
  CameraData c;
  setupCamera(c, *cam);
  for (int64_t i = 0; i < n; i++) {
    Ray r;
    cameraRay(c, v2(screen2[2 * i], screen2[2 * i + 1]), r);
    out6[6 * i + 0] = r.org.x;
    out6[6 * i + 1] = r.org.y;
    out6[6 * i + 2] = r.org.z;
    out6[6 * i + 3] = r.dir.x;
    out6[6 * i + 4] = r.dir.y;
    out6[6 * i + 5] = r.dir.z;
  }
}

// closest hits for rays given as (org3, dir3, t0, t1); out: t,u,v (floats)
// and inst, model, prim (uint32, 0xffffffff on miss).  fp64: 0 the BVH
// traversal of extension rays (near-ties re-resolved in FP64), 1 the
// double-precision brute force, 2 the BVH traversal in float order only
// (the shadow rays' closest hits).
void orc_intersect(void *scene, const float *rays8, int64_t n, int fp64, float *outTUV, uint32_t *outIDs, uint64_t *stats2)
{
  const World &w = ((OrcSceneHandle *)scene)->scene.world;
  uint64_t nodes = 0, prims = 0;
  for (int64_t i = 0; i < n; i++) {
    const float *r = rays8 + 8 * i;
    Ray ray{v3(r[0], r[1], r[2]), v3(r[3], r[4], r[5]), r[6], r[7]};
    const Hit h = fp64 == 1 ? intersectBruteForceFP64(w, ray) : w.intersect(ray, &nodes, &prims, fp64 == 0);
    outTUV[3 * i] = h.hit ? h.t : kInf;
    outTUV[3 * i + 1] = h.u;
    outTUV[3 * i + 2] = h.v;
    outIDs[3 * i] = h.hit ? h.inst : kInvalidID;
    outIDs[3 * i + 1] = h.hit ? h.model : kInvalidID;
    outIDs[3 * i + 2] = h.hit ? h.prim : kInvalidID;
  }
  if (stats2) {
    stats2[0] = nodes;
    stats2[1] = prims;
  }
}


// ---- volume hooks ----------------------------------------------------------------

// trilinear samples of volume `vol` at volume-space points (NaN outside)
void orc_volume_sample(void *scene, int vol, const float *p3, float *out, int64_t n)
{
  const VolumeData &v = ((OrcSceneHandle *)scene)->scene.volumes[vol];
  for (int64_t i = 0; i < n; i++)
    out[i] = sampleVolume(v, v3(p3[3 * i], p3[3 * i + 1], p3[3 * i + 2]));
}

// transfer function: rgba per value; max opacity per value range
void orc_tf_get(void *scene, int vol, const float *values, float *out4, int64_t n)
{
  const VolumeData &v = ((OrcSceneHandle *)scene)->scene.volumes[vol];
  for (int64_t i = 0; i < n; i++) {
    const vec4f c = tfGet(v, values[i]);
    out4[4 * i] = c.x;
    out4[4 * i + 1] = c.y;
    out4[4 * i + 2] = c.z;
    out4[4 * i + 3] = c.w;
  }
}

void orc_tf_max_opacity(void *scene, int vol, const float *ranges2, float *out, int64_t n)
{
  const VolumeData &v = ((OrcSceneHandle *)scene)->scene.volumes[vol];
  for (int64_t i = 0; i < n; i++)
    out[i] = tfMaxOpacity(v, ranges2[2 * i], ranges2[2 * i + 1]);
}

// macrocell intervals of a volume-space ray (org, dir, t0, t1) clipped to the
// box: up to maxn rows (ta, tb, valueLo, valueHi); returns the count
int orc_volume_intervals(void *scene, int vol, const float *ray8, float *out4, int maxn)
{
  const VolumeData &v = ((OrcSceneHandle *)scene)->scene.volumes[vol];
  const vec3f o = v3(ray8[0], ray8[1], ray8[2]), d = v3(ray8[3], ray8[4], ray8[5]);
  float lo, hi;
  if (!boxInterval(v, o, d, ray8[6], ray8[7], lo, hi))
    return 0;
  int n = 0;
  forEachInterval(v, o, d, lo, hi, [&](float ta, float tb, float vlo, float vhi) {
    if (n < maxn) {
      out4[4 * n] = ta;
      out4[4 * n + 1] = tb;
      out4[4 * n + 2] = vlo;
      out4[4 * n + 3] = vhi;
    }
    n++;
    return true;
  });
  return n;
}

// per world ray (org, dir, t0, t1), `trials` independent estimates with the
// PCG stream (seed, i * trials + k): mean transmittance and mean free path
// (inf counted as t1)
void orc_volume_estimates(
    void *scene, const float *rays8, int64_t n, int trials, uint32_t seed, float *outT, float *outHitFrac, float *outMeanDist)
{
  const Scene &s = ((OrcSceneHandle *)scene)->scene;
  for (int64_t i = 0; i < n; i++) {
    const float *r = rays8 + 8 * i;
    Ray ray{v3(r[0], r[1], r[2]), v3(r[3], r[4], r[5]), r[6], r[7]};
    double sumT = 0.0, hits = 0.0, sumD = 0.0;
    for (int k = 0; k < trials; k++) {
      RandomSampler rs;
      RandomSampler_init(rs, seed, (uint32_t)(i * trials + k));
      sumT += volumeTransmittanceHook(s, ray, ray.t0, ray.t, rs);
      RandomSampler_init(rs, seed ^ 0x9e3779b9u, (uint32_t)(i * trials + k));
      Ray fr = ray;
      int vi;
      vec3f alb;
      const float d = volumeFreePathHook(s, fr, ray.t0, ray.t, rs, vi, alb);
      if (d < kInf) {
        hits += 1.0;
        sumD += d;
      }
    }
    outT[i] = (float)(sumT / trials);
    outHitFrac[i] = (float)(hits / trials);
    outMeanDist[i] = hits > 0 ? (float)(sumD / hits) : 0.f;
  }
}

// Henyey-Greenstein: sample (wi, pdf, type) for (g, wo, s) rows; eval pdf
void orc_hg_sample(float g, const float *wo3, const float *s2, float *outWi3, float *outPdf, uint32_t *outType, int64_t n)
{
  for (int64_t i = 0; i < n; i++) {
    vec3f wi;
    hgSampleHook(g, v3(wo3[3 * i], wo3[3 * i + 1], wo3[3 * i + 2]), v2(s2[2 * i], s2[2 * i + 1]), wi, outPdf[i], outType[i]);
    outWi3[3 * i] = wi.x;
    outWi3[3 * i + 1] = wi.y;
    outWi3[3 * i + 2] = wi.z;
  }
}

void orc_hg_eval(float g, const float *wo3, const float *wi3, float *outPdf, int64_t n)
{
  for (int64_t i = 0; i < n; i++)
    outPdf[i] = hgEvalHook(g, v3(wo3[3 * i], wo3[3 * i + 1], wo3[3 * i + 2]), v3(wi3[3 * i], wi3[3 * i + 1], wi3[3 * i + 2]));
}

// Variance_kernel + getAvgError on given accumulated colour and variance
// buffers (RGBA float, w*h each): per-task errors into outTasks
// (ceil(w/8)*ceil(h/8) floats), the frame estimate into *outAvg
void orc_variance_buffers(int w, int h, const float *color4, const float *variance4, float threshold, float *outTasks,
    float *outAvg)
{
  FrameBuffer fb;
  initFrameBuffer(fb, w, h, 0, 1u | 4u | 8u);
  const size_t n = (size_t)w * h;
  for (size_t i = 0; i < n; i++) {
    fb.color[i] = {color4[4 * i], color4[4 * i + 1], color4[4 * i + 2], color4[4 * i + 3]};
    fb.variance[i] = {variance4[4 * i], variance4[4 * i + 1], variance4[4 * i + 2], variance4[4 * i + 3]};
  }
  varianceHook(fb);
  std::memcpy(outTasks, fb.taskVariance.data(), fb.taskVariance.size() * sizeof(float));
  RendererData r{};
  r.errorThreshold = threshold;
  *outAvg = getVariance(fb, r);
}

// Distribution1D_create over `weights` (n values), then Distribution1D_sample
// for each of the m values in s: idx, and (prob, pdf, frac) per sample
float orc_distribution1d(const float *weights, int n, const float *s, int64_t m, int32_t *outIdx, float *out3)
{
  std::vector<float> cdf(weights, weights + n);
  const float sum = Distribution1D_create(n, cdf.data());
  for (int64_t i = 0; i < m; i++) {
    const Sample1D r = Distribution1D_sample(n, cdf.data(), 0, s[i]);
    outIdx[i] = r.idx;
    out3[3 * i] = r.prob;
    out3[3 * i + 1] = r.pdf;
    out3[3 * i + 2] = r.frac;
  }
  return sum;
}

// lightSample of light `light` of the scene's light list ([geometry lights,
// world lights]) at point P with normals Ng, Ns for n sample pairs s2:
// out8 = dir (3), dist, weight (3), pdf per sample
void orc_light_samples(void *scene, int light, const float *P, const float *Ng, const float *Ns, const float *s2,
    int64_t n, float *out8)
{
  const Scene &sc = ((OrcSceneHandle *)scene)->scene;
  for (int64_t i = 0; i < n; i++)
    lightSampleHook(sc, light, v3(P[0], P[1], P[2]), v3(Ng[0], Ng[1], Ng[2]), v3(Ns[0], Ns[1], Ns[2]),
        v2(s2[2 * i], s2[2 * i + 1]), out8 + 8 * i);
}

// the path tracer's light selection (NextEventEstimation.ispc:60-63) for
// each value ss: the light's index in the scene's light list and its
// selection probability; geometryLights selects the CDF with or without
// the emissive geometry (PathTracer.ispc:95-110)
void orc_select_light(void *scene, int geometryLights, const float *ss, int64_t n, int32_t *outIdx, float *outProb)
{
  const Scene &sc = ((OrcSceneHandle *)scene)->scene;
  const std::vector<float> &cdf = geometryLights ? sc.lightsCDF : sc.worldLightsCDF;
  const int first = geometryLights ? 0 : (int)sc.numGeoLights;
  for (int64_t i = 0; i < n; i++) {
    const Sample1D r = Distribution1D_sample((int)cdf.size(), cdf.data(), 0, ss[i]);
    outIdx[i] = first + r.idx;
    outProb[i] = r.prob;
  }
}
} // extern "C"
