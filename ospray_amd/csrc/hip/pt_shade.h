// Surface interaction of the hip path tracer (host/device): post-intersect,
// BSDFs, lights.  Each function follows the reference ISPC source cited on
// it; arithmetic order is kept so CPU and GPU agree bit for bit.
#pragma once

#include "pt_sampler.h"
#include "pt_trace.h"
#include "pt_types.h"

namespace hpt {

enum : uint32_t
{
  SPEC_REFL = 1u << 0,
  GLOSSY_REFL = 1u << 1,
  DIFF_REFL = 1u << 2,
  SPEC_TRANS = 1u << 3,
  GLOSSY_TRANS = 1u << 4,
  DIFF_TRANS = 1u << 5,
  S_SPECULAR = SPEC_REFL | SPEC_TRANS,
  S_REFLECTION = DIFF_REFL | GLOSSY_REFL | SPEC_REFL,
  S_TRANSMISSION = DIFF_TRANS | GLOSSY_TRANS | SPEC_TRANS,
  S_SMOOTH = DIFF_REFL | DIFF_TRANS | GLOSSY_REFL | GLOSSY_TRANS
}; // render/pathtracer/Scattering.ih:5-45

constexpr float kMaxRouletteContProb = 0.95f; // PathTracerDefines.ih:6
constexpr float kCosAngleMax = 0.99999988f; // DirectionalLightShared.h:14

struct DG
{
  vec3f P, Ng, Ns, dPds, dPdt;
  float colorW;
  float epsilon;
  float areaPDF; // of the model as a geometry light (GeometricModel.ih:58)
  int material;
  uint32_t primID, objID, instID;
};

HPT_HD vec3f cartesian(float phi, float sinTheta, float cosTheta)
{
  float s, c;
  fsincos(phi, s, c);
  return v3(c * sinTheta, s * sinTheta, cosTheta);
}
HPT_HD float cos2sin(float c) { return sqrtf(maxf(0.f, 1.f - c * c)); }
HPT_HD vec3f cosineSampleHemisphere(vec2f s) // math/sampling.ih:43-49
{
  const float phi = kTwoPi * s.x;
  const float cosTheta = sqrtf(s.y);
  const float sinTheta = sqrtf(1.0f - s.y);
  return cartesian(phi, sinTheta, cosTheta);
}
HPT_HD vec3f uniformSampleSphere(float radius, vec2f s) // sampling.ih:191-197
{
  const float phi = kTwoPi * s.x;
  const float cosTheta = radius * (1.f - 2.f * s.y);
  const float sinTheta = 2.f * radius * sqrtf(s.y * (1.f - s.y));
  return cartesian(phi, sinTheta, cosTheta);
}
HPT_HD vec3f uniformSampleCone(float cosAngle, vec2f s) // sampling.ih:92-97
{
  const float phi = kTwoPi * s.x;
  const float cosTheta = 1.0f - s.y * (1.0f - cosAngle);
  return cartesian(phi, cos2sin(cosTheta), cosTheta);
}
HPT_HD float rcp_safe(float f) { return rcp(fabsf(f) < 1e-8f ? copysignf(1e-8f, f) : f); }
HPT_HD vec3f reflect(vec3f I, vec3f N) { return (2.0f * dot(I, N)) * N - I; } // Optics.ih:11-20

HPT_HD bool mediumEq(const DMedium &a, const DMedium &b) // materials/Medium.ih:23-26
{
  return a.ior == b.ior && a.attenuation[0] == b.attenuation[0] && a.attenuation[1] == b.attenuation[1]
      && a.attenuation[2] == b.attenuation[2];
}

// postIntersect: common/World.ih:389-465, Instance.ih:44-105,
// GeometricModel.ih:46-76, Mesh.ispc:65-268, Spheres.ispc:24-42
HPT_HD void postIntersect(const DScene &S, vec3f rorg, vec3f rdir, const Hit &hit, DG &dg, bool tangents)
{
  const DInstance &I = S.instances[hit.inst];
  const DModel &M = S.models[I.modelOffset + hit.model];
  dg.primID = hit.prim;
  dg.areaPDF = M.areaPDF;
  dg.material = -1;
  dg.colorW = 1.f;
  dg.dPds = v3(1.f, 0.f, 0.f);
  dg.dPdt = v3(0.f, 1.f, 0.f);
  dg.P = rorg + hit.t * rdir;
  dg.epsilon = 0.f;
  dg.Ng = dg.Ns = hit.Ng;
  if (M.geomType == GEOM_TRIANGLES) {
    const uint32_t *ix = M.index + 3 * (size_t)hit.prim;
    const uint32_t i0 = ix[0], i1 = ix[1], i2 = ix[2];
    const float u = hit.u, v = hit.v;
    const vec3f uv = v3(1.0f - u - v, u, v);
    if (M.normal) {
      const vec3f a = ld3(M.normal + 3 * (size_t)i0), b = ld3(M.normal + 3 * (size_t)i1), c = ld3(M.normal + 3 * (size_t)i2);
      dg.Ns = uv.x * a + uv.y * b + uv.z * c;
    }
    const vec3f va = ld3(M.vertex + 3 * (size_t)i0), vb = ld3(M.vertex + 3 * (size_t)i1), vc = ld3(M.vertex + 3 * (size_t)i2);
    dg.epsilon = reduce_max(vmax(vmax(vabs(va), vabs(vb)), vabs(vc))) * kUlpEpsilon;
    if (tangents) {
      dg.dPds = vb - va;
      dg.dPdt = vc - va;
    }
  } else if (M.geomType == GEOM_QUADS) {
    const uint32_t *ix = M.index + 4 * (size_t)hit.prim;
    const uint32_t i0 = ix[0], i1 = ix[1], i2 = ix[2], i3 = ix[3];
    const float u = hit.u, v = hit.v;
    vec4f uv;
    if (i2 == i3)
      uv = {1.f - u - v, u, v, 0.f};
    else
      uv = {(1 - v) * (1 - u), (1 - v) * u, v * u, v * (1 - u)};
    if (M.normal) {
      const vec3f a = ld3(M.normal + 3 * (size_t)i0), b = ld3(M.normal + 3 * (size_t)i1),
                  c = ld3(M.normal + 3 * (size_t)i2), e = ld3(M.normal + 3 * (size_t)i3);
      dg.Ns = uv.x * a + uv.y * b + uv.z * c + uv.w * e;
    }
    const vec3f va = ld3(M.vertex + 3 * (size_t)i0), vb = ld3(M.vertex + 3 * (size_t)i1),
                vc = ld3(M.vertex + 3 * (size_t)i2), vd = ld3(M.vertex + 3 * (size_t)i3);
    dg.epsilon = reduce_max(vmax(vmax(vabs(va), vabs(vb)), vmax(vabs(vc), vabs(vd)))) * kUlpEpsilon;
    if (tangents) {
      dg.dPds = vb - va;
      dg.dPdt = vd - va;
    }
  } else {
    dg.epsilon = M.spheres[4 * (size_t)hit.prim + 3] * kUlpEpsilon;
  }
  dg.objID = M.userID == kInvalidID ? hit.model : M.userID;
  dg.material = M.material;
  if (M.invertNormals)
    dg.Ns = neg(dg.Ns);
  dg.instID = I.userID == kInvalidID ? hit.inst : I.userID;
  const linear3f rl = {{I.rcp[0], I.rcp[1], I.rcp[2]}, {I.rcp[3], I.rcp[4], I.rcp[5]}, {I.rcp[6], I.rcp[7], I.rcp[8]}};
  dg.Ns = mulT(rl, dg.Ns);
  dg.Ng = mulT(rl, dg.Ng);
  dg.epsilon *= I.epsScale;
  if (tangents) {
    const linear3f xl = {{I.xfm[0], I.xfm[1], I.xfm[2]}, {I.xfm[3], I.xfm[4], I.xfm[5]}, {I.xfm[6], I.xfm[7], I.xfm[8]}};
    dg.dPds = mul(xl, dg.dPds);
    dg.dPdt = mul(xl, dg.dPdt);
  }
  const float tEps = reduce_max(vabs(rdir)) * hit.t;
  dg.epsilon = maxf(dg.epsilon, maxf(reduce_max(vabs(dg.P)), tEps) * kUlpEpsilon);
  vec3f ffnng = normalize(dg.Ng);
  dg.Ng = ffnng;
  dg.Ns = normalize(dg.Ns);
  if (dot(rdir, dg.Ng) >= 0.f) {
    ffnng = neg(ffnng);
    dg.Ng = neg(dg.Ng);
  }
  if (dot(dg.Ng, dg.Ns) < 0.f)
    dg.Ns = neg(dg.Ns);
  dg.P = dg.P + dg.epsilon * ffnng;
}

// ---- BSDFs --------------------------------------------------------------------

struct BSDF
{
  int kind; // 1 OBJ (Lambert + Transmission + Specular), 2 conductor, 3 microfacet conductor, 4 dielectric
  uint32_t scatteringType;
  vec3f albedo;
  linear3f frame;
  float weight[3], importance[3], importanceSum;
  vec3f childAlbedo[3];
  // OBJ specular child (Specular.ih:14-19): its reflectivity R is
  // childAlbedo[2]; the normalised reflectivity and the sample weight are R
  // times these two factors; specNs is the Blinn exponent
  float specKsScale, specWeightScale, specNs;
  vec3f eta, k;
  const float *spec; // spectral conductor: eta[8] then k[8], else null
  vec2f alpha;
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
  vec3f weight, wi;
  float pdf;
  uint32_t type;
};

HPT_HD float fresnelConductor1(float cosI, float eta, float k) // Fresnel.ih:75-84
{
  const float tmp = sqr(eta) + sqr(k);
  const float Rpar = (tmp * sqr(cosI) - eta * (2.0f * cosI) + 1.f) * rcp(tmp * sqr(cosI) + eta * (2.0f * cosI) + 1.f);
  const float Rper = (tmp - 2.0f * eta * cosI + sqr(cosI)) * rcp(tmp + 2.0f * eta * cosI + sqr(cosI));
  return 0.5f * (Rpar + Rper);
}
// the CIE-to-linear-sRGB weights of the eight spectral samples at 430,
// 465, ..., 675 nm (math/spectrum.ispc:11-38)
constexpr float kSpectrumSRGB[3][8] = {
    {0.0598548f, -0.0234574f, -0.220138f, -0.238902f, 0.316327f, 0.738315f, 0.323302f, 0.0446981f},
    {-0.0567346f, -0.0160361f, 0.223861f, 0.531185f, 0.337221f, 0.0149718f, -0.0296053f, -0.00486239f},
    {0.420693f, 0.616597f, 0.0796766f, -0.0496266f, -0.0473149f, -0.0167536f, -0.00295686f, -0.000314818f}};

// the conductor's Fresnel term: per RGB channel (FresnelConductorRGBUniform,
// Fresnel.ih:131-141) or summed over the spectrum and clamped to [0, 1]
// (FresnelConductorSpectral, Fresnel.ih:83-90, 199-213)
// a rolled loop, so the rarely taken spectral sum holds few registers in
// the shading kernels
HPT_HD vec3f fresnelSpectral(const float *spec, float cosI)
{
  vec3f rgb = v3(0.f);
#pragma unroll 1
  for (int l = 0; l < 8; l++) {
    const float F = fresnelConductor1(cosI, spec[l], spec[8 + l]);
    rgb = rgb + F * v3(kSpectrumSRGB[0][l], kSpectrumSRGB[1][l], kSpectrumSRGB[2][l]);
  }
  return v3(clampf(rgb.x), clampf(rgb.y), clampf(rgb.z));
}
HPT_HD vec3f fresnelRGB(const BSDF &b, float cosI)
{
  if (b.spec)
    return fresnelSpectral(b.spec, cosI);
  return v3(fresnelConductor1(cosI, b.eta.x, b.k.x), fresnelConductor1(cosI, b.eta.y, b.k.y),
      fresnelConductor1(cosI, b.eta.z, b.k.z));
}
HPT_HD float fresnelDielectric(float cosI, float eta) // Fresnel.ih:19-40
{
  const float sqrcost = 1.0f - sqr(eta) * (1.0f - sqr(cosI));
  if (sqrcost < 0.0f)
    return 1.0f;
  const float cosT = sqrtf(sqrcost);
  const float Rper = (eta * cosI - cosT) * rcp(eta * cosI + cosT);
  const float Rpar = (cosI - eta * cosT) * rcp(cosI + eta * cosT);
  return 0.5f * (sqr(Rpar) + sqr(Rper));
}

// GGX: render/bsdfs/GGXDistribution.ih
HPT_HD float ggxEval(vec2f a, vec3f wh)
{
  const float cosTheta = wh.z;
  const float cosTheta2 = sqr(cosTheta);
  const float e = (sqr(wh.x / a.x) + sqr(wh.y / a.y)) / cosTheta2;
  return rcp(kPi * a.x * a.y * sqr(cosTheta2 * (1.f + e)));
}
HPT_HD float ggxLambda(vec2f a, vec3f wo)
{
  const float c2 = sqr(wo.z);
  const float invA2 = (sqr(wo.x * a.x) + sqr(wo.y * a.y)) / c2;
  return 0.5f * (-1.f + sqrtf(1.f + invA2));
}
HPT_HD float ggxG1(vec2f a, vec3f wo, float cosThetaOH)
{
  if (wo.z * cosThetaOH <= 0.f)
    return 0.f;
  return rcp(1.f + ggxLambda(a, wo));
}
HPT_HD float ggxG2(vec2f a, vec3f wo, vec3f wi, float cosOH, float cosIH)
{
  if (wo.z * cosOH <= 0.f || wi.z * cosIH <= 0.f)
    return 0.f;
  return rcp(1.f + ggxLambda(a, wo) + ggxLambda(a, wi));
}
HPT_HD float ggxEvalVisible(vec2f a, vec3f wh, vec3f wo, float cosOH, float &pdf)
{
  const float D = ggxEval(a, wh);
  pdf = ggxG1(a, wo, cosOH) * fabsf(cosOH) * D / fabsf(wo.z);
  return D;
}
HPT_HD vec3f ggxSampleVisible(vec2f a, vec3f wo, float &pdf, vec2f s)
{
  const vec3f V = normalize(v3(a.x * wo.x, a.y * wo.y, wo.z));
  const vec3f T1 = (V.z < 0.9999f) ? normalize(cross(V, v3(0.f, 0.f, 1.f))) : v3(1.f, 0.f, 0.f);
  const vec3f T2 = cross(T1, V);
  const float aa = 1.f / (1.f + V.z);
  const float r = sqrtf(s.x);
  const float phi = (s.y < aa) ? s.y / aa * kPi : kPi + (s.y - aa) / (1.f - aa) * kPi;
  float sinPhi, cosPhi;
  fsincos(phi, sinPhi, cosPhi);
  const float P1 = r * cosPhi;
  const float P2 = r * sinPhi * ((s.y < aa) ? 1.f : V.z);
  vec3f wh = P1 * T1 + P2 * T2 + sqrtf(maxf(0.f, 1.f - P1 * P1 - P2 * P2)) * V;
  wh = normalize(v3(a.x * wh.x, a.y * wh.y, maxf(0.f, wh.z)));
  const float cwh = dot(wo, wh);
  pdf = ggxG1(a, wo, cwh) * fabsf(cwh) * ggxEval(a, wh) / fabsf(wo.z);
  return wh;
}

// ---- the OBJ material's specular lobe: a normalised Blinn-Phong BRDF
// f = r (n+2) / (4 pi (2 - 2^(-n/2))) cos(HN)^n / cos(LN) with the Phong
// exponent converted to Blinn (x4), sampled by the power-cosine
// distribution of the half vector (bsdfs/Specular.ih:14-78,
// math/sampling.ih:70-87)
constexpr float kOneOverFourPi = 0.0795774715459476678844f;
constexpr float kOneOverTwoPi = 0.159154943091895335769f;
constexpr float kLnTwoOverTwo = 0.346573590279972654709f;

HPT_HD float powerCosineSampleHemispherePDF(float cosTheta, float n)
{
  return (n + 1.0f) * kOneOverTwoPi * fpow(cosTheta, n);
}

HPT_HD vec3f powerCosineSampleHemisphere(float n, vec2f s)
{
  const float phi = kTwoPi * s.x;
  const float cosTheta = fpow(s.y, 1.0f / (n + 1.0f));
  return cartesian(phi, cos2sin(cosTheta), cosTheta);
}

HPT_HD EvalRes specularEval(const BSDF &b, vec3f wo, vec3f wi) // Specular.ih:23-41
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
  res.value = (b.childAlbedo[2] * b.specKsScale) * fpow(cosThetaH, b.specNs);
  return res;
}

HPT_HD SampleRes specularSample(const BSDF &b, vec3f wo, vec2f s) // Specular.ih:43-66
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
  res.weight = (b.childAlbedo[2] * b.specWeightScale) * cosThetaOH;
  return res;
}

// Material_dispatch_getBSDF (render/MaterialDispatch.ispc:23-72)
// the lobes of an OBJ material: Lambert (kd), specular transmission
// (1 - d, tf) and the specular reflection (ks, ns) (OBJ.ispc:55-98,
// MultiBSDF.ih:11-40)
HPT_HD void objLobes(const DMaterial &m, float colorW, BSDF &b)
{
  b.kind = 1;
  b.albedo = v3(0.f);
  b.scatteringType = 0;
  b.importanceSum = 0.f;
  b.weight[0] = b.weight[1] = b.weight[2] = 0.f;
  b.importance[0] = b.importance[1] = b.importance[2] = 0.f;
  const float d = m.d * colorW;
  const vec3f Kd = ld3(m.Kd) * d * v3(1.f, 1.f, 1.f);
  if (reduce_max(Kd) > 0.0f) {
    b.childAlbedo[0] = Kd;
    b.albedo = b.albedo + Kd * 1.f;
    b.scatteringType |= DIFF_REFL;
    b.weight[0] = 1.f;
    b.importance[0] = luminance(Kd);
    b.importanceSum += b.importance[0];
  }
  const vec3f T = ld3(m.Tf) * d + v3(1.f - d);
  if (reduce_max(T) > 0.0f) {
    b.childAlbedo[1] = T;
    b.albedo = b.albedo + T * 1.f;
    b.scatteringType |= SPEC_TRANS;
    b.weight[1] = 1.f;
    b.importance[1] = luminance(T);
    b.importanceSum += b.importance[1];
  }
  // Specular_Constructor and its MultiBSDF_add as child 2 (OBJ.ispc:86-95)
  const vec3f Ks = d * ld3(m.Ks);
  if (reduce_max(Ks) > 0.0f) {
    b.specNs = m.Ns * 4.f;
    const float tmp = (b.specNs + 2.f) * rcp(2.f - fexp(b.specNs * -kLnTwoOverTwo));
    b.specKsScale = kOneOverFourPi * tmp;
    b.specWeightScale = 2.f * tmp * rcp(b.specNs + 1.f);
    b.childAlbedo[2] = Ks;
    b.albedo = b.albedo + Ks * 1.f;
    b.scatteringType |= GLOSSY_REFL;
    b.weight[2] = 1.f;
    b.importance[2] = luminance(Ks);
    b.importanceSum += b.importance[2];
  }
}

// the conductor lobe of a metal material (Metal.ispc:18-46)
HPT_HD void metalLobe(const DMaterial &m, BSDF &b)
{
  b.eta = ld3(m.eta);
  b.k = ld3(m.k);
  b.spec = m.spectral ? m.etaS : nullptr;
  // neither conductor Fresnel has an evalAvg (Fresnel.ih:307-319): Favg = 0,
  // so the albedo and the Kulla-Conty compensation term are zero
  b.albedo = v3(0.f);
  if (m.roughness == 0.0f) {
    b.kind = 2;
    b.scatteringType = SPEC_REFL;
  } else {
    b.kind = 3;
    b.scatteringType = GLOSSY_REFL;
    const float aspect = sqrtf(1.f - 0.9f * 0.f);
    b.alpha = v2(maxf(sqr(m.roughness) / aspect, 0.001f), maxf(sqr(m.roughness) * aspect, 0.001f));
  }
}

HPT_HD bool getBSDF(const DScene &S, const DG &dg, const DMedium &medium, BSDF &b)
{
  if (dg.material < 0)
    return false;
  const DMaterial &m = S.materials[dg.material];
  if (m.type == MAT_LUMINOUS) { // Luminous.ispc:15-32: Transmission, or no BSDF
    if (!(m.transparency > 0.0f))
      return false;
    b.kind = 5;
    b.albedo = v3(m.transparency);
    b.scatteringType = SPEC_TRANS;
    b.frame = frame(dg.Ns);
    return true;
  }
  if (m.type == MAT_OBJ) { // OBJ.ispc:29-98
    b.kind = 1;
    b.albedo = v3(0.f);
    b.scatteringType = 0;
    { // makeShadingFrame (render/Material.ih:25-37)
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
    objLobes(m, dg.colorW, b);
    return true;
  }
  if (m.type == MAT_METAL) { // Metal.ispc:18-46
    b.frame = frame(dg.Ns);
    metalLobe(m, b);
    return true;
  }
  // Glass.ispc:18-35 -> RobustDielectric (bsdfs/RobustDielectric.ih:162-185)
  const DMedium &inside = S.media[m.inside];
  const DMedium &outside = S.media[m.outside];
  b.kind = 4;
  b.etaRel = mediumEq(medium, outside) ? outside.ior * rcp(inside.ior) : inside.ior * rcp(outside.ior);
  b.frame = frame(dg.Ns);
  b.albedo = v3(1.f);
  b.scatteringType = S_SPECULAR;
  b.Ns = dg.Ns;
  b.Ng = dg.Ng;
  return true;
}

HPT_HD EvalRes lambertEval(const BSDF &b, vec3f albedo, vec3f wi) // Lambert.ih:11-19
{
  const float cosThetaI = maxf(dot(wi, b.frame.vz), 0.f);
  return {albedo * kOneOverPi * cosThetaI, cosThetaI * kOneOverPi};
}

// BSDF_dispatch_eval (bsdfs/BSDF.ispc:22-56)
HPT_HD EvalRes bsdfEval(const BSDF &b, vec3f wo, vec3f wi)
{
  EvalRes z = {v3(0.f), 0.f};
  if (b.kind == 1) {
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
    r.pdf = whPdf * rcp(4.f * fabsf(cosThetaOH));
    r.value = F * (D * G * rcp(4.f * cosThetaO)) + fms;
    return r;
  }
  return z;
}

// A smooth BSDF (OBJ Lambert + Specular, MicrofacetConductor) as far as bsdfEval reads
// it, rebuilt from the material and the shading normal: bsdfEval uses only
// frame.vz of the OBJ shading frame, which is Ns whichever tangents built it
// (Material.ih:25-37), and metal's frame(Ns).  colorW is 1: postIntersect
// has no vertex colours.
HPT_HD void getEvalBSDF(const DScene &S, int material, vec3f Ns, BSDF &b)
{
  const DMaterial &m = S.materials[material];
  b.frame = frame(Ns);
  if (m.type == MAT_OBJ)
    objLobes(m, 1.f, b);
  else
    metalLobe(m, b);
}

// BSDF_dispatch_sample (bsdfs/BSDF.ispc:61-112)
HPT_HD SampleRes bsdfSample(const BSDF &b, vec3f wo, vec2f s, float ss)
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
  if (b.kind == 1) { // MULTIBSDF_SAMPLE (MultiBSDF.ih:80-128)
    SampleRes res = zero;
    const float x = ss * b.importanceSum;
    int choice = 0;
    float prefixSum = 0.f;
    if (b.importanceSum == 0.f)
      return res;
    if (x < (prefixSum += b.importance[0])) {
      choice = 0;
      const vec3f localDir = cosineSampleHemisphere(s); // Lambert_sample (Lambert.ih:21-30)
      res.wi = mul(b.frame, localDir);
      res.pdf = localDir.z * kOneOverPi;
      res.type = DIFF_REFL;
      res.weight = b.childAlbedo[0];
      res.weight = res.weight * b.weight[0];
    } else if (x < (prefixSum += b.importance[1])) {
      choice = 1;
      res.wi = neg(wo); // Transmission_sample (Transmission.ih:11-20)
      res.pdf = kInf;
      res.type = SPEC_TRANS;
      res.weight = b.childAlbedo[1];
      res.weight = res.weight * b.weight[1];
    } else if (x < (prefixSum += b.importance[2])) {
      choice = 2;
      res = specularSample(b, wo, s);
      res.weight = res.weight * b.weight[2];
    }
    if (eq(res.weight, v3(0.f)) || res.pdf == 0.f)
      return zero;
    const float chosen = choice == 0 ? b.importance[0] : choice == 1 ? b.importance[1] : b.importance[2];
    if (res.type & S_SPECULAR) {
      res.weight = res.weight * (b.importanceSum * rcp(chosen));
      return res;
    }
    EvalRes e;
    e.value = res.weight * res.pdf;
    res.pdf *= chosen;
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
    res.pdf = whPdf * rcp(4.f * fabsf(cosThetaOH));
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
    const vec3f par = dot(incomingDir, shadingNormal) * shadingNormal;
    const vec3f dirR = (incomingDir - par) - par;
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
          const float cos2 = sqrtf(clampf(1.f - sin2 * sin2));
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
HPT_HD vec3f getTransparency(const DScene &S, const DG &dg, vec3f rdir, const DMedium &medium)
{
  if (dg.material < 0)
    return v3(0.f);
  const DMaterial &m = S.materials[dg.material];
  if (m.type == MAT_OBJ) { // OBJ.ispc:100-117
    const float d = m.d * dg.colorW;
    return ld3(m.Tf) * d + v3(1.f - d);
  }
  if (m.type == MAT_LUMINOUS) // Luminous.ispc:34-44
    return v3(m.transparency);
  if (m.type == MAT_GLASS) { // Glass.ispc:37-52
    const DMedium &inside = S.media[m.inside];
    const DMedium &outside = S.media[m.outside];
    const float eta = mediumEq(medium, outside) ? outside.ior * rcp(inside.ior) : inside.ior * rcp(outside.ior);
    const float cosThetaO = maxf(-dot(rdir, dg.Ns), 0.0f);
    return v3(1.0f - fresnelDielectric(cosThetaO, eta));
  }
  return v3(0.f);
}

// Glass_selectNextMedium (Glass.ispc:54-64); returns the new medium index
HPT_HD uint32_t selectNextMedium(const DScene &S, const DG &dg, uint32_t mediumIdx)
{
  if (dg.material < 0)
    return mediumIdx;
  const DMaterial &m = S.materials[dg.material];
  if (m.type != MAT_GLASS)
    return mediumIdx;
  return mediumEq(S.media[mediumIdx], S.media[m.outside]) ? (uint32_t)m.inside : (uint32_t)m.outside;
}

// ---- lights (lights/*.ispc) ------------------------------------------------------

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
  float S, k, n0z, n2z, x0, x1, y0, y1, z0;
};

HPT_HD SphericalQuad sphericalQuadCreate(vec3f q0, vec3f e0, vec3f e1, vec3f n, vec3f P) // SphericalQuadSampling.ih:15-78
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
  q.z0 = -fabsf(q.z0);
  q.x0 = dot(d, q.e0);
  q.y0 = dot(d, q.e1);
  q.x1 = q.x0 + e0l;
  q.y1 = q.y0 + e1l;
  const vec3f v00 = v3(q.x0, q.y0, q.z0), v01 = v3(q.x0, q.y1, q.z0), v10 = v3(q.x1, q.y0, q.z0),
              v11 = v3(q.x1, q.y1, q.z0);
  const vec3f n0 = normalize(cross(v00, v10));
  q.n0z = n0.z;
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

HPT_HD float signf(float x) { return (float)(0.f < x) - (float)(x < 0.f); }

HPT_HD vec3f sampleSphericalQuad(const SphericalQuad &q, vec2f s) // SphericalQuadSampling.ih:86-111
{
  const float au = s.x * q.S + q.k;
  float sinau, cosau;
  fsincos(au, sinau, cosau);
  const float fu = (cosau * q.n0z - q.n2z) * rcp(sinau);
  const float cu = clampf(rsqrt(fu * fu + sqr(q.n0z)) * signf(fu), -1.f, 1.f);
  const float x0 = -(cu * q.z0) * rsqrt(1.f - cu * cu);
  const float xu = clampf(x0, q.x0, q.x1);
  const float ds = sqrtf(sqr(xu) + sqr(q.z0));
  const float ds2 = sqr(ds);
  const float h0 = q.y0 * rsqrt(ds2 + sqr(q.y0));
  const float h1 = q.y1 * rsqrt(ds2 + sqr(q.y1));
  const float hv = s.y * (h1 - h0) + h0;
  const float hv2 = sqr(hv);
  const float yv = (hv2 < 1.f - 1e-4f) ? (hv * ds) * rsqrt(1.f - hv2) : q.y1;
  return xu * q.e0 + yv * q.e1 + q.z0 * q.z;
}

HPT_HD float solveQuadratic(float A, float B, float C) // SphericalQuadSampling.ih:113-135
{
  if (A == 0.f)
    return -C / B;
  float rad = B * B - 4.f * A * C;
  if (rad < 0.f)
    return 0.f;
  rad = sqrtf(rad);
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

HPT_HD float fuWarp(float r, float W00, float W01, float W10, float W11)
{
  const float A = W10 + W11 - W01 - W00;
  const float B = 2.f * (W01 + W00);
  if (A + B == 0.f)
    return r;
  const float C = -(W00 + W01 + W10 + W11) * r;
  return solveQuadratic(A, B, C);
}

HPT_HD float fvWarp(float r, float u, float W00, float W01, float W10, float W11)
{
  const float A = (1.f - u) * W01 + u * W11 - (1.f - u) * W00 - u * W10;
  const float B = 2.f * (1.f - u) * W00 + 2.f * u * W10;
  if (A + B == 0.f)
    return r;
  const float C = -((1.f - u) * W01 + u * W11 + (1.f - u) * W00 + u * W10) * r;
  return solveQuadratic(A, B, C);
}

HPT_HD vec3f cosineWeightedRNG(vec3f q0, vec3f e0, vec3f e1, vec3f P, vec3f Ng, vec2f s) // SphericalQuadSampling.ih:157-196
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

// Light_dispatch_sample: QuadLight.ispc:82-136, DirectionalLight.ispc:14-27,
// AmbientLight.ispc:17-38
HPT_HD vec2f uniformSampleTriangleUV(vec2f s) // math/sampling.ih:165-173
{
  const float su = sqrtf(s.x);
  return v2(s.y * su, (1.0f - s.y) * su);
}

// the vertices a, b, c, d of a mesh primitive as Mesh_sampleArea reads them
// (a triangle repeats its last vertex: index (x, y, z, z))
HPT_HD bool meshPrim(const float *vertex, const uint32_t *index, int geomType, uint32_t p, vec3f &a, vec3f &b,
    vec3f &c, vec3f &d)
{
  uint32_t i0, i1, i2, i3;
  bool quad = false;
  if (geomType == GEOM_TRIANGLES) {
    const uint32_t *q = index + 3 * (size_t)p;
    i0 = q[0], i1 = q[1], i2 = q[2], i3 = q[2];
  } else {
    const uint32_t *q = index + 4 * (size_t)p;
    i0 = q[0], i1 = q[1], i2 = q[2], i3 = q[3];
    quad = i2 != i3;
  }
  a = ld3(vertex + 3 * (size_t)i0);
  b = ld3(vertex + 3 * (size_t)i1);
  d = ld3(vertex + 3 * (size_t)i3);
  c = quad ? ld3(vertex + 3 * (size_t)i2) : d;
  return quad;
}

// world-space area of a primitive (Mesh_getAreas, geometry/Mesh.ispc:377-415;
// Spheres_getAreas, geometry/Spheres.ispc:76-103); xfm: columns vx, vy, vz
HPT_HD float primArea(
    const float *vertex, const uint32_t *index, const float *spheres, int geomType, uint32_t p, const linear3f &xl)
{
  if (geomType == GEOM_SPHERES) {
    const vec3f sc = v3(length(xl.vx), length(xl.vy), length(xl.vz));
    const float scale = 4.0f * kPi * sqr((sc.x + sc.y + sc.z) / 3.f);
    return scale * sqr(spheres[4 * (size_t)p + 3]);
  }
  vec3f a, b, c, d;
  const bool quad = meshPrim(vertex, index, geomType, p, a, b, c, d);
  const vec3f e1 = mul(xl, a - d);
  const vec3f e2 = mul(xl, b - d);
  float area = length(cross(e1, e2));
  if (quad) {
    const vec3f e3 = mul(xl, c - d);
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
// (geometry/Spheres.ispc:44-74) under instance I
HPT_HD SampleAreaRes sampleArea(const DModel &M, const DInstance &I, uint32_t p, vec2f s)
{
  SampleAreaRes res;
  const linear3f xl = {{I.xfm[0], I.xfm[1], I.xfm[2]}, {I.xfm[3], I.xfm[4], I.xfm[5]}, {I.xfm[6], I.xfm[7], I.xfm[8]}};
  const vec3f xp = v3(I.xfm[9], I.xfm[10], I.xfm[11]);
  const float xfmScale = maxf(fabsf(xl.vx.x), maxf(fabsf(xl.vy.y), fabsf(xl.vz.z)));
  if (M.geomType == GEOM_SPHERES) {
    const float *sp = M.spheres + 4 * (size_t)p;
    const vec3f localNormal = uniformSampleSphere(sp[3], s);
    const linear3f rl = {{I.rcp[0], I.rcp[1], I.rcp[2]}, {I.rcp[3], I.rcp[4], I.rcp[5]}, {I.rcp[6], I.rcp[7], I.rcp[8]}};
    res.pos = mul(xl, v3(sp[0], sp[1], sp[2]) + localNormal) + xp;
    res.normal = normalize(mulT(rl, localNormal)); // transposed(rcp_xfm.l) * n
    res.epsilon = sp[3] * kUlpEpsilon;
    res.epsilon *= xfmScale;
    return res;
  }
  vec3f a, b, c, d;
  const bool quad = meshPrim(M.vertex, M.index, M.geomType, p, a, b, c, d);
  res.epsilon = reduce_max(vmax(vmax(vabs(a), vabs(b)), vmax(vabs(c), vabs(d)))) * kUlpEpsilon;
  res.epsilon *= xfmScale;
  const vec3f e1 = mul(xl, a - d);
  const vec3f e2 = mul(xl, b - d);
  const vec3f m1 = cross(e1, e2);
  res.normal = m1;
  vec3f v0 = a, v1 = b, v2 = d;
  vec2f sp = s;
  if (quad) {
    const vec3f e3 = mul(xl, c - d);
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
  res.pos = mul(xl, localPos) + xp;
  return res;
}

HPT_HD LightSample lightSample(const DScene &S, const DLight &l, vec3f P, vec3f Ng, vec3f Ns, vec2f s)
{
  LightSample res;
  if (l.type == LIGHT_GEOMETRY) { // GeometryLight_sample (render/pathtracer/GeometryLight.ispc:60-123)
    const DModel &M = S.models[l.geoModel];
    const Sample1D smp = distribution1DSample(l.geoNumPrims, S.geoCDF + l.geoCdfOffset, 0, s.x);
    const SampleAreaRes as = sampleArea(M, S.instances[l.geoInst], (uint32_t)smp.idx, v2(smp.frac, s.y));
    const vec3f dir = as.pos - P;
    const float dist = length(dir);
    res.dir = dir / dist;
    const float rcosd = rcp(fabsf(dot(as.normal, res.dir)));
    res.pdf = l.geoPdf * sqr(dist) * rcosd;
    // calcEpsilon(P, t) (common/DifferentialGeometry.ih:48-51)
    const float epsP = maxf(reduce_max(vabs(as.pos)), dist) * kUlpEpsilon;
    res.dist = dist - maxf(as.epsilon, epsP) * rcosd;
    res.weight = ld3(S.materials[M.material].radiance) * rcp(res.pdf);
    return res;
  }
  if (l.type == LIGHT_QUAD) {
    const vec3f position = ld3(l.position), edge1 = ld3(l.edge1), edge2 = ld3(l.edge2), nn = ld3(l.nnormal);
    if (s.x == 0.0f && s.y == 0.0f) {
      const vec3f p = position + edge1 * 0.5f + edge2 * 0.5f;
      const vec3f dir = p - P;
      const float dist = length(dir);
      res.dir = dir / dist;
      res.dist = dist;
      const float cosd = dot(nn, res.dir);
      res.pdf = l.ppdf * sqr(dist) / fabsf(cosd);
      const float weight = cosd > 0.f ? rcp(res.pdf) : 0.f;
      res.weight = ld3(l.radiance) * weight;
      return res;
    }
    const SphericalQuad quad = sphericalQuadCreate(position, edge1, edge2, neg(nn), P);
    const vec3f cosW = cosineWeightedRNG(position, edge1, edge2, P, Ng, s);
    const vec3f dir = sampleSphericalQuad(quad, v2(cosW.x, cosW.y));
    const float dist = length(dir);
    res.dir = dir / dist;
    res.dist = dist;
    const float cosd = dot(nn, res.dir);
    res.pdf = quad.S == 0.f ? 0.f : cosW.z / quad.S;
    res.weight = (cosd > 0.f) && (res.pdf != 0.f) ? ld3(l.radiance) * rcp(res.pdf) : v3(0.f);
    return res;
  }
  if (l.type == LIGHT_DISTANT) {
    const linear3f fr = {ld3(l.frame), ld3(l.frame + 3), ld3(l.frame + 6)};
    res.dir = fr.vz;
    res.dist = kInf;
    res.pdf = l.pdf;
    if (l.cosAngle < kCosAngleMax)
      res.dir = mul(fr, uniformSampleCone(l.cosAngle, s));
    res.weight = ld3(l.irradiance);
    return res;
  }
  if (reduce_max(vabs(Ns)) > 0.f) {
    const vec3f localDir = cosineSampleHemisphere(s);
    res.dir = mul(frame(Ns), localDir);
    res.pdf = localDir.z * kOneOverPi;
  } else {
    res.dir = uniformSampleSphere(1.f, s);
    res.pdf = rcp(kFourPi * sqr(1.f));
  }
  res.dist = kInf;
  res.weight = ld3(l.radiance) * rcp(res.pdf);
  return res;
}

// Light_dispatch_eval: QuadLight.ispc:167-209, DirectionalLight.ispc:58-70,
// AmbientLight.ispc:40-55
HPT_HD LightEval lightEval(const DLight &l, vec3f P, vec3f Ns, vec3f dir, float minDist, float maxDist)
{
  LightEval res;
  res.radiance = v3(0.f);
  res.pdf = 0.f;
  if (l.type == LIGHT_QUAD) {
    const vec3f position = ld3(l.position), edge1 = ld3(l.edge1), edge2 = ld3(l.edge2), nn = ld3(l.nnormal);
    const float cosd = dot(nn, dir);
    if (cosd <= 0.f)
      return res;
    const vec3f c = position - P;
    const vec3f r = cross(c, dir);
    const float rcosd = rcp(cosd);
    const float u = dot(r, edge1) * rcosd;
    const float v = -dot(r, edge2) * rcosd;
    if (minf(u, v) < 0.f || maxf(u, v) * l.ppdf > 1.0f)
      return res;
    const float dist = dot(nn, c) * rcosd;
    if (dist <= minDist || dist > maxDist)
      return res;
    const SphericalQuad quad = sphericalQuadCreate(position, edge1, edge2, neg(nn), P);
    if (quad.S > 0.f) {
      res.radiance = ld3(l.radiance);
      res.pdf = rcp(quad.S);
    }
    return res;
  }
  if (l.type == LIGHT_DISTANT) {
    const vec3f vz = ld3(l.frame + 6);
    if (kInf <= maxDist && l.cosAngle < kCosAngleMax && dot(vz, dir) > l.cosAngle) {
      res.radiance = ld3(l.irradiance) * l.pdf;
      res.pdf = l.pdf;
    }
    return res;
  }
  res.radiance = kInf <= maxDist ? ld3(l.radiance) : v3(0.f);
  if (reduce_max(vabs(Ns)) > 0.f)
    res.pdf = maxf(dot(Ns, dir), 0.f) * kOneOverPi;
  else
    res.pdf = rcp(kFourPi * sqr(1.f));
  return res;
}

HPT_HD float misHeuristic(float pdf1, float pdf2) // PathTracerUtil.ih:30-46
{
  const float p = sqr(pdf1) * rcp(sqr(pdf1) + sqr(pdf2));
  return pdf1 > 1e17f ? 1.0f : p;
}

// evaluateGeometryLights (render/pathtracer/GeometryLight.ispc:22-56): the
// emission of a hit luminous surface, MIS-weighted against NEE; lastP,
// lastPdf, lastNLS describe the previous vertex
HPT_HD vec3f evaluateGeometryLights(const DScene &S, const DG &dg, vec3f rdir, vec3f thr, vec3f lastP, float lastPdf,
    uint32_t lastNLS)
{
  if (dg.material < 0)
    return v3(0.f);
  const DMaterial &m = S.materials[dg.material];
  const vec3f radiance = ld3(m.radiance);
  if (m.type != MAT_LUMINOUS || !(reduce_max(radiance) > 0.f))
    return v3(0.f);
  float misWeight = 1.0f;
  if (S.numGeoLights > 0 && lastNLS > 0) {
    const float lightSelectionPDF = (float)lastNLS / (float)S.numLights;
    const float cosd = dot(dg.Ng, rdir);
    const vec3f vdis = lastP - dg.P;
    const float lePdf = dg.areaPDF * dot(vdis, vdis) / fabsf(cosd);
    misWeight = misHeuristic(lastPdf, lightSelectionPDF * lePdf);
  }
  return thr * radiance * misWeight;
}

} // namespace hpt
