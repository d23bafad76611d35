// Neural radiance cache MLP on the gfx950 matrix cores.
//
// Fully fused: one wave runs a tile of 64 queries through every layer, the
// activations staying in LDS between layers.  Every product is a
// v_mfma_f32_16x16x32_bf16 (16x16 tile, K = 32, bf16 operands, fp32
// accumulation).  Operand layout of that instruction, lane l = 16 g + r:
//   A (16 x 32): row r, k = 8 g .. 8 g + 7      (one 16-byte load)
//   B (32 x 16): column r, k = 8 g .. 8 g + 7   (one 16-byte load)
//   D (16 x 16): column r, rows 4 g .. 4 g + 3  (4 fp32 registers)
// so an operand is read where its K index is contiguous: weights
// row-major (out, in) for the forward pass, their transposes for the
// backward pass, activations query-major [q][n] as the B operand of the
// next layer and neuron-major [n][q] (K = queries) for the weight
// gradients.
//
// Network (include/ospray_hip/hip_nrc.h): 160 -> 64 -> 64 -> 64 -> 64 -> 3,
// ReLU, no biases (input 156 is a constant 1).  Loss: relative L2,
// mean over queries and channels of (y - t)^2 / (y^2 + 0.01), the
// denominator held constant.  Training writes one gradient row per tile;
// the Adam kernel sums the rows in tile order (deterministic).
#include <hip/hip_runtime.h>

#include "ospray_hip/hip_nrc.h"
#include "pt_math.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kIn = HPT_NRC_INPUTS, kH = HPT_NRC_HIDDEN, kOP = 16, kP = HPT_NRC_PARAMS;
// element offsets of W0..W4 in the weights, of W1^T..W4^T in the transposes
__constant__ const int kOff[5] = {0, 10240, 14336, 18432, 22528};
__constant__ const int kOffT[4] = {0, 4096, 8192, 12288};
constexpr int kS = 72; // LDS row stride (bf16) of 64-wide rows: 144 B, 16-byte aligned
constexpr int kDyS = 40; // row stride of the query-major output gradient (32 used)
constexpr float kLossEps = 0.01f;

__device__ __forceinline__ f32x4 mma(const bf16x8 &a, const bf16x8 &b, const f32x4 &c)
{
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ bf16x8 ld8(const __bf16 *p) { return *reinterpret_cast<const bf16x8 *>(p); }
__device__ __forceinline__ bf16x8 zero8()
{
  bf16x8 z;
  for (int i = 0; i < 8; i++)
    z[i] = (__bf16)0.f;
  return z;
}
__device__ __forceinline__ f32x4 zero4() { return f32x4{0.f, 0.f, 0.f, 0.f}; }

// layer 0: acc[nt][qt] = W0 (64 x 160) x X^T for the tile's 64 queries
// read from the feature rows in HBM (zeros past n)
__device__ __forceinline__ void layer0(const __bf16 *W, const __bf16 *X, int q0, int n, f32x4 acc[4][4], int r, int g)
{
  for (int nt = 0; nt < 4; nt++)
    for (int qt = 0; qt < 4; qt++)
      acc[nt][qt] = zero4();
  for (int kb = 0; kb < kIn / 32; kb++) {
    bf16x8 b[4];
    for (int qt = 0; qt < 4; qt++) {
      const int q = q0 + qt * 16 + r;
      b[qt] = q < n ? ld8(X + (size_t)q * kIn + kb * 32 + g * 8) : zero8();
    }
    for (int nt = 0; nt < 4; nt++) {
      const bf16x8 a = ld8(W + (nt * 16 + r) * kIn + kb * 32 + g * 8);
      for (int qt = 0; qt < 4; qt++)
        acc[nt][qt] = mma(a, b[qt], acc[nt][qt]);
    }
  }
}

// hidden layer: acc[nt][qt] = W (64 x 64) x act^T, act query-major in LDS
__device__ __forceinline__ void layer64(const __bf16 *W, const __bf16 *act, f32x4 acc[4][4], int r, int g)
{
  for (int nt = 0; nt < 4; nt++)
    for (int qt = 0; qt < 4; qt++)
      acc[nt][qt] = zero4();
  for (int kb = 0; kb < 2; kb++) {
    bf16x8 b[4];
    for (int qt = 0; qt < 4; qt++)
      b[qt] = ld8(act + (qt * 16 + r) * kS + kb * 32 + g * 8);
    for (int nt = 0; nt < 4; nt++) {
      const bf16x8 a = ld8(W + (nt * 16 + r) * kH + kb * 32 + g * 8);
      for (int qt = 0; qt < 4; qt++)
        acc[nt][qt] = mma(a, b[qt], acc[nt][qt]);
    }
  }
}

// relu(acc) in bf16 to act[q][n] and, when given, to hT[n][q]
__device__ __forceinline__ void storeAct(const f32x4 acc[4][4], __bf16 *act, __bf16 *hT, int r, int g)
{
  for (int nt = 0; nt < 4; nt++)
    for (int qt = 0; qt < 4; qt++) {
      const int q = qt * 16 + r, n0 = nt * 16 + 4 * g;
      bf16x4 v;
      for (int e = 0; e < 4; e++)
        v[e] = (__bf16)fmaxf(acc[nt][qt][e], 0.f);
      *reinterpret_cast<bf16x4 *>(act + q * kS + n0) = v;
      if (hT)
        for (int e = 0; e < 4; e++)
          hT[(n0 + e) * kS + q] = v[e];
    }
}

// output layer: o[qt] = W4 (16 x 64) x act^T; rows 0..2 are the outputs
__device__ __forceinline__ void layerOut(const __bf16 *W4, const __bf16 *act, f32x4 o[4], int r, int g)
{
  for (int qt = 0; qt < 4; qt++)
    o[qt] = zero4();
  for (int kb = 0; kb < 2; kb++) {
    const bf16x8 a = ld8(W4 + r * kH + kb * 32 + g * 8);
    for (int qt = 0; qt < 4; qt++)
      o[qt] = mma(a, ld8(act + (qt * 16 + r) * kS + kb * 32 + g * 8), o[qt]);
  }
}

__global__ __launch_bounds__(64) void k_nrc_infer(HptNrc N, const __bf16 *X, int n, float *out)
{
  __shared__ __attribute__((aligned(16))) __bf16 act[64 * kS];
  const int lane = threadIdx.x, r = lane & 15, g = lane >> 4;
  const __bf16 *W = reinterpret_cast<const __bf16 *>(N.w);
  for (int tile = blockIdx.x; tile * 64 < n; tile += gridDim.x) {
    const int q0 = tile * 64;
    f32x4 acc[4][4];
    layer0(W + kOff[0], X, q0, n, acc, r, g);
    __syncthreads();
    storeAct(acc, act, nullptr, r, g);
    __syncthreads();
    for (int l = 1; l < 4; l++) {
      layer64(W + kOff[l], act, acc, r, g);
      __syncthreads();
      storeAct(acc, act, nullptr, r, g);
      __syncthreads();
    }
    f32x4 o[4];
    layerOut(W + kOff[4], act, o, r, g);
    if (g == 0)
      for (int qt = 0; qt < 4; qt++) {
        const int q = q0 + qt * 16 + r;
        if (q < n) {
          out[3 * (size_t)q] = o[qt][0];
          out[3 * (size_t)q + 1] = o[qt][1];
          out[3 * (size_t)q + 2] = o[qt][2];
        }
      }
    __syncthreads();
  }
}

// one tile of 64 (features, target) pairs per block iteration: forward,
// relative-L2 gradient, backward; the tile's weight gradients go to row
// `tile` of N.partial, its loss to N.tileLoss[tile]
__global__ __launch_bounds__(64) void k_nrc_train(HptNrc N, const __bf16 *X, const float *T, int n)
{
  __shared__ __attribute__((aligned(16))) __bf16 xT[kIn * kS]; // [k][q]
  __shared__ __attribute__((aligned(16))) __bf16 hT[4][kH * kS]; // h1..h4, [n][q]
  __shared__ __attribute__((aligned(16))) __bf16 act[64 * kS]; // [q][n]; backward: dz [q][n]
  __shared__ __attribute__((aligned(16))) __bf16 dzT[kH * kS]; // [n][q]
  __shared__ __attribute__((aligned(16))) __bf16 dyT[kOP * kS]; // [c][q]
  __shared__ __attribute__((aligned(16))) __bf16 dy[64 * kDyS]; // [q][c], c < 32
  const int lane = threadIdx.x, r = lane & 15, g = lane >> 4;
  const __bf16 *W = reinterpret_cast<const __bf16 *>(N.w);
  const __bf16 *WT = reinterpret_cast<const __bf16 *>(N.wT);
  const float scale = 1.f / (3.f * (float)n);
  for (int tile = blockIdx.x; tile * 64 < n; tile += gridDim.x) {
    const int q0 = tile * 64;
    float *grad = N.partial + (size_t)tile * kP;
    // the tile's features, transposed to [k][q]
    {
      const int q = q0 + lane;
      for (int c = 0; c < kIn / 8; c++) {
        const bf16x8 v = q < n ? ld8(X + (size_t)q * kIn + c * 8) : zero8();
        for (int j = 0; j < 8; j++)
          xT[(c * 8 + j) * kS + lane] = v[j];
      }
    }
    // forward
    f32x4 acc[4][4];
    layer0(W + kOff[0], X, q0, n, acc, r, g);
    __syncthreads();
    storeAct(acc, act, hT[0], r, g);
    __syncthreads();
    for (int l = 1; l < 4; l++) {
      layer64(W + kOff[l], act, acc, r, g);
      __syncthreads();
      storeAct(acc, act, hT[l], r, g);
      __syncthreads();
    }
    f32x4 o[4];
    layerOut(W + kOff[4], act, o, r, g);
    // output gradient dy = 2 (y - t) / (y^2 + eps) / (3 n) in rows 0..2
    float loss = 0.f;
    for (int qt = 0; qt < 4; qt++) {
      const int q = qt * 16 + r, qg = q0 + q;
      bf16x4 dv;
      for (int e = 0; e < 4; e++) {
        float d = 0.f;
        if (g == 0 && e < 3 && qg < n) {
          const float y = o[qt][e], t = T[3 * (size_t)qg + e];
          const float den = y * y + kLossEps;
          d = 2.f * (y - t) / den * scale;
          loss += (y - t) * (y - t) / den * scale;
        }
        dv[e] = (__bf16)d;
        dyT[(4 * g + e) * kS + q] = dv[e];
        dy[q * kDyS + 16 + 4 * g + e] = (__bf16)0.f;
      }
      *reinterpret_cast<bf16x4 *>(dy + q * kDyS + 4 * g) = dv;
    }
    for (int o2 = 32; o2 > 0; o2 >>= 1)
      loss += __shfl_xor(loss, o2);
    if (lane == 0)
      N.tileLoss[tile] = loss;
    __syncthreads();
    // dW4 (16 x 64) = dy^T h4
    for (int jt = 0; jt < 4; jt++) {
      f32x4 a4 = zero4();
      for (int kb = 0; kb < 2; kb++)
        a4 = mma(ld8(dyT + r * kS + kb * 32 + g * 8), ld8(hT[3] + (jt * 16 + r) * kS + kb * 32 + g * 8), a4);
      for (int e = 0; e < 4; e++)
        grad[kOff[4] + (4 * g + e) * kH + jt * 16 + r] = a4[e];
    }
    // dz4 = (W4^T dy) * [h4 > 0]
    for (int nt = 0; nt < 4; nt++) {
      const bf16x8 a = ld8(WT + kOffT[3] + (nt * 16 + r) * 32 + g * 8);
      for (int qt = 0; qt < 4; qt++)
        acc[nt][qt] = mma(a, ld8(dy + (qt * 16 + r) * kDyS + g * 8), zero4());
    }
    for (int l = 3;; l--) {
      // mask with the layer's input activation, store dz both ways
      __syncthreads();
      for (int nt = 0; nt < 4; nt++)
        for (int qt = 0; qt < 4; qt++) {
          const int q = qt * 16 + r, n0 = nt * 16 + 4 * g;
          bf16x4 v;
          for (int e = 0; e < 4; e++) {
            const float h = (float)hT[l][(n0 + e) * kS + q];
            v[e] = (__bf16)(h > 0.f ? acc[nt][qt][e] : 0.f);
            dzT[(n0 + e) * kS + q] = v[e];
          }
          *reinterpret_cast<bf16x4 *>(act + q * kS + n0) = v;
        }
      __syncthreads();
      // dW_l (64 x K_in) = dz h_l^T, h_0 = x
      const int kin = l ? kH : kIn;
      const __bf16 *hin = l ? hT[l - 1] : xT;
      for (int jt = 0; jt < kin / 16; jt++) {
        bf16x8 b[2];
        for (int kb = 0; kb < 2; kb++)
          b[kb] = ld8(hin + (jt * 16 + r) * kS + kb * 32 + g * 8);
        for (int nt = 0; nt < 4; nt++) {
          f32x4 a4 = zero4();
          for (int kb = 0; kb < 2; kb++)
            a4 = mma(ld8(dzT + (nt * 16 + r) * kS + kb * 32 + g * 8), b[kb], a4);
          for (int e = 0; e < 4; e++)
            grad[kOff[l] + (nt * 16 + 4 * g + e) * kin + jt * 16 + r] = a4[e];
        }
      }
      if (l == 0)
        break;
      // dh_l = W_l^T dz, masked by h_l at the top of the next iteration
      for (int nt = 0; nt < 4; nt++)
        for (int qt = 0; qt < 4; qt++)
          acc[nt][qt] = zero4();
      for (int kb = 0; kb < 2; kb++) {
        bf16x8 b[4];
        for (int qt = 0; qt < 4; qt++)
          b[qt] = ld8(act + (qt * 16 + r) * kS + kb * 32 + g * 8);
        for (int nt = 0; nt < 4; nt++) {
          const bf16x8 a = ld8(WT + kOffT[l - 1] + (nt * 16 + r) * kH + kb * 32 + g * 8);
          for (int qt = 0; qt < 4; qt++)
            acc[nt][qt] = mma(a, b[qt], acc[nt][qt]);
        }
      }
    }
    __syncthreads();
  }
}

// parameter p of the flat layout -> (layer, row, column)
__device__ __forceinline__ void paramIndex(int p, int &l, int &row, int &col)
{
  l = p < 10240 ? 0 : (p < 14336 ? 1 : (p < 18432 ? 2 : (p < 22528 ? 3 : 4)));
  const int k = l ? kH : kIn;
  const int o = p - kOff[l];
  row = o / k;
  col = o - row * k;
}

__device__ __forceinline__ void writeWeight(const HptNrc &N, int p, float w)
{
  int l, row, col;
  paramIndex(p, l, row, col);
  if (l == 4 && row >= HPT_NRC_OUTPUTS)
    w = 0.f; // padded output rows stay zero
  N.master[p] = w;
  const __bf16 b = (__bf16)w;
  reinterpret_cast<__bf16 *>(N.w)[p] = b;
  if (l >= 1) // transposes: W_l^T [in][out], W4^T [64][32]
    reinterpret_cast<__bf16 *>(N.wT)[kOffT[l - 1] + col * (l == 4 ? 32 : kH) + row] = b;
}

__global__ __launch_bounds__(256) void k_nrc_init(HptNrc N, uint32_t seed)
{
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < kP; p += gridDim.x * blockDim.x) {
    int l, row, col;
    paramIndex(p, l, row, col);
    // PCG-style hash of (seed, p) -> uniform [0, 1)
    uint64_t s = (uint64_t)seed * 0x9E3779B97F4A7C15ull + (uint64_t)p * 0xD1B54A32D192ED03ull + 1442695040888963407ull;
    s ^= s >> 33;
    s *= 0xff51afd7ed558ccdull;
    s ^= s >> 33;
    s *= 0xc4ceb9fe1a85ec53ull;
    s ^= s >> 33;
    const float u = (float)(uint32_t)(s >> 40) * (1.f / 16777216.f);
    const float lim = sqrtf(6.f / (float)(l ? kH : kIn));
    N.adamM[p] = 0.f;
    N.adamV[p] = 0.f;
    writeWeight(N, p, (2.f * u - 1.f) * lim);
  }
  // W4^T columns 16..31: zero
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < kH * 16; i += gridDim.x * blockDim.x)
    reinterpret_cast<__bf16 *>(N.wT)[kOffT[3] + (i / 16) * 32 + 16 + (i & 15)] = (__bf16)0.f;
}

// 64 parameters per block, the tiles' gradient rows split over 4 groups of
// 64 threads (group k sums tiles k, k + 4, ... in order) and combined in a
// fixed order: deterministic
__global__ __launch_bounds__(256) void k_nrc_adam(HptNrc N, int tiles, float lr, float bc1, float bc2)
{
  __shared__ float part[4][64];
  const int col = threadIdx.x & 63, grp = threadIdx.x >> 6;
  for (int p0 = blockIdx.x * 64; p0 < kP; p0 += gridDim.x * 64) {
    const int p = p0 + col;
    float gs = 0.f;
    if (p < kP)
      for (int t = grp; t < tiles; t += 4)
        gs += N.partial[(size_t)t * kP + p];
    part[grp][col] = gs;
    __syncthreads();
    if (grp == 0 && p < kP) {
      const float g = (part[0][col] + part[1][col]) + (part[2][col] + part[3][col]);
      const float m = 0.9f * N.adamM[p] + 0.1f * g;
      const float v = 0.99f * N.adamV[p] + 0.01f * g * g;
      N.adamM[p] = m;
      N.adamV[p] = v;
      writeWeight(N, p, N.master[p] - lr * (m / bc1) / (sqrtf(v / bc2) + 1e-8f));
    }
    __syncthreads();
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    float L = 0.f;
    for (int t = 0; t < tiles; t++)
      L += N.tileLoss[t];
    N.loss[0] = L;
  }
}

// feature group j (0..19: 8 features) of a query whose raw inputs are
// position, direction, normal, albedo (raw12): groups 0..17 hold 4 bands
// x (sin, cos) of coordinate j / 3 (position mapped over [lo, hi], then
// direction); 18 and 19 the albedo, normal, mapped position, direction, 1.
// j may differ across lanes: the coordinate is picked by selects, so the
// raw inputs stay in registers.
__device__ __forceinline__ bf16x8 encodeGroup(const float (&raw12)[12], int j, const float lo[3], const float inv[3])
{
  const float px = (raw12[0] - lo[0]) * inv[0], py = (raw12[1] - lo[1]) * inv[1], pz = (raw12[2] - lo[2]) * inv[2];
  bf16x8 v;
  if (j < 18) {
    const int d = j / 3, h = j - 3 * d;
    float c = px;
    c = d == 1 ? py : c;
    c = d == 2 ? pz : c;
    c = d == 3 ? raw12[3] : c;
    c = d == 4 ? raw12[4] : c;
    c = d == 5 ? raw12[5] : c;
    // band 4h directly, the next three by the double-angle formulas
    // sin 2x = 2 sin x cos x, cos 2x = (cos x - sin x)(cos x + sin x)
    float sn, cs;
    hpt::fsincos((float)(1 << (4 * h)) * hpt::kPi * c, sn, cs);
    for (int b = 0; b < 4; b++) {
      v[2 * b] = (__bf16)sn;
      v[2 * b + 1] = (__bf16)cs;
      const float s2 = 2.f * sn * cs, c2 = (cs - sn) * (cs + sn);
      sn = s2;
      cs = c2;
    }
    return v;
  }
  const bool first = j == 18;
  const float t[8] = {first ? raw12[9] : pz, first ? raw12[10] : raw12[3], first ? raw12[11] : raw12[4],
      first ? raw12[6] : raw12[5], first ? raw12[7] : 1.f, first ? raw12[8] : 0.f, first ? px : 0.f,
      first ? py : 0.f};
  for (int b = 0; b < 8; b++)
    v[b] = (__bf16)t[b];
  return v;
}

__global__ __launch_bounds__(256) void k_nrc_encode(const float *raw, int n, float lx, float ly, float lz, float hx,
    float hy, float hz, __bf16 *X)
{
  const float lo[3] = {lx, ly, lz}, inv[3] = {1.f / (hx - lx), 1.f / (hy - ly), 1.f / (hz - lz)};
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
    float rq[12];
    for (int k = 0; k < 12; k++)
      rq[k] = raw[12 * (size_t)q + k];
    bf16x8 *out = reinterpret_cast<bf16x8 *>(X + (size_t)q * kIn);
    for (int j = 0; j < kIn / 8; j++)
      out[j] = encodeGroup(rq, j, lo, inv);
  }
}

// The path tracer's cache queries in one kernel: gather the queued vertex
// records (per path slot: P, thr.x | wo, thr.y | Ns, thr.z | albedo, -),
// encode them in registers, run the network with the weights staged once
// per block in LDS (8 waves, one 64-query tile each per iteration), and add
// throughput * max(prediction, 0) to the path's contribution.
constexpr int kQueryWaves = 8;
__global__ __launch_bounds__(64 * kQueryWaves) void k_nrc_query(HptNrc N, const float *rec16, const uint32_t *queue,
    int n, float lx, float ly, float lz, float hx, float hy, float hz, float *contrib4)
{
  __shared__ __attribute__((aligned(16))) __bf16 wl[kP];
  __shared__ __attribute__((aligned(16))) __bf16 actAll[kQueryWaves][64 * kS];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
  __bf16 *act = actAll[wave];
  {
    const bf16x8 *src = reinterpret_cast<const bf16x8 *>(N.w);
    bf16x8 *dst = reinterpret_cast<bf16x8 *>(wl);
    for (int i = threadIdx.x; i < kP / 8; i += blockDim.x)
      dst[i] = src[i];
  }
  __syncthreads();
  const float lo[3] = {lx, ly, lz}, inv[3] = {1.f / (hx - lx), 1.f / (hy - ly), 1.f / (hz - lz)};
  // uniform trip count over the block: every wave reaches every barrier
  for (int base = blockIdx.x * kQueryWaves * 64; base < n; base += gridDim.x * kQueryWaves * 64) {
    const int q0 = base + wave * 64;
    // raw inputs of this lane's 4 queries (tile rows qt * 16 + r)
    float raw[4][12];
    for (int qt = 0; qt < 4; qt++) {
      const int q = q0 + qt * 16 + r;
      if (q < n) {
        const float *rec = rec16 + 16 * (size_t)queue[q];
        for (int k = 0; k < 4; k++) {
          const float4 v = *reinterpret_cast<const float4 *>(rec + 4 * k);
          raw[qt][3 * k] = v.x;
          raw[qt][3 * k + 1] = v.y;
          raw[qt][3 * k + 2] = v.z;
        }
      } else {
        for (int k = 0; k < 12; k++)
          raw[qt][k] = 0.f;
      }
    }
    f32x4 acc[4][4];
    for (int nt = 0; nt < 4; nt++)
      for (int qt = 0; qt < 4; qt++)
        acc[nt][qt] = zero4();
    for (int kb = 0; kb < kIn / 32; kb++) {
      bf16x8 b[4];
      for (int qt = 0; qt < 4; qt++)
        b[qt] = q0 + qt * 16 + r < n ? encodeGroup(raw[qt], 4 * kb + g, lo, inv) : zero8();
      for (int nt = 0; nt < 4; nt++) {
        const bf16x8 a = ld8(wl + kOff[0] + (nt * 16 + r) * kIn + kb * 32 + g * 8);
        for (int qt = 0; qt < 4; qt++)
          acc[nt][qt] = mma(a, b[qt], acc[nt][qt]);
      }
    }
    __syncthreads();
    storeAct(acc, act, nullptr, r, g);
    __syncthreads();
    for (int l = 1; l < 4; l++) {
      layer64(wl + kOff[l], act, acc, r, g);
      __syncthreads();
      storeAct(acc, act, nullptr, r, g);
      __syncthreads();
    }
    f32x4 o[4];
    layerOut(wl + kOff[4], act, o, r, g);
    if (g == 0)
      for (int qt = 0; qt < 4; qt++) {
        const int q = q0 + qt * 16 + r;
        if (q < n) {
          const uint32_t p = queue[q];
          const float *rec = rec16 + 16 * (size_t)p;
          float *c = contrib4 + 4 * (size_t)p;
          for (int k = 0; k < 3; k++)
            c[k] = c[k] + rec[4 * k + 3] * fmaxf(o[qt][k], 0.f);
        }
      }
    __syncthreads();
  }
}

int gridFor(long long n, int per, int cap)
{
  const long long b = (n + per - 1) / per;
  return (int)(b < 1 ? 1 : (b > cap ? cap : b));
}

} // namespace

extern "C" {

hipError_t hpt_nrc_init(const HptNrc *N, uint32_t seed, hipStream_t s)
{
  hipLaunchKernelGGL(k_nrc_init, dim3(gridFor(kP, 256, 1024)), dim3(256), 0, s, *N, seed);
  return hipGetLastError();
}

hipError_t hpt_nrc_encode(const float *raw12, int n, const float *lo3, const float *hi3, uint16_t *features, hipStream_t s)
{
  if (n <= 0)
    return hipSuccess;
  hipLaunchKernelGGL(k_nrc_encode, dim3(gridFor(n, 256, 8192)), dim3(256), 0, s, raw12, n, lo3[0], lo3[1], lo3[2],
      hi3[0], hi3[1], hi3[2], reinterpret_cast<__bf16 *>(features));
  return hipGetLastError();
}

hipError_t hpt_nrc_query(const HptNrc *N, const float *rec16, const uint32_t *queue, int n, const float *lo3,
    const float *hi3, float *contrib4, hipStream_t s)
{
  if (n <= 0)
    return hipSuccess;
  hipLaunchKernelGGL(k_nrc_query, dim3(gridFor(n, 64 * kQueryWaves, 4096)), dim3(64 * kQueryWaves), 0, s, *N, rec16,
      queue, n, lo3[0], lo3[1], lo3[2], hi3[0], hi3[1], hi3[2], contrib4);
  return hipGetLastError();
}

hipError_t hpt_nrc_infer(const HptNrc *N, const uint16_t *features, int n, float *out3, hipStream_t s)
{
  if (n <= 0)
    return hipSuccess;
  hipLaunchKernelGGL(k_nrc_infer, dim3(gridFor(n, 64, 8192)), dim3(64), 0, s, *N,
      reinterpret_cast<const __bf16 *>(features), n, out3);
  return hipGetLastError();
}

hipError_t hpt_nrc_train(const HptNrc *N, const uint16_t *features, const float *target3, int n, hipStream_t s)
{
  if (n <= 0)
    return hipSuccess;
  if ((n + 63) / 64 > N->maxTiles)
    return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_nrc_train, dim3(gridFor(n, 64, N->maxTiles)), dim3(64), 0, s, *N,
      reinterpret_cast<const __bf16 *>(features), target3, n);
  return hipGetLastError();
}

hipError_t hpt_nrc_adam(const HptNrc *N, int n, float lr, int step, hipStream_t s)
{
  const int tiles = (n + 63) / 64;
  if (tiles <= 0 || tiles > N->maxTiles || step < 1)
    return hipErrorInvalidValue;
  const float bc1 = 1.f - powf(0.9f, (float)step), bc2 = 1.f - powf(0.99f, (float)step);
  hipLaunchKernelGGL(k_nrc_adam, dim3(gridFor(kP, 64, 1024)), dim3(256), 0, s, *N, tiles, lr, bc1, bc2);
  return hipGetLastError();
}

} // extern "C"
