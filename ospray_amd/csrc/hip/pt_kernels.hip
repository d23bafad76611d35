// Wavefront path-tracing kernels for gfx950 and their C-ABI launch layer.
//
// One frame = for each pass of samples: generate -> (extend -> shade ->
// shadow -> resolve) x (maxPathLength+1) -> finish_pass; then fb_accumulate,
// variance (odd frames) and colour conversion.  This splits the reference's
// per-pixel megakernel (PathTracer_renderTask, PathTracer.ispc:218-249;
// samplePath, PathSampler.ispc:109-415) into coherent stages:
//   extend  closest-hit traversal only (LDS stack, lean registers);
//   shade   post-intersect, virtual lights, BSDF, NEE light samples (emits
//           shadow-ray records and the next extension ray, compacted with
//           one wave-wide ballot + atomic per queue);
//   shadow  occlusion with opaque early exit, then the exact closest-hit
//           transparency loop (TransparentShadow.ispc:18-85) when only
//           non-opaque surfaces are in the way;
//   resolve adds the NEE samples of a vertex in light-sample order, so the
//           float sums match the reference's sequential order.
// Queues live in HBM; counts are read by the kernels themselves, so no host
// synchronisation is needed between stages (grid-stride loops over a fixed
// resident grid).
#include <hip/hip_runtime.h>

#include <mutex>

#include "ospray_hip/hip_pt.h"
#include "pt_shade.h"
#include "pt_volume.h"

using namespace hpt;

namespace {

__device__ __forceinline__ int laneId() { return (int)__lane_id(); }

// wave-aggregated queue append: one atomic per wave
__device__ __forceinline__ int enqueue(int32_t *counter, bool pred)
{
  const unsigned long long mask = __ballot(pred);
  if (mask == 0)
    return -1;
  const int lane = laneId();
  const int leader = __ffsll((long long)mask) - 1;
  int base = 0;
  if (lane == leader)
    base = atomicAdd(counter, __popcll(mask));
  base = __shfl(base, leader);
  return pred ? base + __popcll(mask & ((1ull << lane) - 1ull)) : -1;
}

// wave-reduced counter add: one atomic per wave; every lane of the wave calls it
__device__ __forceinline__ void waveAdd(unsigned long long *counter, uint32_t v)
{
  unsigned long long s = v;
  for (int o = 32; o > 0; o >>= 1)
    s += __shfl_xor(s, o);
  if (laneId() == 0)
    atomicAdd(counter, s);
}

// Block-aggregated append of up to one entry per thread to an extension
// queue and of `nee` entries per thread to the NEE sample queue, with one
// 64-bit atomic per block and call on the counter pair (ext, nee): atomics
// on one counter serialise in the L2, so one per wave costs milliseconds
// per frame.  Every thread of the block calls it, the same number of times.
__device__ __forceinline__ void blockEnqueuePair(int32_t *pair, bool ext, uint32_t nee, int &extSlot, int &neeSlot)
{
  __shared__ unsigned long long waveBase[16];
  const int lane = laneId(), wave = (int)(threadIdx.x >> 6), nw = (int)((blockDim.x + 63) >> 6);
  const unsigned long long mask = __ballot(ext);
  uint32_t x = nee; // inclusive scan of nee over the wave
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o);
    if (lane >= o)
      x += y;
  }
  const uint32_t totalNee = __shfl(x, 63); // all lanes take part in the shuffle
  if (lane == 0)
    waveBase[wave] = (unsigned long long)__popcll(mask) | ((unsigned long long)totalNee << 32);
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long total = 0; // (ext, nee) counts packed: neither half overflows
    for (int w = 0; w < nw; w++)
      total += waveBase[w];
    unsigned long long run = total ? atomicAdd(reinterpret_cast<unsigned long long *>(pair), total) : 0ull;
    for (int w = 0; w < nw; w++) {
      const unsigned long long t = waveBase[w];
      waveBase[w] = run;
      run += t;
    }
  }
  __syncthreads();
  const unsigned long long b = waveBase[wave];
  extSlot = (int)(uint32_t)b + __popcll(mask & ((1ull << lane) - 1ull));
  neeSlot = (int)(uint32_t)(b >> 32) + (int)(x - nee);
  __syncthreads(); // waveBase is reused by the next call
}

__device__ __forceinline__ float4 ld4(const float *p) { return *reinterpret_cast<const float4 *>(p); }
__device__ __forceinline__ uint4 ld4u(const uint32_t *p) { return *reinterpret_cast<const uint4 *>(p); }
__device__ __forceinline__ void st4(float *p, float a, float b, float c, float d)
{
  *reinterpret_cast<float4 *>(p) = make_float4(a, b, c, d);
}
__device__ __forceinline__ void st4u(uint32_t *p, uint32_t a, uint32_t b, uint32_t c, uint32_t d)
{
  *reinterpret_cast<uint4 *>(p) = make_uint4(a, b, c, d);
}

__device__ __forceinline__ int gridStride() { return (int)(gridDim.x * blockDim.x); }
__device__ __forceinline__ int globalId() { return (int)(blockIdx.x * blockDim.x + threadIdx.x); }

// path slot <-> pixel: slots are 8x8 tiles, one wave per tile and sample
__device__ __forceinline__ void slotPixel(const DFrame &F, int q, int &ix, int &iy)
{
  const int tile = q >> 6, w = q & 63, sh = F.slotShift;
  ix = ((tile % F.slotTilesX) << sh) + (w & ((1 << sh) - 1));
  iy = (tile / F.slotTilesX) * (64 >> sh) + (w >> sh);
}
__device__ __forceinline__ int pixelSlot(const DFrame &F, int ix, int iy)
{
  const int sh = F.slotShift, th = 64 >> sh;
  return ((iy / th) * F.slotTilesX + (ix >> sh)) * 64 + (iy % th) * (1 << sh) + (ix & ((1 << sh) - 1));
}

// the frame index (accumID): the host's, or the device-resident one of a
// captured batch frame
__device__ __forceinline__ uint32_t frameAccumID(const DFrame &F) { return F.frameState ? *F.frameState : F.accumID; }

// does this frame accumulate into the variance buffer (FrameBuffer.cpp:
// 126-136): even frames draw a coin, odd frames are the other half
__device__ __forceinline__ bool frameAccVar(const DFrame &F)
{
  if (!F.frameState)
    return F.accumulateVariance != 0;
  const uint32_t id = *F.frameState;
  const bool pickOdd = ((F.varianceCoins[id >> 6] >> ((id >> 1) & 31u)) & 1u) != 0u;
  return ((id & 1u) == 0u) != pickOdd;
}

// the adaptive-accumulation threshold of this frame (0: every task)
__device__ __forceinline__ float frameAdaptive(const DFrame &F)
{
  if (!F.frameState)
    return F.adaptiveThreshold;
  return *F.frameState >= (uint32_t)F.minAdaptiveFrames ? F.errorThreshold : 0.f;
}

// adaptive accumulation: is the 8x8 task of pixel (ix, iy) rendered this
// frame (LocalFB.cpp:238-261)?  threshold 0: every task is
__device__ __forceinline__ bool taskActive(const DFrame &F, float threshold, int ix, int iy)
{
  return threshold == 0.f || F.taskVariance[(iy >> 3) * F.tilesX + (ix >> 3)] > threshold;
}

__device__ __forceinline__ uint32_t clz32(uint32_t x) { return x ? (uint32_t)__clz((int)x) : 32u; }

// the path's PCG32 volume sampler in P.rng (state lo, hi, stream)
__device__ __forceinline__ RandomSampler loadRng(const DPaths &P, uint32_t p)
{
  const uint4 r = *reinterpret_cast<const uint4 *>(P.rng + 4 * (size_t)p);
  RandomSampler rs;
  rs.state = (uint64_t)r.x | ((uint64_t)r.y << 32);
  rs.stream = r.z;
  return rs;
}
__device__ __forceinline__ void storeRng(const DPaths &P, uint32_t p, const RandomSampler &rs)
{
  *reinterpret_cast<uint4 *>(P.rng + 4 * (size_t)p) =
      make_uint4((uint32_t)rs.state, (uint32_t)(rs.state >> 32), rs.stream, 0u);
}

// pixel filter (PixelFilterDispatch.ispc:7-19, LUTPixelFilter.ispc:9-15,
// Distribution2D.ispc:9-25); cdfX / cdfY: the LUT's CDFs (n x n, n)
__device__ __forceinline__ vec2f pixelFilterSample(const DFrame &F, vec2f s, const float *cdfX, const float *cdfY)
{
  if (F.filterType == 0)
    return v2(0.f, 0.f);
  if (F.filterType == 1)
    return v2((s.x + -0.5f) * F.filterWidth, (s.y + -0.5f) * F.filterWidth);
  const int n = F.lutSize;
  const Sample1D sy = distribution1DSample(n, cdfY, 0, s.y);
  const Sample1D sx = distribution1DSample(n, cdfX, sy.idx * n, s.x);
  const float rcpN = 1.f / (float)n;
  const vec2f uv = v2(((float)sx.idx + sx.frac) * rcpN, ((float)sy.idx + sy.frac) * rcpN);
  return v2((uv.x + -0.5f) * F.filterWidth, (uv.y + -0.5f) * F.filterWidth);
}

} // namespace

// ---- kernels ----------------------------------------------------------------------

// Zeroes the counters whose bit is set in `mask`.  A cancelled frame (the
// host wrote its generation into *cancelFlag, or the latch is already set)
// latches *cancelLatch and also zeroes counter `keep` (the entry count of
// the extension queue being read), so the frame's remaining kernels find no
// work.  One wave: lane 0's reads are broadcast.
__global__ void k_reset(int32_t *counters, uint32_t mask, int32_t keep, const uint32_t *cancelFlag, uint32_t gen,
    uint32_t *cancelLatch)
{
  const int i = globalId();
  uint32_t cancelled = 0u;
  if (cancelLatch) {
    if (i == 0) {
      cancelled = __hip_atomic_load(cancelLatch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (!cancelled && cancelFlag)
        cancelled = __hip_atomic_load(cancelFlag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) == gen ? 1u : 0u;
    }
    cancelled = __shfl(cancelled, 0);
  }
  if (i < kNumCounters && (((mask >> i) & 1u) || (cancelled && i == keep)))
    counters[i] = 0;
  if (cancelled && i == 0)
    __hip_atomic_store(cancelLatch, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// the frame's cancellation latch (read by kernels that run once per pass)
__device__ __forceinline__ bool frameCancelled(const uint32_t *latch)
{
  return latch && __hip_atomic_load(latch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u;
}

__global__ void k_fill_u32(uint32_t *p, size_t n, uint32_t v)
{
  for (size_t i = globalId(); i < n; i += gridStride())
    p[i] = v;
}

// plain device copy, 16 B per lane and iteration (bandwidth and HBM-counter
// calibration)
__global__ __launch_bounds__(256) void k_copy(uint4 *dst, const uint4 *src, size_t n16)
{
  for (size_t i = globalId(); i < n16; i += gridStride())
    dst[i] = src[i];
}

// PathTracer_renderPixel up to the first samplePath iteration
// (PathTracer.ispc:75-170) for samples [passS0, passS0+passCount)
constexpr int kLutLds = 1024; // floats of pixel-filter LUT staged in LDS

__global__ __launch_bounds__(256) void k_generate(DFrame F, DPaths P)
{
  // a cancelled frame generates no rays (block-uniform: the LUT staging
  // below has a barrier)
  __shared__ uint32_t cancelled;
  if (threadIdx.x == 0)
    cancelled = frameCancelled(P.cancelLatch) ? 1u : 0u;
  __syncthreads();
  if (cancelled)
    return;
  // the filter LUT's CDFs in LDS: every camera sample runs two binary
  // searches over them
  __shared__ float lut[kLutLds];
  const float *cdfX = F.lutCdfX, *cdfY = F.lutCdfY;
  if (F.filterType == 2 && F.lutSize * (F.lutSize + 1) <= kLutLds) { // uniform over the grid
    const int nn = F.lutSize * F.lutSize;
    for (int k = threadIdx.x; k < nn + F.lutSize; k += blockDim.x)
      lut[k] = k < nn ? F.lutCdfX[k] : F.lutCdfY[k - nn];
    __syncthreads();
    cdfX = lut;
    cdfY = lut + nn;
  }
  const int total = F.passCount * F.numPixPadded;
  const uint32_t accumID = frameAccumID(F);
  const float adaptive = frameAdaptive(F);
  for (int i = globalId(); i < total; i += gridStride()) {
    const int sLocal = i / F.numPixPadded;
    const int q = i - sLocal * F.numPixPadded;
    int ix, iy;
    slotPixel(F, q, ix, iy);
    const bool valid = ix < F.width && iy < F.height && taskActive(F, adaptive, ix, iy);
    if (valid) {
      const uint32_t spp = (uint32_t)F.spp;
      const uint32_t s = (uint32_t)(F.passS0 + sLocal);
      uint32_t seed = 0x49c5f37eu;
      uint32_t sampleOffs = 0, sampleOffsLight = 0;
      if (F.targetFrames) {
        const uint32_t targetSpp = (uint32_t)F.targetFrames * spp;
        sampleOffs = OwenScramble4(interleaveZOrder16((uint32_t)ix, (uint32_t)iy), seed) * targetSpp;
        sampleOffsLight = sampleOffs * F.numFirstLightSamples;
      } else {
        seed ^= (uint32_t)F.width * (uint32_t)iy + (uint32_t)ix;
      }
      const uint32_t sampleID = accumID * spp + s;
      // LDSampler_init (random.ih:283-294); the index masks are frame-uniform
      LDSampler lds;
      lds.revIndex = reverseBits(sampleOffs + sampleID);
      lds.scramble = seed;
      lds.baseIndexLight = sampleOffsLight + sampleID * F.numFirstLightSamples;
      lds.idxMask = F.idxMask;
      lds.idxMaskLight = F.idxMaskLight;
      const vec2f center = v2(((float)ix + 0.5f) * (1.f / (float)F.width), ((float)iy + 0.5f) * (1.f / (float)F.height));
      const vec2f pixelSample = ldGetNext5Samples(lds);
      const vec2f pf = pixelFilterSample(F, pixelSample, cdfX, cdfY);
      const vec2f screen = v2(center.x + pf.x * (1.f / (float)F.width), center.y + pf.y * (1.f / (float)F.height));
      // PerspectiveCamera_initRay (PerspectiveCamera.ispc:16-64), default image region
      const float sx = (1.f - screen.x) * 0.f + screen.x * 1.f;
      const float sy = (1.f - screen.y) * 0.f + screen.y * 1.f;
      const vec3f dir = normalize(ld3(F.camDir00) + sx * ld3(F.camDu) + sy * ld3(F.camDv));
      const vec3f org = ld3(F.camOrg);
      st4(P.org + 4 * (size_t)i, org.x, org.y, org.z, F.nearClip);
      st4(P.dir + 4 * (size_t)i, dir.x, dir.y, dir.z, 0.f);
      st4(P.thr + 4 * (size_t)i, 1.f, 1.f, 1.f, 1.f);
      st4(P.contrib + 4 * (size_t)i, 0.f, 0.f, 0.f, kInf);
      st4u(P.ld + 4 * (size_t)i, lds.revIndex, lds.scramble, lds.baseIndexLight,
          FLAG_FIRST_BOUNCE_LIGHT | FLAG_STRAIGHT | FLAG_SPEC_TRANS_PATH);
      st4(P.lastP + 4 * (size_t)i, org.x, org.y, org.z, kInf);
      st4(P.lastNs + 4 * (size_t)i, dir.x, dir.y, dir.z, __uint_as_float(0u));
      st4u(P.ids + 4 * (size_t)i, kInvalidID, kInvalidID, kInvalidID, 0u);
      if (F.hasVolumes) { // RandomSampler_init (PathTracer.ispc:157)
        RandomSampler rs;
        randomSamplerInit(rs, seed, sampleOffs + sampleID);
        storeRng(P, (uint32_t)i, rs);
      }
      if (P.aux) {
        float *a = P.aux + 16 * (size_t)i;
        for (int k = 0; k < 16; k += 4)
          st4(a + k, 0.f, 0.f, 0.f, 0.f);
      }
    }
    if (adaptive == 0.f) {
      // every pixel of the image is queued: the compacted position follows
      // from the slot layout (whole 8-row tile rows above, full-width tiles
      // to the left, then row-major inside the tile)
      if (valid) {
        const int tx = ix >> 3, ty = iy >> 3, px = ix & 7, py = iy & 7;
        const int h = min(8, F.height - 8 * ty), wt = min(8, F.width - 8 * tx);
        const int pos = sLocal * F.width * F.height + 8 * ty * F.width + 8 * tx * h + py * wt + px;
        P.queue[0][pos] = (uint32_t)i;
      }
    } else {
      const int slot = enqueue(&P.counters[0], valid); // extension queue 0
      if (valid)
        P.queue[0][slot] = (uint32_t)i;
    }
  }
  if (adaptive == 0.f && globalId() == 0)
    P.counters[0] = F.passCount * F.width * F.height;
}

// Work distribution of the traversal kernels: each wave takes batches of
// `size` (64 unless the queue is short) consecutive queue entries
// (neighbouring pixels' rays, which share node fetches) with one atomic
// per batch, until the queue is exhausted.
__device__ __forceinline__ int nextBatch(int32_t *counter, int size = 64)
{
  int base = 0;
  if (laneId() == 0)
    base = atomicAdd(counter, size);
  return __shfl(base, 0);
}

#ifndef HPT_SPREAD_SHORT_QUEUES
#define HPT_SPREAD_SHORT_QUEUES 1
#endif

// Lanes of a wave that wait for a ray before a traversal round refills
// them: extension rays (long closest-hit traversals) and shadow rays
// (shorter, ended early by an occluder) are refilled at different levels
#ifndef HPT_EXTEND_REFILL
#define HPT_EXTEND_REFILL 16
#endif
#ifndef HPT_SHADOW_REFILL
#define HPT_SHADOW_REFILL 48
#endif
constexpr int kExtendRefill = HPT_EXTEND_REFILL, kShadowRefill = HPT_SHADOW_REFILL;

// Traversal of every entry of a queue with the lanes of each wave kept
// busy: a lane whose ray has ended waits; once at least REFILL lanes
// wait (any number once the whole wave waits), the rays that ended are
// finished together and the waiting lanes take the next entries of the
// wave's current batch (64 consecutive queue entries, fetched with one
// atomic and read ahead with one coalesced load) before the next
// traversal round.  load(entry, org, dir, t0, tfar) gives an entry's ray;
// finish(ended, entry, best, occluded, state) runs on every lane of the
// wave (`ended` on the lanes whose ray ended), so it may use wave-wide
// appends.  A queue shorter than 64 entries per wave of the grid (the
// last bounces' few paths) is spread over every wave instead: each wave
// takes one batch of ceil(n / waves) entries and no more, so the few rays
// run in many waves side by side rather than 64 to a wave (where each
// round waits for the slowest of 64 node fetches and the longest of 64
// rays sets the wave's time).  Each ray's traversal is the one traceRay
// runs: the result does not depend on which rays share a wave.
template <int MODE, bool STATS, bool RECORD, int TIES, int REFILL, class Load, class Finish>
__device__ __forceinline__ void traceQueue(const DScene &S, const uint32_t *queue, int n, int32_t *fetch, uint32_t *lds,
    uint32_t *ovf, TraceStats &st, Load load, Finish finish)
{
  const int wl = laneId();
  const unsigned long long ltMask = (1ull << wl) - 1ull;
  const int waves = (int)((gridDim.x * blockDim.x) >> 6);
  const bool spread = HPT_SPREAD_SHORT_QUEUES && n < 64 * waves;
  const int bCap = spread ? (n + waves - 1) / waves : 64; // entries per batch, 1..64
  int bBase = nextBatch(fetch, bCap), bNext = 0; // wave-uniform batch cursor
  uint32_t bEntry = wl < bCap && bBase + wl < n ? queue[bBase + wl] : 0u;
  TraceState s;
  s.cur = kTraceDone;
  s.post = 0;
  s.sp = 0;
  Hit best;
  uint32_t entry = 0;
  bool busy = false, ended = false, occluded = false;
  for (;;) {
    unsigned long long idle = __ballot(!busy);
    if (idle == ~0ull || (bBase < n && __popcll(idle) >= REFILL)) {
      if (__any(ended))
        finish(ended, entry, best, occluded, s);
      ended = false;
      while (idle != 0ull && bBase < n) {
        const int bSize = n - bBase < bCap ? n - bBase : bCap;
        const int avail = bSize - bNext;
        const int rank = __popcll(idle & ltMask);
        const int src = bNext + rank;
        const uint32_t e = __shfl(bEntry, src < 63 ? src : 63);
        if (!busy && rank < avail) {
          entry = e;
          busy = true;
          occluded = false;
          vec3f org, dir;
          float t0, tfar;
          load(entry, org, dir, t0, tfar);
          traceBegin(s, S, org, dir, t0, tfar, best);
        }
        const int waiting = __popcll(idle);
        bNext += waiting < avail ? waiting : avail;
        idle = __ballot(!busy);
        if (bNext == bSize) {
          // a spread queue: the first batches of all waves cover it
          bBase = spread ? n : nextBatch(fetch);
          bNext = 0;
          bEntry = !spread && bBase + wl < n ? queue[bBase + wl] : 0u;
        }
      }
      if (idle == ~0ull) // the queue is exhausted and every lane waits
        break;
    }
    traceDescend<TIES, STATS>(s, S, best, lds, ovf, st);
    if (busy && (s.post == 0 ? s.cur == kTraceDone : traceLeaf<MODE, STATS, RECORD, TIES>(s, S, best, lds, ovf, st, occluded))) {
      busy = false;
      ended = true;
      s.cur = kTraceDone;
      s.post = 0;
      s.sp = 0;
    }
  }
}

// closest hit of every queued path (the rtcIntersectV of
// traceGeometryRayIntervals, World.ih:247-286)
template <bool STATS>
__global__ __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(5))) void k_extend(DScene S, DPaths P, int qin)
{
  __shared__ uint32_t lds[kLdsDepth * kTraceBlock];
  const int n = P.counters[2 * qin];
  if (n == 0) // an empty queue: no work fetch
    return;
  uint32_t *ovf = P.overflow + (size_t)globalId() * kThreadRow;
  TraceStats st = {0, 0};
  uint32_t rays = 0;
  traceQueue<MODE_CLOSEST, STATS, false, TIES_REPORT, kExtendRefill>(S, P.queue[qin], n, &P.counters[kCntFetchExtend], lds, ovf, st,
      [&](uint32_t p, vec3f &org, vec3f &dir, float &t0, float &tfar) {
        const float4 o = ld4(P.org + 4 * (size_t)p);
        const float4 d = ld4(P.dir + 4 * (size_t)p);
        org = v3(o.x, o.y, o.z);
        dir = v3(d.x, d.y, d.z);
        t0 = o.w;
        tfar = kInf;
      },
      [&](bool ended, uint32_t p, const Hit &h, bool, const TraceState &ts) {
        if (!ended)
          return;
        st4(P.hitA + 4 * (size_t)p, h.hit ? h.t : kInf, h.u, h.v, __uint_as_float(h.hit ? h.inst : kInvalidID));
        st4(P.hitB + 4 * (size_t)p, h.Ng.x, h.Ng.y, h.Ng.z, __uint_as_float(h.model));
        P.hitPrim[p] = h.prim | (h.half << 31);
        rays++;
        if (ts.tie) // rare: k_ties re-resolves the hit in FP64
          P.tieQueue[atomicAdd(&P.counters[kCntTies], 1)] = p;
      });
  if (STATS) {
    waveAdd(&P.stats[0], st.nodes);
    waveAdd(&P.stats[1], st.prims);
    waveAdd(&P.stats[2], rays);
  }
}

// FP64 re-resolution of the extension hits k_extend found near-tied
// (resolveNearTie): every hit within the window of the float closest hit
// gets its FP64 distance, the smallest (FP64 t, key) replaces the record.
template <bool STATS>
__global__ __launch_bounds__(kTraceBlock) void k_ties(DScene S, DPaths P)
{
  __shared__ uint32_t lds[kLdsDepth * kTraceBlock];
  const int n = P.counters[kCntTies];
  uint32_t *ovf = P.overflow + (size_t)globalId() * kThreadRow;
  TraceStats st = {0, 0};
  for (int i = globalId(); i < n; i += gridStride()) {
    const uint32_t p = P.tieQueue[i];
    const float4 o = ld4(P.org + 4 * (size_t)p);
    const float4 d = ld4(P.dir + 4 * (size_t)p);
    const float4 a = ld4(P.hitA + 4 * (size_t)p);
    const float4 b = ld4(P.hitB + 4 * (size_t)p);
    const uint32_t hp = P.hitPrim[p];
    Hit h;
    h.hit = true;
    h.t = a.x;
    h.u = a.y;
    h.v = a.z;
    h.inst = __float_as_uint(a.w);
    h.Ng = v3(b.x, b.y, b.z);
    h.model = __float_as_uint(b.w);
    h.prim = hp & 0x7fffffffu;
    h.half = hp >> 31;
    resolveNearTie<STATS>(S, v3(o.x, o.y, o.z), v3(d.x, d.y, d.z), o.w, kInf, h, lds, ovf, st);
    st4(P.hitA + 4 * (size_t)p, h.t, h.u, h.v, __uint_as_float(h.inst));
    st4(P.hitB + 4 * (size_t)p, h.Ng.x, h.Ng.y, h.Ng.z, __uint_as_float(h.model));
    P.hitPrim[p] = h.prim | (h.half << 31);
  }
  if (STATS) {
    waveAdd(&P.stats[0], st.nodes);
    waveAdd(&P.stats[1], st.prims);
  }
}

// volumeSampleFreePath (VolumeSampler.ispc:96-199) of every queued ray
// after its closest surface hit: delta tracking through the volume
// intervals up to the hit.  A collision replaces the hit record by a
// volume vertex (t, pair index, albedo).  The ray range of the clipping
// pass is [t0, inf) (no clipping geometry).
__global__ __launch_bounds__(256) void k_volume(DScene S, DPaths P, int qin, int stats)
{
  const int n = P.counters[2 * qin];
  VolCount cnt;
  for (int w0 = globalId() - laneId(); w0 < n; w0 += gridStride()) {
    const int i = w0 + laneId();
    if (i >= n)
      continue;
    const uint32_t p = P.queue[qin][i];
    const size_t p4 = 4 * (size_t)p;
    const float4 o4 = ld4(P.org + p4);
    const float4 d4 = ld4(P.dir + p4);
    const float4 hA = ld4(P.hitA + p4);
    const bool surface = __float_as_uint(hA.w) != kInvalidID;
    RandomSampler rs = loadRng(P, p);
    int pair;
    vec3f albedo;
    const float dist = volumeFreePath(S, v3(o4.x, o4.y, o4.z), v3(d4.x, d4.y, d4.z), o4.w, surface ? hA.x : kInf,
        surface, o4.w, kInf, rs, pair, albedo, cnt);
    storeRng(P, p, rs);
    if (dist < kInf) {
      st4(P.hitA + p4, dist, 0.f, 0.f, __uint_as_float(kVolumeHit | (uint32_t)pair));
      st4(P.hitB + p4, albedo.x, albedo.y, albedo.z, 0.f);
    }
  }
  if (stats) {
    waveAdd(&P.stats[9], cnt.samples);
    waveAdd(&P.stats[10], cnt.cells);
    waveAdd(&P.stats[11], cnt.tracks);
  }
}

// one iteration of samplePath (PathSampler.ispc:143-407) for every queued
// path.  At a smooth vertex it stores the NEE vertex record and appends the
// vertex's light-sample slots to the NEE sample queue; k_nee evaluates them
// (NextEventEstimation.ispc:25-160).  The loop runs over whole waves so the
// appends can scan across the wave.  Every instance holds 3 waves per SIMD:
// the volume instances (whose transmittance walks would take 215
// registers) spill the registers of their cold branches instead of
// dropping to 2 waves.
template <bool VOL, bool NRC>
__global__ __launch_bounds__(256, 3) void k_shade(DScene S, DFrame F, DPaths P, int qin)
{
  const int n = P.counters[2 * qin];
  const int qout = qin ^ 1;
  VolCount vcnt; // transmittance work of the virtual-light MIS weights
  // block-uniform iterations: the queue appends are aggregated per block
  for (int b0 = (int)(blockIdx.x * blockDim.x); b0 < n; b0 += (int)(gridDim.x * blockDim.x)) {
    const int i = b0 + (int)threadIdx.x;
    bool cont = false, nrcQuery = false, nrcTrain = false;
    uint32_t nNee = 0;
    uint32_t p = 0;
    if (i < n) {
      p = P.queue[qin][i];
      const size_t p4 = 4 * (size_t)p;
      const float4 o4 = ld4(P.org + p4);
      const float4 d4 = ld4(P.dir + p4);
      float4 thr4 = ld4(P.thr + p4);
      float4 con4 = ld4(P.contrib + p4);
      uint4 ld4v = ld4u(P.ld + p4);
      const float4 lp4 = ld4(P.lastP + p4);
      const float4 ln4 = ld4(P.lastNs + p4);
      const float4 hA = ld4(P.hitA + p4);
      const float4 hB = ld4(P.hitB + p4);
      const uint32_t hp = P.hitPrim[p];

      vec3f rorg = v3(o4.x, o4.y, o4.z);
      const vec3f rdir = v3(d4.x, d4.y, d4.z);
      vec3f thr = v3(thr4.x, thr4.y, thr4.z);
      float alpha = thr4.w;
      vec3f contribution = v3(con4.x, con4.y, con4.z);
      uint32_t flags = ld4v.w;
      const uint32_t depth = flags & FLAG_DEPTH_MASK;
      uint32_t scat = (flags >> FLAG_SCAT_SHIFT) & 0xffu;
      uint32_t medium = flags >> FLAG_MEDIUM_SHIFT;
      LDSampler lds;
      lds.revIndex = ld4v.x;
      lds.scramble = ld4v.y;
      lds.baseIndexLight = ld4v.z;
      lds.idxMask = F.idxMask;
      lds.idxMaskLight = F.idxMaskLight;

      Hit hit;
      hit.hit = __float_as_uint(hA.w) != kInvalidID;
      hit.t = hA.x;
      hit.u = hA.y;
      hit.v = hA.z;
      hit.inst = __float_as_uint(hA.w);
      hit.Ng = v3(hB.x, hB.y, hB.z);
      hit.model = __float_as_uint(hB.w);
      hit.prim = hp & 0x7fffffffu;
      hit.half = hp >> 31;
      const float rayT = hit.hit ? hit.t : kInf;
      // a volume collision of k_volume (PathSampler.ispc:178-193)
      const bool volHit = VOL && hit.hit && (hit.inst & kVolumeHit) != 0;
      constexpr bool volumes = VOL;
      RandomSampler rs;
      if (volumes)
        rs = loadRng(P, p);
      if (volHit)
        thr = thr * hit.Ng;

      const bool straight = (flags & FLAG_STRAIGHT) != 0;
      if (depth == 0)
        con4.w = F.projectedDepth ? rayT * dot(rdir, ld3(F.camDirection)) : rayT;

      // background (PathSampler.ispc:211-223)
      if (!hit.hit && straight) {
        alpha -= luminance(thr) * (1.0f - F.bgColor[3]);
        const vec3f bg3 = thr * v3(F.bgColor[0], F.bgColor[1], F.bgColor[2]);
        contribution = contribution + bg3;
      }
      // evaluateVirtualLights (VirtualLight.ispc:38-107), not at volume vertices
      if (!volHit && S.numLights > S.numGeoLights) {
        const vec3f lastP = v3(lp4.x, lp4.y, lp4.z);
        const vec3f lastNs = v3(ln4.x, ln4.y, ln4.z);
        const float lastPdf = lp4.w;
        const uint32_t lastNLS = __float_as_uint(ln4.w);
        vec3f L = v3(0.f);
        const float lower = distance(lastP, rorg);
        const float upper = !hit.hit ? kInf : distance(lastP, rorg + rayT * rdir);
        const float selectionPDF = (float)lastNLS / (float)S.numLights;
        for (int li = S.numGeoLights; li < S.numLights; li++) {
          const DLight &l = S.lights[li];
          if (!straight || l.visible) {
            const LightEval le = lightEval(l, lastP, lastNs, rdir, lower, upper);
            if (reduce_max(le.radiance) > 0.0f) {
              float mis = 1.0f;
              if (selectionPDF != 0.f) {
                // transmittance over the light's distance range
                // (VirtualLight.ispc:76-99)
                float T = 1.f;
                if (volumes)
                  T = volumeTransmittance(S, lastP, rdir, lower, upper, lower, upper, rs, vcnt);
                mis = misHeuristic(lastPdf * T, selectionPDF * le.pdf);
              }
              L = L + thr * le.radiance * mis;
            }
          }
        }
        contribution = contribution + L;
      }

      vec3f newOrg = rorg, newDir = rdir;
      float newLastPdf = lp4.w;
      vec3f newLastP = v3(lp4.x, lp4.y, lp4.z), newLastNs = v3(ln4.x, ln4.y, ln4.z);
      uint32_t newLastNLS = __float_as_uint(ln4.w);
      do {
        if (!hit.hit)
          break;
        // with geometry lights the emission at maxDepth is still evaluated
        // (PathSampler.ispc:249-255)
        if (depth >= (uint32_t)F.maxDepth && S.numGeoLights == 0)
          break;
        DG dg;
        float g = 0.f; // anisotropy of a volume vertex
        const int pair = (int)(hit.inst & ~kVolumeHit);
        if (volHit) { // PathSampler.ispc:51-68, VolumetricModel.ih:16-33
          const DVolInst vi = S.volInsts[pair];
          const DVolume &vol = S.volumes[vi.vol];
          g = vol.anisotropy;
          dg.P = rorg + rayT * rdir;
          const uint32_t instUser = S.instances[vi.inst].userID;
          dg.instID = instUser != kInvalidID ? instUser : (uint32_t)vi.inst;
          dg.Ng = dg.Ns = dg.dPds = dg.dPdt = v3(0.f);
          dg.colorW = 0.f;
          dg.primID = 0;
          dg.objID = vol.userID != kInvalidID ? vol.userID : (uint32_t)vi.local;
          dg.areaPDF = 0.f;
          dg.epsilon = 0.f;
          dg.material = -2 - pair;
        } else {
          postIntersect(S, rorg, rdir, hit, dg, true);
          contribution = contribution
              + evaluateGeometryLights(S, dg, rdir, thr, v3(lp4.x, lp4.y, lp4.z), lp4.w, __float_as_uint(ln4.w));
        }
        if (depth == 0) {
          P.ids[p4 + 0] = dg.primID;
          P.ids[p4 + 1] = dg.objID;
          P.ids[p4 + 2] = dg.instID;
          if (P.aux) {
            float *a = P.aux + 16 * (size_t)p;
            a[0] = dg.P.x;
            a[1] = dg.P.y;
            a[2] = dg.P.z;
            a[3] = dg.Ns.x;
            a[4] = dg.Ns.y;
            a[5] = dg.Ns.z;
          }
        }
        if (depth >= (uint32_t)F.maxDepth || scat >= (uint32_t)F.maxScatteringEvents)
          break;
        BSDF b;
        if (!volHit && !getBSDF(S, dg, S.media[medium], b))
          break;
        const bool suffixQuery = NRC && (flags & FLAG_NRC_TRAIN) && depth == (uint32_t)F.nrcBounce + 1u;
        if (NRC && (depth == (uint32_t)F.nrcBounce || suffixQuery)) {
          // the radiance cache at this vertex (before its NEE): position,
          // direction towards the previous vertex, shading normal, albedo
          // and the throughput.  The path ends here with a cache query;
          // a training slot runs on and becomes a training record (with
          // self-training, its own query one vertex later ends it).
          const bool query = suffixQuery || (p + F.nrcTrainOffset) % F.nrcTrainStride != 0;
          const vec3f alb = volHit ? hit.Ng : b.albedo;
          float *rec = (query ? P.nrcRec : P.nrcTrainRec) + 16 * (size_t)p;
          st4(rec, dg.P.x, dg.P.y, dg.P.z, thr.x);
          st4(rec + 4, -rdir.x, -rdir.y, -rdir.z, thr.y);
          st4(rec + 8, dg.Ns.x, dg.Ns.y, dg.Ns.z, thr.z);
          st4(rec + 12, alb.x, alb.y, alb.z, 0.f);
          if (query) {
            nrcQuery = true;
            break;
          }
          nrcTrain = true;
          // a training path from a volume vertex runs to its end: through a
          // medium whose albedo is near one, a target that includes the
          // cache's own estimate feeds the cache's bias back into it
          if (F.nrcSelfTrain && !volHit)
            flags |= FLAG_NRC_TRAIN;
          st4(P.nrcBefore + 4 * (size_t)p, contribution.x, contribution.y, contribution.z, 0.f);
        }
        ldNextGroup(lds);
        const vec3f wo = neg(rdir);
        uint32_t pvNLS = 0;
        // isSmooth (PathTracerUtil.ih:15-28)
        if (volHit ? hgIsSmooth(g) : (b.scatteringType & S_SMOOTH) != 0) {
          if (scat == 0 && P.aux) { // updateAuxilliaryData (PathTracerUtil.ih:48-60)
            float *a = P.aux + 16 * (size_t)p;
            const vec3f nn = thr * (volHit ? dg.Ns : b.frame.vz);
            const vec3f al = thr * (volHit ? hit.Ng : b.albedo);
            a[6] = nn.x;
            a[7] = nn.y;
            a[8] = nn.z;
            a[9] = al.x;
            a[10] = al.y;
            a[11] = al.z;
          }
          const uint32_t nls = (flags & FLAG_FIRST_BOUNCE_LIGHT) ? F.numFirstLightSamples : F.numIndirectLightSamples;
          pvNLS = nls;
          nNee = nls;
          if (nls) {
            // NEE vertex record: everything k_nee needs of this vertex
            st4(P.vtxP + p4, dg.P.x, dg.P.y, dg.P.z, dg.epsilon);
            st4(P.vtxNg + p4, dg.Ng.x, dg.Ng.y, dg.Ng.z, __int_as_float(dg.material));
            st4(P.vtxNs + p4, dg.Ns.x, dg.Ns.y, dg.Ns.z, __uint_as_float(medium));
            st4(P.vtxThr + p4, thr.x, thr.y, thr.z, __uint_as_float(nls));
            st4(P.vtxWo + p4, wo.x, wo.y, wo.z, 0.f);
            st4u(P.vtxLd + p4, lds.revIndex, lds.scramble, lds.baseIndexLight, 0u);
          }
          flags &= ~FLAG_FIRST_BOUNCE_LIGHT;
        }
        // sampleDirection (PathSampler.ispc:72-107)
        float ss;
        uint32_t uu3;
        const vec2f s = ldGetNext4Samples(lds, ss, uu3);
        SampleRes fs;
        if (volHit)
          fs = hgSample(g, wo, s);
        else {
          fs = bsdfSample(b, wo, s, ss);
          const bool sameSide = dot(fs.wi, dg.Ng) >= 0.f;
          const bool reflection = (fs.type & S_REFLECTION) != 0;
          if (sameSide != reflection)
            fs.weight = v3(0.0f);
        }
        if (reduce_max(fs.weight) <= 0.0f || fs.pdf <= 0.0f)
          break;
        thr = thr * fs.weight;
        if (depth >= (uint32_t)F.rouletteDepth) {
          const float rr = ldFinalizeDim3(lds, uu3);
          const float contProb = minf(reduce_max(thr), kMaxRouletteContProb);
          if (rr > contProb)
            break;
          thr = thr * rcp(contProb);
        }
        const DMedium &cm = S.media[medium];
        if (minf(minf(cm.attenuation[0], cm.attenuation[1]), cm.attenuation[2]) < 0.f) {
          const vec3f a = ld3(cm.attenuation) * rayT;
          thr = thr * v3(fexp(a.x), fexp(a.y), fexp(a.z));
        }
        vec3f rayOrg = dg.P;
        if (!volHit && (fs.type & S_TRANSMISSION)) {
          rayOrg = rayOrg - (2.0f * dg.epsilon) * dg.Ng;
          medium = selectNextMedium(S, dg, medium);
        }
        if ((fs.type & ~SPEC_TRANS) || volHit) {
          newLastP = dg.P;
          newLastNs = dg.Ns;
          newLastPdf = fs.pdf;
          newLastNLS = pvNLS;
        }
        if (!eq(rdir, fs.wi))
          flags &= ~FLAG_STRAIGHT;
        if (!(fs.type & SPEC_TRANS))
          flags &= ~FLAG_SPEC_TRANS_PATH;
        newOrg = rayOrg;
        newDir = fs.wi;
        const uint32_t nd = depth + 1;
        if (fs.type & S_SMOOTH)
          scat++;
        flags = (flags & ~(FLAG_DEPTH_MASK | (0xffu << FLAG_SCAT_SHIFT) | (0xfffu << FLAG_MEDIUM_SHIFT))) | nd
            | (scat << FLAG_SCAT_SHIFT) | (medium << FLAG_MEDIUM_SHIFT);
        cont = reduce_max(thr) > F.minContribution;
      } while (false);

      st4(P.thr + p4, thr.x, thr.y, thr.z, alpha);
      st4(P.contrib + p4, contribution.x, contribution.y, contribution.z, con4.w);
      if (volumes)
        storeRng(P, p, rs);
      if (cont) {
        st4(P.org + p4, newOrg.x, newOrg.y, newOrg.z, 0.f);
        st4(P.dir + p4, newDir.x, newDir.y, newDir.z, 0.f);
        st4u(P.ld + p4, lds.revIndex, lds.scramble, lds.baseIndexLight, flags);
        st4(P.lastP + p4, newLastP.x, newLastP.y, newLastP.z, newLastPdf);
        st4(P.lastNs + p4, newLastNs.x, newLastNs.y, newLastNs.z, __uint_as_float(newLastNLS));
      }
      P.neeN[p] = nNee;
    }
    if (NRC) {
      const int qs = enqueue(&P.counters[kCntNrcQuery], nrcQuery);
      if (nrcQuery)
        P.nrcQueryQueue[qs] = p;
      const int ts = enqueue(&P.counters[kCntNrcTrain], nrcTrain);
      if (nrcTrain)
        P.nrcTrainQueue[ts] = p;
    }
    // the continuing path to extension queue qout; the light-sample slots of
    // this vertex (records p*maxRec + [0, nNee)) to the NEE sample queue
    int so, first;
    blockEnqueuePair(&P.counters[2 * qout], cont, nNee, so, first);
    if (cont)
      P.queue[qout][so] = p;
    for (uint32_t k = 0; k < nNee; k++)
      P.neeSampleQueue[first + k] = p * (uint32_t)F.maxRec + k;
  }
  if (VOL && F.stats) {
    waveAdd(&P.stats[12], vcnt.samples);
    waveAdd(&P.stats[13], vcnt.cells);
    waveAdd(&P.stats[14], vcnt.tracks);
  }
}

// radiance cache: the queued vertex records as raw cache inputs
// (position, direction, normal, albedo)
__global__ __launch_bounds__(256) void k_nrc_gather(const float *records, const uint32_t *queue, int n, float *raw12)
{
  for (int i = globalId(); i < n; i += gridStride()) {
    const float *rec = records + 16 * (size_t)queue[i];
    float *o = raw12 + 12 * (size_t)i;
    for (int k = 0; k < 4; k++) {
      o[3 * k] = rec[4 * k];
      o[3 * k + 1] = rec[4 * k + 1];
      o[3 * k + 2] = rec[4 * k + 2];
    }
  }
}

// a query's path: contribution += throughput * max(prediction, 0)
__global__ __launch_bounds__(256) void k_nrc_apply(DPaths P, int n, const float *pred3)
{
  for (int i = globalId(); i < n; i += gridStride()) {
    const uint32_t p = P.nrcQueryQueue[i];
    const float *rec = P.nrcRec + 16 * (size_t)p;
    float *c = P.contrib + 4 * (size_t)p;
    for (int k = 0; k < 3; k++)
      c[k] = c[k] + rec[4 * k + 3] * fmaxf(pred3[3 * (size_t)i + k], 0.f);
  }
}

// a training record's target: the radiance its path gathered after the
// vertex, divided by the throughput at the vertex
__global__ __launch_bounds__(256) void k_nrc_targets(DPaths P, int n, float *target3)
{
  for (int i = globalId(); i < n; i += gridStride()) {
    const uint32_t p = P.nrcTrainQueue[i];
    const float *rec = P.nrcTrainRec + 16 * (size_t)p;
    const float *b = P.nrcBefore + 4 * (size_t)p;
    const float *c = P.contrib + 4 * (size_t)p;
    for (int k = 0; k < 3; k++) {
      const float thr = rec[4 * k + 3];
      const float t = thr > 0.f ? (c[k] - b[k]) / thr : 0.f;
      target3[3 * (size_t)i + k] = t == t ? fminf(fmaxf(t, 0.f), 1e4f) : 0.f;
    }
  }
}

// One NEE light sample per thread (NextEventEstimation.ispc:54-146): pick a
// light, sample it, evaluate the BSDF and write the shadow-ray record;
// records that can contribute go to the shadow queue.
constexpr int kStage = 2048;

// Writes the block's staged shadow-ray records to the shadow queue (one
// atomic for the block), stably partitioned by the light they sample (the
// light index modulo kLightBuckets): a wave of the occlusion pass then
// traces rays from neighbouring vertices towards one light, which share
// node fetches and branch alike.  Thread t owns the kStagePerThread
// entries from t * kStagePerThread; the partition offsets come from an
// exclusive scan of the per-(bucket, thread) counts in bucket-major order.
// Results do not depend on the queue order.  Called by every thread of the
// (256-thread) block.
constexpr int kLightBuckets = 8;
constexpr int kNeeBlock = 256;
constexpr int kStagePerThread = kStage / kNeeBlock;
static_assert(kStage % kNeeBlock == 0 && kLightBuckets <= 256, "stage layout");

__device__ __attribute__((noinline)) void flushShadowStage(DPaths &P, const uint32_t *stage, const uint8_t *stageKey,
    uint32_t cnt, uint32_t *table, uint32_t *partial, uint32_t &stageBase)
{
  // column t of table (entries b * 256 + t) is thread t's alone until the
  // scan: its per-bucket counts
  const int t = threadIdx.x;
  const uint32_t e0 = (uint32_t)t * kStagePerThread;
  for (int b = 0; b < kLightBuckets; b++)
    table[b * kNeeBlock + t] = 0;
  for (int k = 0; k < kStagePerThread; k++)
    if (e0 + k < cnt)
      table[stageKey[e0 + k] * kNeeBlock + t]++;
  __syncthreads();
  // exclusive scan in bucket-major order: thread t sums the table entries
  // [t * kLightBuckets, (t + 1) * kLightBuckets), the sums are scanned
  // within each wave by shuffles, then over the waves
  uint32_t run = 0;
  for (int k = 0; k < kLightBuckets; k++)
    run += table[t * kLightBuckets + k];
  uint32_t v = run;
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t u = __shfl_up(v, off);
    if (laneId() >= off)
      v += u;
  }
  if (laneId() == 63)
    partial[t >> 6] = v;
  __syncthreads();
  uint32_t before = v - run;
  for (int w = 0; w < (t >> 6); w++)
    before += partial[w];
  for (int k = 0; k < kLightBuckets; k++) {
    const uint32_t x = table[t * kLightBuckets + k];
    table[t * kLightBuckets + k] = before;
    before += x;
  }
  if (t == 0)
    stageBase = (uint32_t)atomicAdd(&P.counters[kCntShadow], (int)cnt);
  __syncthreads();
  // column t now holds thread t's first position in each bucket
  for (int k = 0; k < kStagePerThread; k++)
    if (e0 + k < cnt) {
      uint32_t &pos = table[stageKey[e0 + k] * kNeeBlock + t];
      P.shadowQueue[stageBase + pos] = stage[e0 + k];
      pos++;
    }
  __syncthreads();
}

__global__ __launch_bounds__(256) void k_nee(DScene S, DFrame F, DPaths P, int qin)
{
  __shared__ uint32_t stage[kStage];
  __shared__ uint8_t stageKey[kStage];
  __shared__ uint32_t table[kLightBuckets * kNeeBlock], partial[kNeeBlock / 64];
  __shared__ uint32_t stageCount, stageBase;
  if (threadIdx.x == 0)
    stageCount = 0;
  __syncthreads();
  const int n = P.counters[2 * (qin ^ 1) + 1];
  for (int b0 = blockIdx.x * blockDim.x; b0 < n; b0 += gridDim.x * blockDim.x) {
    const int i = b0 + (int)threadIdx.x;
    bool emit = false;
    uint32_t r = 0, light = 0;
    if (i < n) {
      r = P.neeSampleQueue[i];
      const uint32_t p = r / (uint32_t)F.maxRec;
      const uint32_t li = r - p * (uint32_t)F.maxRec;
      const size_t p4 = 4 * (size_t)p;
      const float4 vP = ld4(P.vtxP + p4);
      const float4 vNg = ld4(P.vtxNg + p4);
      const float4 vNs = ld4(P.vtxNs + p4);
      const float4 vThr = ld4(P.vtxThr + p4);
      const uint4 vLd = ld4u(P.vtxLd + p4);
      const vec3f Pv = v3(vP.x, vP.y, vP.z), Ng = v3(vNg.x, vNg.y, vNg.z), Ns = v3(vNs.x, vNs.y, vNs.z);
      const uint32_t nls = __float_as_uint(vThr.w);
      LDSampler lds;
      lds.revIndex = vLd.x;
      lds.scramble = vLd.y;
      lds.baseIndexLight = vLd.z;
      lds.idxMask = F.idxMask;
      lds.idxMaskLight = F.idxMaskLight;
      float ss;
      const vec2f s = ldGet3LightSamples(lds, li, nls > 1, ss);
      const Sample1D sel = distribution1DSample(S.numLights, S.lightsCDF, 0, ss);
      light = (uint32_t)sel.idx;
      const float selProb = sel.prob * (float)nls;
      LightSample ls = lightSample(S, S.lights[sel.idx], Pv, Ng, Ns, s);
      ls.weight = ls.weight / selProb;
      ls.pdf *= selProb;
      st4(P.recResult + 4 * (size_t)r, 0.f, 0.f, 0.f, 0.f);
      emit = !(reduce_max(ls.weight) <= 0.0f || ls.pdf <= 0.0f);
      EvalRes fe = {v3(0.f), 0.f};
      const int mat = __float_as_int(vNg.w); // <= -2: volume vertex of pair -2 - mat
      if (emit) {
        const float4 vWo = ld4(P.vtxWo + p4);
        if (mat <= -2) {
          fe = hgEval(S.volumes[S.volInsts[-2 - mat].vol].anisotropy, v3(vWo.x, vWo.y, vWo.z), ls.dir);
        } else {
          BSDF b;
          getEvalBSDF(S, mat, Ns, b);
          fe = bsdfEval(b, v3(vWo.x, vWo.y, vWo.z), ls.dir);
        }
        emit = !(reduce_max(fe.value) <= 0.0f);
      }
      if (emit) {
        vec3f org = Pv;
        if (mat > -2 && dot(Ng, ls.dir) < 0.f)
          org = org - (2.0f * vP.w) * Ng;
        const vec3f throughput = v3(vThr.x, vThr.y, vThr.z) * fe.value;
        const vec3f unshaded = throughput * ls.weight;
        st4(P.recOrg + 4 * (size_t)r, org.x, org.y, org.z, ls.dist);
        st4(P.recDir + 4 * (size_t)r, ls.dir.x, ls.dir.y, ls.dir.z, misHeuristic(ls.pdf, fe.pdf));
        st4(P.recContrib + 4 * (size_t)r, unshaded.x, unshaded.y, unshaded.z, vNs.w);
        if (S.numVolInsts > 0)
          st4(P.recVol + 4 * (size_t)r, ls.pdf, fe.pdf, 0.f, 0.f);
      }
    }
    {
      const unsigned long long mask = __ballot(emit);
      uint32_t wpos = 0;
      if (laneId() == 0 && mask)
        wpos = atomicAdd(&stageCount, (uint32_t)__popcll(mask));
      wpos = __shfl(wpos, 0);
      if (emit) {
        const uint32_t at = wpos + __popcll(mask & ((1ull << laneId()) - 1ull));
        stage[at] = r;
        stageKey[at] = (uint8_t)(light % (uint32_t)kLightBuckets);
      }
    }
    __syncthreads();
    // every wave reads the count before any wave can append again in the
    // next iteration, so the flush branch below is block-uniform
    const uint32_t cnt = stageCount;
    __syncthreads();
    if (cnt > (uint32_t)(kStage - blockDim.x)) {
      flushShadowStage(P, stage, stageKey, cnt, table, partial, stageBase);
      if (threadIdx.x == 0)
        stageCount = 0;
      __syncthreads();
    }
  }
  const uint32_t cnt = stageCount;
  if (cnt > 0)
    flushShadowStage(P, stage, stageKey, cnt, table, partial, stageBase);
}

// transparentShadow (TransparentShadow.ispc:18-85) per NEE record, first
// pass: occlusion traversal with early exit on an opaque primitive.
// Unoccluded records get their result; records that hit only non-opaque
// primitives go to the transparency queue (k_shadow_trans).
template <bool STATS>
__global__ __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(6))) void k_shadow(DScene S, DFrame F, DPaths P)
{
  __shared__ uint32_t lds[kLdsDepth * kTraceBlock];
  const int n = P.counters[kCntShadow];
  if (n == 0) // an empty queue: no work fetch
    return;
  uint32_t *ovf = P.overflow + (size_t)globalId() * kThreadRow;
  TraceStats st = {0, 0};
  uint32_t rays = 0;
  traceQueue<MODE_OCCLUSION, STATS, true, TIES_OFF, kShadowRefill>(S, P.shadowQueue, n, &P.counters[kCntFetchShadow], lds, ovf, st,
      [&](uint32_t r, vec3f &org, vec3f &dir, float &t0, float &tfar) {
        const float4 o4 = ld4(P.recOrg + 4 * (size_t)r);
        const float4 d4 = ld4(P.recDir + 4 * (size_t)r);
        org = v3(o4.x, o4.y, o4.z);
        dir = v3(d4.x, d4.y, d4.z);
        t0 = 0.f;
        tfar = o4.w;
      },
      [&](bool ended, uint32_t r, const Hit &, bool occluded, const TraceState &ts) {
        bool toTrans = false;
        const bool sawT = ended && ts.sawTransparent;
        if (ended) {
          rays++;
          if (!occluded) {
            if (sawT) {
              toTrans = true;
            } else {
              const float4 c4 = ld4(P.recContrib + 4 * (size_t)r);
              const vec3f lc = v3(c4.x, c4.y, c4.z);
              if (reduce_max(lc) > 0) {
                // with volumes, transmittance and MIS are applied in k_resolve
                const vec3f res = S.numVolInsts > 0 ? lc : lc * P.recDir[4 * (size_t)r + 3];
                st4(P.recResult + 4 * (size_t)r, res.x, res.y, res.z, 1.f);
              }
            }
          }
        }
        const int slot = enqueue(&P.counters[kCntTrans], toTrans);
        if (toTrans) {
          P.transQueue[slot] = r;
          if (slot < P.transCap) {
            // hand the hits the traversal recorded to the transparency pass
            const uint32_t nrec = *recordedCount(ovf);
            if (nrec <= (uint32_t)kTransRecHits) {
              const float4 *rec = recordedHits(ovf);
              float4 *dst = reinterpret_cast<float4 *>(P.transHits) + (size_t)slot * kTransRecHits;
              for (uint32_t k = 0; k < nrec; k++)
                dst[k] = rec[k];
            }
            P.transCount[slot] = nrec <= (uint32_t)kTransRecHits ? nrec : 0xffffffffu;
          }
        }
        // the recorded-hit count is 0 when a traversal starts: the rows are
        // cleared when allocated, and only rays that recorded hits reset it
        if (sawT)
          *recordedCount(ovf) = 0;
      });
  if (STATS) {
    waveAdd(&P.stats[3], st.nodes);
    waveAdd(&P.stats[4], st.prims);
    waveAdd(&P.stats[5], rays);
  }
}

// transparentShadow second pass: the exact closest-hit loop through
// non-opaque surfaces (TransparentShadow.ispc:36-84) for the records of the
// transparency queue
template <bool STATS>
__global__ __launch_bounds__(kTraceBlock) void k_shadow_trans(DScene S, DFrame F, DPaths P)
{
  __shared__ uint32_t lds[kLdsDepth * kTraceBlock];
  const int n = P.counters[kCntTrans];
  if (n == 0) // an empty queue: no work fetch
    return;
  uint32_t *ovf = P.overflow + (size_t)globalId() * kThreadRow;
  TraceStats st = {0, 0};
  uint32_t rays = 0;
  for (int base = nextBatch(&P.counters[kCntFetchTrans]); base < n; base = nextBatch(&P.counters[kCntFetchTrans])) {
    const int i = base + laneId();
    if (i >= n)
      continue;
    const uint32_t r = P.transQueue[i];
    const float4 o4 = ld4(P.recOrg + 4 * (size_t)r);
    const float4 d4 = ld4(P.recDir + 4 * (size_t)r);
    const float4 c4 = ld4(P.recContrib + 4 * (size_t)r);
    const vec3f org = v3(o4.x, o4.y, o4.z), dir = v3(d4.x, d4.y, d4.z);
    vec3f lc = v3(c4.x, c4.y, c4.z);
    uint32_t medium = __float_as_uint(c4.w);
    uint32_t maxDepth = (uint32_t)F.maxDepth;
    const float tOriginal = o4.w;
    float t0 = 0.f;
    Hit h;
    // the hits the occlusion pass recorded, if it recorded them all: the
    // closest hit in (t0, tOriginal] is the smallest recorded (t, instance,
    // model, primitive, half) with t > t0 (the first recorded on a full tie:
    // a sphere's entry before its exit), exactly what the traversal returns
    const uint32_t nrec = i < P.transCap ? P.transCount[i] : 0xffffffffu;
    const float4 *rec = reinterpret_cast<const float4 *>(P.transHits) + (size_t)i * kTransRecHits;
    for (;;) {
      if (nrec <= (uint32_t)kTransRecHits) {
        h.hit = false;
        float4 sel = make_float4(0.f, 0.f, 0.f, 0.f);
        for (int k = 0; k < (int)nrec; k++) {
          const float4 a = rec[k];
          if (!(a.x > t0))
            continue;
          const DPrim &Pk = S.prims[__float_as_uint(a.z)];
          const uint32_t prim = __float_as_uint(reinterpret_cast<const float4 *>(&Pk)[0].w);
          const uint32_t mk = __float_as_uint(reinterpret_cast<const float4 *>(&Pk)[1].w);
          const uint32_t half = (mk >> 28) == PRIM_QUAD1 ? 1u : 0u;
          if (better(a.x, __float_as_uint(a.y), mk & 0x0fffffffu, prim, half, h)) {
            h.hit = true;
            h.t = a.x;
            h.inst = __float_as_uint(a.y);
            h.model = mk & 0x0fffffffu;
            h.prim = prim;
            h.half = half;
            sel = a;
          }
        }
        if (h.hit)
          hitFromRecord(S, org, dir, sel, h);
      } else {
        bool dummy;
        traceRay<MODE_CLOSEST, STATS>(S, org, dir, t0, tOriginal, h, lds, ovf, dummy, st);
        rays++;
      }
      if (!h.hit)
        break;
      DG dg;
      postIntersect(S, org, dir, h, dg, false);
      const vec3f tr = getTransparency(S, dg, dir, S.media[medium]);
      lc = lc * tr;
      const DMedium &cm = S.media[medium];
      if (!(cm.attenuation[0] == 0.f && cm.attenuation[1] == 0.f && cm.attenuation[2] == 0.f)) {
        const vec3f a = ld3(cm.attenuation) * (h.t - t0);
        lc = lc * v3(fexp(a.x), fexp(a.y), fexp(a.z));
      }
      if (reduce_max(lc) <= F.minContribution || --maxDepth <= 0) {
        lc = v3(0.f);
        break;
      }
      medium = selectNextMedium(S, dg, medium);
      t0 = h.t + dg.epsilon;
    }
    if (reduce_max(lc) > 0) {
      const vec3f res = S.numVolInsts > 0 ? lc : lc * d4.w;
      st4(P.recResult + 4 * (size_t)r, res.x, res.y, res.z, 1.f);
      if (S.numVolInsts > 0)
        P.recVol[4 * (size_t)r + 2] = t0; // the last segment [t0, tOriginal]
    }
  }
  if (STATS) {
    waveAdd(&P.stats[6], st.nodes);
    waveAdd(&P.stats[7], st.prims);
    waveAdd(&P.stats[8], rays);
  }
}

// sum of the NEE samples of a vertex in light-sample order, added to the
// path contribution (NextEventEstimation.ispc:150-159, PathSampler.ispc:338).
// With volumes each visible sample is weighted by two independent
// transmittance estimates along its last shadow segment, one for the
// sample and one inside the MIS weight (NextEventEstimation.ispc:126-149),
// drawn from the path's sampler in sample order.
template <bool VOL>
__global__ __launch_bounds__(256) void k_resolve(DScene S, DFrame F, DPaths P, int qin)
{
  const int n = P.counters[2 * qin];
  VolCount vcnt;
  for (int i = globalId(); i < n; i += gridStride()) {
    const uint32_t p = P.queue[qin][i];
    const uint32_t nls = P.neeN[p];
    if (nls == 0)
      continue;
    vec3f L = v3(0.f);
    if (VOL) {
      RandomSampler rs = loadRng(P, p);
      for (uint32_t k = 0; k < nls; k++) {
        const size_t r4 = 4 * ((size_t)p * F.maxRec + k);
        const float4 r = ld4(P.recResult + r4);
        if (r.w == 0.f)
          continue;
        const float4 o4 = ld4(P.recOrg + r4);
        const float4 d4 = ld4(P.recDir + r4);
        const float4 v4 = ld4(P.recVol + r4);
        const vec3f org = v3(o4.x, o4.y, o4.z), dir = v3(d4.x, d4.y, d4.z);
        const float T = volumeTransmittance(S, org, dir, v4.z, o4.w, 0.f, o4.w, rs, vcnt);
        if (T > 0.f) {
          const float Tmis = volumeTransmittance(S, org, dir, v4.z, o4.w, 0.f, o4.w, rs, vcnt);
          L = L + (T * v3(r.x, r.y, r.z)) * misHeuristic(v4.x, v4.y * Tmis);
        }
      }
      storeRng(P, p, rs);
    } else {
      for (uint32_t k = 0; k < nls; k++) {
        const float4 r = ld4(P.recResult + 4 * ((size_t)p * F.maxRec + k));
        if (r.w != 0.f)
          L = L + v3(r.x, r.y, r.z);
      }
    }
    float *c = P.contrib + 4 * (size_t)p;
    c[0] = c[0] + L.x;
    c[1] = c[1] + L.y;
    c[2] = c[2] + L.z;
  }
  if (VOL && F.stats) {
    waveAdd(&P.stats[12], vcnt.samples);
    waveAdd(&P.stats[13], vcnt.cells);
    waveAdd(&P.stats[14], vcnt.tracks);
  }
}

// end of samplePath + ScreenSample_accumulate over the pass's samples
// (PathSampler.ispc:409-414, PathTracer.ispc:196-197, ScreenSample.ih:59-77)
__global__ __launch_bounds__(256) void k_finish_pass(DFrame F, DPaths P)
{
  if (frameCancelled(F.cancelLatch))
    return;
  const int npix = F.width * F.height;
  const float adaptive = frameAdaptive(F);
  for (int pix = globalId(); pix < npix; pix += gridStride()) {
    const int ix = pix % F.width, iy = pix / F.width;
    if (!taskActive(F, adaptive, ix, iy))
      continue;
    const int q = pixelSlot(F, ix, iy);
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    float aux[12];
    for (int k = 0; k < 12; k++)
      aux[k] = 0.f;
    float z = kInf;
    uint32_t ids[3] = {kInvalidID, kInvalidID, kInvalidID};
    if (F.passS0 > 0) {
      acc = ld4(F.accRGBA + 4 * (size_t)pix);
      if (F.accAux)
        for (int k = 0; k < 12; k++)
          aux[k] = F.accAux[12 * (size_t)pix + k];
      z = F.accZ[pix];
      for (int k = 0; k < 3; k++)
        ids[k] = F.accIDs[3 * (size_t)pix + k];
    }
    for (int sl = 0; sl < F.passCount; sl++) {
      const size_t p = (size_t)sl * F.numPixPadded + q;
      const float4 c = ld4(P.contrib + 4 * p);
      float alpha = P.thr[4 * p + 3];
      vec3f rgb = v3(c.x, c.y, c.z);
      if (rgb.x != rgb.x || rgb.y != rgb.y || rgb.z != rgb.z) {
        rgb = v3(0.f);
        alpha = 1.0f;
      }
      rgb = vmin(rgb, v3(F.maxRadiance));
      acc.x = acc.x + rgb.x;
      acc.y = acc.y + rgb.y;
      acc.z = acc.z + rgb.z;
      acc.w = acc.w + alpha;
      if (F.accAux) {
        const float *a = P.aux + 16 * p;
        for (int k = 0; k < 12; k++)
          aux[k] = aux[k] + a[k];
      }
      if (fabsf(c.w) < fabsf(z)) {
        z = c.w;
        ids[0] = P.ids[4 * p + 0];
        ids[1] = P.ids[4 * p + 1];
        ids[2] = P.ids[4 * p + 2];
      }
    }
    st4(F.accRGBA + 4 * (size_t)pix, acc.x, acc.y, acc.z, acc.w);
    if (F.accAux)
      for (int k = 0; k < 12; k++)
        F.accAux[12 * (size_t)pix + k] = aux[k];
    F.accZ[pix] = z;
    for (int k = 0; k < 3; k++)
      F.accIDs[3 * (size_t)pix + k] = ids[k];
  }
}

__device__ __forceinline__ float lerpAcc(float old, float nw, float f) { return (1.f - f) * old + f * nw; }

// ScreenSample_normalize + LocalFB_accumulateSample (ScreenSample.ih:79-92,
// fb/LocalFB.ispc:14-78)
__global__ __launch_bounds__(256) void k_fb_accumulate(DFrame F)
{
  if (frameCancelled(F.cancelLatch)) // a cancelled frame leaves the framebuffer as it was
    return;
  const int npix = F.width * F.height;
  const float rspp = 1.0f / (float)F.spp;
  const int accumID = (int)frameAccumID(F);
  const float accScale = 1.0f / (float)(accumID + 1);
  const float accVarScale = 1.0f / (float)(accumID / 2 + 1);
  const bool accVar = frameAccVar(F);
  const float adaptive = frameAdaptive(F);
  for (int pix = globalId(); pix < npix; pix += gridStride()) {
    if (!taskActive(F, adaptive, pix % F.width, pix / F.width))
      continue;
    const float4 a = ld4(F.accRGBA + 4 * (size_t)pix);
    const float r = a.x * rspp, g = a.y * rspp, b = a.z * rspp, al = a.w * rspp;
    if (F.fbColor) {
      float *c = F.fbColor + 4 * (size_t)pix;
      if (accScale > 0.8f)
        st4(c, r, g, b, al);
      else {
        const float4 old = ld4(c);
        st4(c, lerpAcc(old.x, r, accScale), lerpAcc(old.y, g, accScale), lerpAcc(old.z, b, accScale),
            lerpAcc(old.w, al, accScale));
      }
    }
    if (F.fbVariance && accVar) {
      float *c = F.fbVariance + 4 * (size_t)pix;
      if (accVarScale > 0.8f)
        st4(c, r, g, b, al);
      else {
        const float4 old = ld4(c);
        st4(c, lerpAcc(old.x, r, accVarScale), lerpAcc(old.y, g, accVarScale), lerpAcc(old.z, b, accVarScale),
            lerpAcc(old.w, al, accVarScale));
      }
    }
    const float z = F.accZ[pix];
    bool updateIDs = accumID == 0;
    if (F.fbDepth) {
      if (accumID > 0 && fabsf(z) < fabsf(F.fbDepth[pix]))
        updateIDs = true;
      if (updateIDs)
        F.fbDepth[pix] = z;
    }
    if (updateIDs) {
      if (F.fbPrimID)
        F.fbPrimID[pix] = F.accIDs[3 * (size_t)pix];
      if (F.fbObjID)
        F.fbObjID[pix] = F.accIDs[3 * (size_t)pix + 1];
      if (F.fbInstID)
        F.fbInstID[pix] = F.accIDs[3 * (size_t)pix + 2];
    }
    float *dsts[4] = {F.fbPosition, F.fbFirstNormal, F.fbNormal, F.fbAlbedo};
    for (int k = 0; k < 4; k++) {
      float *dst = dsts[k];
      if (!dst)
        continue;
      for (int c = 0; c < 3; c++) {
        const float v = F.accAux[12 * (size_t)pix + 3 * k + c] * rspp;
        float *d = dst + 3 * (size_t)pix + c;
        *d = accScale > 0.8f ? v : lerpAcc(*d, v, accScale);
      }
    }
  }
}

// Variance_kernel (fb/frame_ops/Variance.ispc:11-59), one thread per task;
// run after odd frames (LocalFB.cpp:459-466): a captured batch frame
// checks its device-resident index
__global__ __launch_bounds__(256) void k_variance(DFrame F)
{
  if (F.frameState && (*F.frameState & 1u) == 0u)
    return;
  const int tx = (F.width + 7) / 8, ty = (F.height + 7) / 8;
  const int ntask = tx * ty;
  for (int t = globalId(); t < ntask; t += gridStride()) {
    const int bx = (t % tx) * 8, by = (t / tx) * 8;
    float v = 0.f;
    int cnt = 0;
    for (int y = 0; y < 8; y++) {
      const int fy = by + y;
      if (fy >= F.height)
        continue;
      for (int x = 0; x < 8; x++) {
        const int fx = bx + x;
        if (fx >= F.width)
          continue;
        const size_t id = (size_t)fy * F.width + fx;
        const float4 c = ld4(F.fbColor + 4 * id);
        const float4 va = ld4(F.fbVariance + 4 * id);
        const float den2 = (c.x + c.y + c.z) + (1.f - c.w);
        if (den2 > 0.f) {
          cnt++;
          const float diff = fabsf(c.x - va.x) + fabsf(c.y - va.y) + fabsf(c.z - va.z) + fabsf(c.w - va.w);
          v += diff * rsqrt(den2);
        }
      }
    }
    if (cnt)
      v *= rcp((float)cnt);
    F.taskVariance[t] = v * 64.f;
  }
}

// LiveVarianceFrameOp::getAvgError (Variance.cpp:32-46) with an FP64 sum;
// single block
__global__ __launch_bounds__(256) void k_variance_reduce(const float *taskVar, int n, float threshold, float *out,
    const uint32_t *frameState)
{
  if (frameState && (*frameState & 1u) == 0u) // even frame of a captured batch: nothing to reduce
    return;
  __shared__ double sSum[256];
  __shared__ float sMax[256];
  __shared__ int sCnt[256];
  double sum = 0.0;
  float mx = 0.f;
  int cnt = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const float e = taskVar[i];
    mx = maxf(mx, e);
    if (e >= threshold) {
      sum += (double)e;
      cnt++;
    }
  }
  sSum[threadIdx.x] = sum;
  sMax[threadIdx.x] = mx;
  sCnt[threadIdx.x] = cnt;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sSum[threadIdx.x] += sSum[threadIdx.x + s];
      sMax[threadIdx.x] = maxf(sMax[threadIdx.x], sMax[threadIdx.x + s]);
      sCnt[threadIdx.x] += sCnt[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    out[0] = sCnt[0] ? (float)(sSum[0] / (double)sCnt[0]) : sMax[0];
}

__device__ __forceinline__ uint32_t cvtU8(float f) { return (uint32_t)(255.f * clampf(f) + 0.5f); }
__device__ __forceinline__ float linearToSrgb(float x)
{
  x = clampf(x);
  return x <= 0.0031308f ? 12.92f * x : 1.055f * fexp(flog(x) * (1.f / 2.4f)) - 0.055f;
}

// ColorConversion_kernel (fb/frame_ops/ColorConversion.ispc:11-36)
__global__ __launch_bounds__(256) void k_color_convert(const float *color, int npix, int format, void *out)
{
  for (int i = globalId(); i < npix; i += gridStride()) {
    const float4 v = ld4(color + 4 * (size_t)i);
    if (format == 3) {
      st4((float *)out + 4 * (size_t)i, maxf(v.x, 0.f), maxf(v.y, 0.f), maxf(v.z, 0.f), minf(maxf(v.w, 0.f), 1.f));
    } else if (format == 1) {
      ((uint32_t *)out)[i] = cvtU8(v.x) | (cvtU8(v.y) << 8) | (cvtU8(v.z) << 16) | (cvtU8(v.w) << 24);
    } else {
      ((uint32_t *)out)[i] = cvtU8(linearToSrgb(v.x)) | (cvtU8(linearToSrgb(v.y)) << 8)
          | (cvtU8(linearToSrgb(v.z)) << 16) | (cvtU8(v.w) << 24);
    }
  }
}

// closest hits of an explicit ray list (first-hit parity checks and BVH
// statistics): rays8 = org.xyz, dir.xyz, t0, t1; out: t,u,v + inst,model,prim
__global__ __launch_bounds__(kTraceBlock) void k_trace_rays(DScene S, const float *rays8, int n, float *outTUV,
    uint32_t *outIDs, uint32_t *ovfBase, unsigned long long *stats)
{
  __shared__ uint32_t lds[kLdsDepth * kTraceBlock];
  uint32_t *ovf = ovfBase + (size_t)globalId() * kThreadRow;
  TraceStats st = {0, 0};
  for (int i = globalId(); i < n; i += gridStride()) {
    const float *r = rays8 + 8 * (size_t)i;
    Hit h;
    bool sawT;
    traceRay<MODE_CLOSEST, true, false, TIES_RESOLVE>(S, v3(r[0], r[1], r[2]), v3(r[3], r[4], r[5]), r[6], r[7], h, lds, ovf,
        sawT, st);
    outTUV[3 * (size_t)i] = h.hit ? h.t : kInf;
    outTUV[3 * (size_t)i + 1] = h.u;
    outTUV[3 * (size_t)i + 2] = h.v;
    outIDs[3 * (size_t)i] = h.hit ? h.inst : kInvalidID;
    outIDs[3 * (size_t)i + 1] = h.hit ? h.model : kInvalidID;
    outIDs[3 * (size_t)i + 2] = h.hit ? h.prim : kInvalidID;
  }
  if (stats) {
    waveAdd(&stats[0], st.nodes);
    waveAdd(&stats[1], st.prims);
  }
}

// The traversal scratch rows of hpt_trace_rays_noalloc: a file-scope device
// array, kThreadRow words for each thread of at most kNoAllocTraceBlocks
// resident blocks (62 MiB).  Every GPU that runs the module holds its own
// copy until the process ends; host code reaches it only through the
// address hipGetSymbolAddress returns on that GPU, looked up once per
// device on the first call there.
constexpr int kNoAllocTraceBlocks = 1024;
static __device__ __align__(256) uint32_t scratch[(size_t)kNoAllocTraceBlocks * kTraceBlock * kThreadRow];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratchMutex;
uint32_t *g_scratchPtr[kMaxDevices] = {};

hipError_t noAllocScratch(uint32_t **out)
{
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess)
    return e;
  if (dev < 0 || dev >= kMaxDevices)
    return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratchMutex);
  if (!g_scratchPtr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess)
      return e;
    if (!p)
      return hipErrorInvalidSymbol;
    g_scratchPtr[dev] = static_cast<uint32_t *>(p);
  }
  *out = g_scratchPtr[dev];
  return hipSuccess;
}
} // namespace

// ---- C-ABI launch layer -------------------------------------------------------------

namespace {
inline int gridFor(size_t n, int block, int maxBlocks)
{
  size_t b = (n + block - 1) / block;
  if (b < 1)
    b = 1;
  if (b > (size_t)maxBlocks)
    b = maxBlocks;
  return (int)b;
}
} // namespace

extern "C" {

namespace {
// counter masks: the shadow side of a bounce (records, transparency queue and
// their work fetch) and the extension side of the bounce reading queue q
// (the next queue's pair, the extension work fetch, the near-tie queue)
constexpr uint32_t kShadowSideMask =
    (1u << kCntShadow) | (1u << kCntFetchShadow) | (1u << kCntTrans) | (1u << kCntFetchTrans);
inline uint32_t extensionSideMask(int q)
{
  return (3u << (2 * (q ^ 1))) | (1u << kCntFetchExtend) | (1u << kCntTies);
}
hipError_t launchReset(const DPaths &p, uint32_t mask, int keep, hipStream_t s)
{
  hipLaunchKernelGGL(k_reset, dim3(1), dim3(kNumCounters), 0, s, p.counters, mask, keep, p.cancelFlag, p.cancelGen,
      p.cancelLatch);
  return hipGetLastError();
}
} // namespace

hipError_t hpt_reset_counters(const HptPaths *P, int keepIndex, hipStream_t s)
{
  const DPaths &p = *reinterpret_cast<const DPaths *>(P);
  // a full reset clears every counter; the per-bounce reset both sides of a
  // bounce, keeping the pair of the extension queue being read
  if (keepIndex < 0)
    return launchReset(p, (1u << kNumCounters) - 1u, -1, s);
  return launchReset(p, extensionSideMask(keepIndex) | kShadowSideMask, 2 * keepIndex, s);
}

hipError_t hpt_reset_extension_counters(const HptPaths *P, int keepIndex, hipStream_t s)
{
  const DPaths &p = *reinterpret_cast<const DPaths *>(P);
  return launchReset(p, (1u << kCntFetchExtend) | (1u << kCntTies), 2 * keepIndex, s);
}

hipError_t hpt_reset_append_counters(const HptPaths *P, int keepIndex, hipStream_t s)
{
  const DPaths &p = *reinterpret_cast<const DPaths *>(P);
  hipLaunchKernelGGL(k_reset, dim3(1), dim3(kNumCounters), 0, s, p.counters, 3u << (2 * (keepIndex ^ 1)), -1,
      (const uint32_t *)nullptr, 0u, (uint32_t *)nullptr);
  return hipGetLastError();
}

hipError_t hpt_reset_shadow_counters(const HptPaths *P, hipStream_t s)
{
  const DPaths &p = *reinterpret_cast<const DPaths *>(P);
  hipLaunchKernelGGL(k_reset, dim3(1), dim3(kNumCounters), 0, s, p.counters, kShadowSideMask, -1,
      (const uint32_t *)nullptr, 0u, (uint32_t *)nullptr);
  return hipGetLastError();
}

hipError_t hpt_fill_u32(uint32_t *ptr, size_t n, uint32_t value, hipStream_t s)
{
  hipLaunchKernelGGL(k_fill_u32, dim3(gridFor(n, 256, 4096)), dim3(256), 0, s, ptr, n, value);
  return hipGetLastError();
}

hipError_t hpt_copy(void *dst, const void *src, size_t bytes, hipStream_t s)
{
  if (bytes % 16 != 0)
    return hipErrorInvalidValue;
  const size_t n16 = bytes / 16;
  hipLaunchKernelGGL(k_copy, dim3(gridFor(n16, 256, 16384)), dim3(256), 0, s, (uint4 *)dst, (const uint4 *)src, n16);
  return hipGetLastError();
}

hipError_t hpt_generate(const HptFrame *F, const HptPaths *P, hipStream_t s)
{
  const DFrame &f = *reinterpret_cast<const DFrame *>(F);
  const size_t n = (size_t)f.passCount * f.numPixPadded;
  hipLaunchKernelGGL(k_generate, dim3(gridFor(n, 256, 16384)), dim3(256), 0, s, f, *reinterpret_cast<const DPaths *>(P));
  return hipGetLastError();
}

hipError_t hpt_extend(const HptScene *S, const HptPaths *P, int queue, int residentBlocks, int stats, hipStream_t s)
{
  const DScene &sc = *reinterpret_cast<const DScene *>(S);
  const DPaths &p = *reinterpret_cast<const DPaths *>(P);
  if (stats)
    hipLaunchKernelGGL(k_extend<true>, dim3(residentBlocks), dim3(kTraceBlock), 0, s, sc, p, queue);
  else
    hipLaunchKernelGGL(k_extend<false>, dim3(residentBlocks), dim3(kTraceBlock), 0, s, sc, p, queue);
  return hipGetLastError();
}

hipError_t hpt_ties(const HptScene *S, const HptPaths *P, int residentBlocks, int stats, hipStream_t s)
{
  const DScene &sc = *reinterpret_cast<const DScene *>(S);
  const DPaths &p = *reinterpret_cast<const DPaths *>(P);
  if (stats)
    hipLaunchKernelGGL(k_ties<true>, dim3(residentBlocks), dim3(kTraceBlock), 0, s, sc, p);
  else
    hipLaunchKernelGGL(k_ties<false>, dim3(residentBlocks), dim3(kTraceBlock), 0, s, sc, p);
  return hipGetLastError();
}

hipError_t hpt_shade(const HptScene *S, const HptFrame *F, const HptPaths *P, int queue, int blocks, hipStream_t s)
{
  const DScene &sc = *reinterpret_cast<const DScene *>(S);
  const DFrame &f = *reinterpret_cast<const DFrame *>(F);
  const DPaths &p = *reinterpret_cast<const DPaths *>(P);
  // the volume variant carries the sampler and transmittance code (more
  // registers), the cache variant the cache records (the caller passes a
  // frame with nrcBounce -1 for the other bounces); surface-only worlds run
  // the lean one
  const bool vol = sc.numVolInsts > 0;
  const bool nrc = f.nrcBounce >= 0;
  if (vol && nrc)
    hipLaunchKernelGGL((k_shade<true, true>), dim3(blocks), dim3(256), 0, s, sc, f, p, queue);
  else if (vol)
    hipLaunchKernelGGL((k_shade<true, false>), dim3(blocks), dim3(256), 0, s, sc, f, p, queue);
  else if (nrc)
    hipLaunchKernelGGL((k_shade<false, true>), dim3(blocks), dim3(256), 0, s, sc, f, p, queue);
  else
    hipLaunchKernelGGL((k_shade<false, false>), dim3(blocks), dim3(256), 0, s, sc, f, p, queue);
  return hipGetLastError();
}

hipError_t hpt_nrc_gather(const HptPaths *P, int which, int n, float *raw12, hipStream_t s)
{
  const DPaths &p = *reinterpret_cast<const DPaths *>(P);
  if (n <= 0)
    return hipSuccess;
  hipLaunchKernelGGL(k_nrc_gather, dim3(gridFor(n, 256, 8192)), dim3(256), 0, s,
      (const float *)(which ? p.nrcTrainRec : p.nrcRec), which ? p.nrcTrainQueue : p.nrcQueryQueue, n, raw12);
  return hipGetLastError();
}

hipError_t hpt_nrc_apply(const HptPaths *P, int n, const float *pred3, hipStream_t s)
{
  if (n <= 0)
    return hipSuccess;
  hipLaunchKernelGGL(k_nrc_apply, dim3(gridFor(n, 256, 8192)), dim3(256), 0, s, *reinterpret_cast<const DPaths *>(P),
      n, pred3);
  return hipGetLastError();
}

hipError_t hpt_nrc_targets(const HptPaths *P, int n, float *target3, hipStream_t s)
{
  if (n <= 0)
    return hipSuccess;
  hipLaunchKernelGGL(k_nrc_targets, dim3(gridFor(n, 256, 8192)), dim3(256), 0, s,
      *reinterpret_cast<const DPaths *>(P), n, target3);
  return hipGetLastError();
}

hipError_t hpt_nee(const HptScene *S, const HptFrame *F, const HptPaths *P, int queue, int blocks, hipStream_t s)
{
  hipLaunchKernelGGL(k_nee, dim3(blocks), dim3(kNeeBlock), 0, s, *reinterpret_cast<const DScene *>(S),
      *reinterpret_cast<const DFrame *>(F), *reinterpret_cast<const DPaths *>(P), queue);
  return hipGetLastError();
}

hipError_t hpt_shadow(
    const HptScene *S, const HptFrame *F, const HptPaths *P, int residentBlocks, int stats, hipStream_t s)
{
  const DScene &sc = *reinterpret_cast<const DScene *>(S);
  const DFrame &f = *reinterpret_cast<const DFrame *>(F);
  const DPaths &p = *reinterpret_cast<const DPaths *>(P);
  if (stats) {
    hipLaunchKernelGGL(k_shadow<true>, dim3(residentBlocks), dim3(kTraceBlock), 0, s, sc, f, p);
  } else {
    hipLaunchKernelGGL(k_shadow<false>, dim3(residentBlocks), dim3(kTraceBlock), 0, s, sc, f, p);
  }
  return hipGetLastError();
}

hipError_t hpt_shadow_trans(
    const HptScene *S, const HptFrame *F, const HptPaths *P, int residentBlocks, int stats, hipStream_t s)
{
  const DScene &sc = *reinterpret_cast<const DScene *>(S);
  const DFrame &f = *reinterpret_cast<const DFrame *>(F);
  const DPaths &p = *reinterpret_cast<const DPaths *>(P);
  if (stats)
    hipLaunchKernelGGL(k_shadow_trans<true>, dim3(residentBlocks), dim3(kTraceBlock), 0, s, sc, f, p);
  else
    hipLaunchKernelGGL(k_shadow_trans<false>, dim3(residentBlocks), dim3(kTraceBlock), 0, s, sc, f, p);
  return hipGetLastError();
}

hipError_t hpt_volume(const HptScene *S, const HptPaths *P, int queue, int blocks, int stats, hipStream_t s)
{
  hipLaunchKernelGGL(k_volume, dim3(blocks), dim3(256), 0, s, *reinterpret_cast<const DScene *>(S),
      *reinterpret_cast<const DPaths *>(P), queue, stats);
  return hipGetLastError();
}

hipError_t hpt_resolve(const HptScene *S, const HptFrame *F, const HptPaths *P, int queue, int blocks, hipStream_t s)
{
  const DScene &sc = *reinterpret_cast<const DScene *>(S);
  const DFrame &f = *reinterpret_cast<const DFrame *>(F);
  const DPaths &p = *reinterpret_cast<const DPaths *>(P);
  if (sc.numVolInsts > 0)
    hipLaunchKernelGGL(k_resolve<true>, dim3(blocks), dim3(256), 0, s, sc, f, p, queue);
  else
    hipLaunchKernelGGL(k_resolve<false>, dim3(blocks), dim3(256), 0, s, sc, f, p, queue);
  return hipGetLastError();
}

hipError_t hpt_finish_pass(const HptFrame *F, const HptPaths *P, hipStream_t s)
{
  const DFrame &f = *reinterpret_cast<const DFrame *>(F);
  hipLaunchKernelGGL(k_finish_pass, dim3(gridFor((size_t)f.width * f.height, 256, 16384)), dim3(256), 0, s, f,
      *reinterpret_cast<const DPaths *>(P));
  return hipGetLastError();
}

hipError_t hpt_fb_accumulate(const HptFrame *F, hipStream_t s)
{
  const DFrame &f = *reinterpret_cast<const DFrame *>(F);
  hipLaunchKernelGGL(k_fb_accumulate, dim3(gridFor((size_t)f.width * f.height, 256, 16384)), dim3(256), 0, s, f);
  return hipGetLastError();
}

hipError_t hpt_variance(const HptFrame *F, hipStream_t s)
{
  const DFrame &f = *reinterpret_cast<const DFrame *>(F);
  const size_t nt = (size_t)((f.width + 7) / 8) * ((f.height + 7) / 8);
  hipLaunchKernelGGL(k_variance, dim3(gridFor(nt, 256, 4096)), dim3(256), 0, s, f);
  return hipGetLastError();
}

hipError_t hpt_variance_reduce(const float *taskVariance, int n, float threshold, float *out,
    const uint32_t *frameState, hipStream_t s)
{
  hipLaunchKernelGGL(k_variance_reduce, dim3(1), dim3(256), 0, s, taskVariance, n, threshold, out, frameState);
  return hipGetLastError();
}

hipError_t hpt_color_convert(const float *color, int npix, int format, void *out, hipStream_t s)
{
  hipLaunchKernelGGL(k_color_convert, dim3(gridFor((size_t)npix, 256, 16384)), dim3(256), 0, s, color, npix, format, out);
  return hipGetLastError();
}

hipError_t hpt_trace_rays(const HptScene *S, const float *rays8, int n, float *outTUV, uint32_t *outIDs,
    uint32_t *overflow, int residentBlocks, unsigned long long *stats, hipStream_t s)
{
  hipLaunchKernelGGL(k_trace_rays, dim3(residentBlocks), dim3(kTraceBlock), 0, s, *reinterpret_cast<const DScene *>(S),
      rays8, n, outTUV, outIDs, overflow, stats);
  return hipGetLastError();
}

hipError_t hpt_trace_rays_noalloc(const HptScene *S, const float *rays8, int n, float *outTUV, uint32_t *outIDs,
    int residentBlocks, unsigned long long *stats, hipStream_t s)
{
  if (!S || n < 0 || residentBlocks < 1 || residentBlocks > kNoAllocTraceBlocks)
    return hipErrorInvalidValue;
  if (n > 0 && (!rays8 || !outTUV || !outIDs))
    return hipErrorInvalidValue;
  uint32_t *rows = nullptr;
  const hipError_t e = noAllocScratch(&rows);
  if (e != hipSuccess)
    return e;
  // closest-hit traversal uses only the stack overflow of a row, which it
  // writes before reading: nothing in the rows needs clearing
  hipLaunchKernelGGL(k_trace_rays, dim3(residentBlocks), dim3(kTraceBlock), 0, s,
      *reinterpret_cast<const DScene *>(S), rays8, n, outTUV, outIDs, rows, stats);
  return hipGetLastError();
}

int hpt_trace_noalloc_max_blocks() { return kNoAllocTraceBlocks; }

int hpt_trace_block_size() { return kTraceBlock; }
int hpt_overflow_depth() { return kThreadRow; }
int hpt_trans_record_hits() { return kTransRecHits; }
int hpt_stack_capacity() { return kLdsDepth + kOverflowDepth; }

} // extern "C"
