// Device BVH build: Morton sort + PLOC clustering + 4-wide collapse, written
// in the traversal kernels' node and primitive layout (pt_types.h).  The
// reference commits each group's BLAS through Embree (common/Group.cpp:
// 93-143); this is the GPU build that replaces it for the hip device.
//
// Stages (all on one stream; the host reads back one word per clustering
// step and per collapse level):
//   items    per item (triangle, quad half, sphere): box, and the centre
//            bounds reduced over all items
//   morton   30-bit Morton code of the box centre on a cubic grid over the
//            centre bounds
//   sort     stable LSD radix sort, 8-bit digits: per-block digit
//            histograms, one exclusive scan, then a scatter whose ranks
//            come from wave ballots (stable within a wave, waves and rounds
//            in order)
//   cluster  PLOC: every cluster takes the neighbour within +-radius of
//            the sorted order whose union box has the least surface area
//            (ties to the smaller index, so the globally closest link is
//            always mutual and every step merges); mutual pairs become new
//            inner nodes, the cluster list is compacted by a scan
//   collapse top-down over levels: a 4-wide node starts from a binary
//            node's two children and opens the child of largest surface
//            area while it has fewer than four; subtrees of at most
//            kMaxLeaf items become leaves, whose primitive records are
//            written at the subtree's first position in depth-first leaf
//            order (so a leaf's records are adjacent)
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "ospray_hip/hip_bvh.h"
#include "pt_types.h"

namespace {

using hpt::DNode4;
using hpt::DPrim;

constexpr int kBlock = 256;
constexpr int kMaxRadius = 32; // PLOC search window: up to kMaxRadius clusters on either side
constexpr uint32_t kMaxLeaf = 2; // items per leaf (as the host build)
constexpr int kSortItems = 16;
constexpr int kSortTile = kBlock * kSortItems;
constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;
constexpr float kInfF = __builtin_inff();

inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// ---- items --------------------------------------------------------------------------

struct Geo
{
  const float *vertex;
  const uint32_t *index;
  const float *spheres;
  int32_t geomType;
  uint32_t firstItem; // items before this geometry
  uint32_t halves; // items per primitive (2 for quads)
  uint32_t pad;
};

__device__ __forceinline__ int findGeo(const Geo *g, int ng, uint32_t item)
{
  int k = 0;
  while (k + 1 < ng && g[k + 1].firstItem <= item)
    k++;
  return k;
}

// the three vertices of triangle-like item (prim, half)
__device__ __forceinline__ void itemVerts(const Geo &g, uint32_t prim, uint32_t half, uint32_t vi[3])
{
  if (g.geomType == hpt::GEOM_TRIANGLES) {
    const uint32_t *ix = g.index + 3 * (size_t)prim;
    vi[0] = ix[0], vi[1] = ix[1], vi[2] = ix[2];
  } else {
    const uint32_t *ix = g.index + 4 * (size_t)prim;
    if (half == 0)
      vi[0] = ix[0], vi[1] = ix[1], vi[2] = ix[3];
    else
      vi[0] = ix[2], vi[1] = ix[3], vi[2] = ix[1];
  }
}

__global__ __launch_bounds__(kBlock) void k_items(const Geo *geos, int ng, uint32_t n, float4 *lo, float4 *hi,
    float *partial)
{
  __shared__ float red[6][kBlock];
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  float c[3] = {kInfF, kInfF, kInfF}, C[3] = {-kInfF, -kInfF, -kInfF};
  if (i < n) {
    const Geo &g = geos[findGeo(geos, ng, i)];
    const uint32_t local = i - g.firstItem, prim = local / g.halves, half = local - prim * g.halves;
    float b[3], B[3];
    if (g.geomType == hpt::GEOM_SPHERES) {
      const float *s = g.spheres + 4 * (size_t)prim;
      for (int k = 0; k < 3; k++) {
        b[k] = s[k] - s[3];
        B[k] = s[k] + s[3];
      }
    } else {
      uint32_t vi[3];
      itemVerts(g, prim, half, vi);
      for (int k = 0; k < 3; k++) {
        const float x0 = g.vertex[3 * (size_t)vi[0] + k], x1 = g.vertex[3 * (size_t)vi[1] + k],
                    x2 = g.vertex[3 * (size_t)vi[2] + k];
        b[k] = fminf(x0, fminf(x1, x2));
        B[k] = fmaxf(x0, fmaxf(x1, x2));
      }
    }
    lo[i] = make_float4(b[0], b[1], b[2], 0.f);
    hi[i] = make_float4(B[0], B[1], B[2], 0.f);
    for (int k = 0; k < 3; k++)
      c[k] = C[k] = 0.5f * (b[k] + B[k]);
  }
  for (int k = 0; k < 3; k++) {
    red[k][threadIdx.x] = c[k];
    red[3 + k][threadIdx.x] = C[k];
  }
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o)
      for (int k = 0; k < 3; k++) {
        red[k][threadIdx.x] = fminf(red[k][threadIdx.x], red[k][threadIdx.x + o]);
        red[3 + k][threadIdx.x] = fmaxf(red[3 + k][threadIdx.x], red[3 + k][threadIdx.x + o]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6)
    partial[6 * (size_t)blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// one block: centre bounds from the per-block partials
__global__ __launch_bounds__(kBlock) void k_bounds(const float *partial, uint32_t nb, float *bounds6)
{
  __shared__ float red[6][kBlock];
  float c[3] = {kInfF, kInfF, kInfF}, C[3] = {-kInfF, -kInfF, -kInfF};
  for (uint32_t b = threadIdx.x; b < nb; b += kBlock)
    for (int k = 0; k < 3; k++) {
      c[k] = fminf(c[k], partial[6 * (size_t)b + k]);
      C[k] = fmaxf(C[k], partial[6 * (size_t)b + 3 + k]);
    }
  for (int k = 0; k < 3; k++) {
    red[k][threadIdx.x] = c[k];
    red[3 + k][threadIdx.x] = C[k];
  }
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o)
      for (int k = 0; k < 3; k++) {
        red[k][threadIdx.x] = fminf(red[k][threadIdx.x], red[k][threadIdx.x + o]);
        red[3 + k][threadIdx.x] = fmaxf(red[3 + k][threadIdx.x], red[3 + k][threadIdx.x + o]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6)
    bounds6[threadIdx.x] = red[threadIdx.x][0];
}

__device__ __forceinline__ uint32_t spread10(uint32_t x)
{
  x &= 0x3ffu;
  x = (x | (x << 16)) & 0x030000ffu;
  x = (x | (x << 8)) & 0x0300f00fu;
  x = (x | (x << 4)) & 0x030c30c3u;
  x = (x | (x << 2)) & 0x09249249u;
  return x;
}

__global__ __launch_bounds__(kBlock) void k_morton(const float4 *lo, const float4 *hi, uint32_t n, const float *bounds6,
    uint32_t *keys, uint32_t *vals)
{
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n)
    return;
  const float4 a = lo[i], b = hi[i];
  const float p[3] = {0.5f * (a.x + b.x), 0.5f * (a.y + b.y), 0.5f * (a.z + b.z)};
  // one cubic grid over the centre bounds (the largest extent on every
  // axis): a thin axis then takes few code bits instead of splitting
  // spatial neighbours at coarse levels
  const float ext = fmaxf(bounds6[3] - bounds6[0], fmaxf(bounds6[4] - bounds6[1], bounds6[5] - bounds6[2]));
  uint32_t q[3];
  for (int k = 0; k < 3; k++) {
    const float t = ext > 0.f ? (p[k] - bounds6[k]) / ext : 0.f;
    const float s = fminf(fmaxf(t * 1024.f, 0.f), 1023.f);
    q[k] = (uint32_t)s;
  }
  keys[i] = (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
  vals[i] = i;
}

// ---- exclusive scan (uint32 or uint64 sums), tiles of kScanTile ----------------------

template <typename T>
__device__ T blockExclusiveScan(T v, T *waveSums, T &total)
{
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  T x = v;
  for (int o = 1; o < 64; o <<= 1) {
    const T y = __shfl_up(x, o);
    if (lane >= o)
      x += y;
  }
  if (lane == 63)
    waveSums[wave] = x;
  __syncthreads();
  T before = 0, all = 0;
  for (int w = 0; w < kBlock / 64; w++) {
    if (w < wave)
      before += waveSums[w];
    all += waveSums[w];
  }
  __syncthreads();
  total = all;
  return before + x - v;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_scan_reduce(const T *in, uint64_t n, T *sums)
{
  __shared__ T ws[kBlock / 64];
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
  T s = 0;
  for (int k = 0; k < kScanItems; k++)
    if (base + k < n)
      s += in[base + k];
  T total;
  blockExclusiveScan<T>(s, ws, total);
  if (threadIdx.x == 0)
    sums[blockIdx.x] = total;
}

// out = exclusive scan of in (in-place allowed) plus offsets[block]; the
// last block writes the grand total to *total
template <typename T>
__global__ __launch_bounds__(kBlock) void k_scan_apply(const T *in, T *out, uint64_t n, const T *offsets, T *total)
{
  __shared__ T ws[kBlock / 64];
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
  T v[kScanItems];
  T s = 0;
  for (int k = 0; k < kScanItems; k++) {
    v[k] = base + k < n ? in[base + k] : (T)0;
    s += v[k];
  }
  T blockTotal;
  T run = blockExclusiveScan<T>(s, ws, blockTotal) + (offsets ? offsets[blockIdx.x] : (T)0);
  for (int k = 0; k < kScanItems; k++) {
    if (base + k < n)
      out[base + k] = run;
    run += v[k];
  }
  if (total && blockIdx.x == gridDim.x - 1 && threadIdx.x == kBlock - 1)
    *total = run;
}

// workspace of the scan of n values: the block sums of every level
template <typename T>
size_t scanScratchElems(uint64_t n)
{
  size_t e = 0;
  while (n > 1) {
    n = cdiv(n, kScanTile);
    e += n + 1;
  }
  return e + 1;
}

template <typename T>
hipError_t exclusiveScan(const T *in, T *out, uint64_t n, T *scratch, T *total, hipStream_t s)
{
  const uint32_t nb = cdiv(n, kScanTile);
  if (nb <= 1) {
    k_scan_apply<T><<<1, kBlock, 0, s>>>(in, out, n, nullptr, total);
    return hipGetLastError();
  }
  T *sums = scratch, *rest = scratch + nb + 1;
  k_scan_reduce<T><<<nb, kBlock, 0, s>>>(in, n, sums);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return e;
  e = exclusiveScan<T>(sums, sums, nb, rest, nullptr, s);
  if (e != hipSuccess)
    return e;
  k_scan_apply<T><<<nb, kBlock, 0, s>>>(in, out, n, sums, total);
  return hipGetLastError();
}

// ---- radix sort ------------------------------------------------------------------------

__global__ __launch_bounds__(kBlock) void k_sort_hist(const uint32_t *keys, uint32_t n, int shift, uint32_t *hist)
{
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t base = blockIdx.x * (uint32_t)kSortTile;
  for (int k = 0; k < kSortItems; k++) {
    const uint32_t i = base + k * kBlock + threadIdx.x;
    if (i < n)
      atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[threadIdx.x * gridDim.x + blockIdx.x] = h[threadIdx.x]; // digit-major
}

__global__ __launch_bounds__(kBlock) void k_sort_scatter(const uint32_t *keysIn, const uint32_t *valsIn,
    uint32_t *keysOut, uint32_t *valsOut, uint32_t n, int shift, const uint32_t *offsets)
{
  __shared__ uint32_t base[256];
  __shared__ uint32_t wcnt[kBlock / 64][256];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  base[threadIdx.x] = offsets[threadIdx.x * gridDim.x + blockIdx.x];
  const uint32_t tile = blockIdx.x * (uint32_t)kSortTile;
  for (int k = 0; k < kSortItems; k++) {
    for (int x = threadIdx.x; x < (kBlock / 64) * 256; x += kBlock)
      (&wcnt[0][0])[x] = 0;
    __syncthreads();
    const uint32_t i = tile + k * kBlock + threadIdx.x;
    const bool ok = i < n;
    const uint32_t key = ok ? keysIn[i] : 0u, val = ok ? valsIn[i] : 0u;
    const uint32_t d = (key >> shift) & 255u;
    unsigned long long peers = __ballot(ok);
    for (int b = 0; b < 8; b++) {
      const bool bit = (d >> b) & 1u;
      const unsigned long long m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    const uint32_t rank = (uint32_t)__popcll(peers & ((1ull << lane) - 1ull));
    if (ok && rank == 0)
      wcnt[wave][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    { // thread t owns digit t: wave offsets of this round, then advance
      uint32_t run = base[threadIdx.x];
      for (int w = 0; w < kBlock / 64; w++) {
        const uint32_t c = wcnt[w][threadIdx.x];
        wcnt[w][threadIdx.x] = run;
        run += c;
      }
      base[threadIdx.x] = run;
    }
    __syncthreads();
    if (ok) {
      const uint32_t pos = wcnt[wave][d] + rank;
      keysOut[pos] = key;
      valsOut[pos] = val;
    }
    __syncthreads();
  }
}

// ---- PLOC clustering ------------------------------------------------------------------

__device__ __forceinline__ float unionHalfArea(float4 a0, float4 a1, float4 b0, float4 b1)
{
  const float dx = fmaxf(a1.x, b1.x) - fminf(a0.x, b0.x);
  const float dy = fmaxf(a1.y, b1.y) - fminf(a0.y, b0.y);
  const float dz = fmaxf(a1.z, b1.z) - fminf(a0.z, b0.z);
  const float a = dx * dy + dy * dz + dz * dx;
  return a >= 0.f ? a : kInfF; // NaN -> +inf
}

// leaves in sorted order: leaf node p holds item vals[p]
__global__ __launch_bounds__(kBlock) void k_leaves(const uint32_t *vals, uint32_t n, const float4 *itemLo,
    const float4 *itemHi, float4 *lo, float4 *hi, uint32_t *count, uint32_t *cluster)
{
  const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= n)
    return;
  const uint32_t it = vals[p];
  lo[p] = itemLo[it];
  hi[p] = itemHi[it];
  count[p] = 1;
  cluster[p] = p;
}

__global__ __launch_bounds__(kBlock) void k_ploc_nn(const uint32_t *cluster, uint32_t c, const float4 *lo,
    const float4 *hi, uint32_t *nn, int radius)
{
  __shared__ float4 sLo[kBlock + 2 * kMaxRadius], sHi[kBlock + 2 * kMaxRadius];
  const int64_t b0 = (int64_t)blockIdx.x * kBlock;
  for (int t = threadIdx.x; t < kBlock + 2 * kMaxRadius; t += kBlock) {
    const int64_t j = b0 - kMaxRadius + t;
    if (j >= 0 && j < (int64_t)c) {
      const uint32_t node = cluster[j];
      sLo[t] = lo[node];
      sHi[t] = hi[node];
    }
  }
  __syncthreads();
  const int64_t i = b0 + threadIdx.x;
  if (i >= (int64_t)c)
    return;
  const float4 a0 = sLo[threadIdx.x + kMaxRadius], a1 = sHi[threadIdx.x + kMaxRadius];
  unsigned long long best = ~0ull;
  for (int o = -radius; o <= radius; o++) {
    const int64_t j = i + o;
    if (o == 0 || j < 0 || j >= (int64_t)c)
      continue;
    const float area = unionHalfArea(a0, a1, sLo[threadIdx.x + kMaxRadius + o], sHi[threadIdx.x + kMaxRadius + o]);
    const unsigned long long key = ((unsigned long long)__float_as_uint(area) << 32) | (unsigned long long)j;
    best = key < best ? key : best;
  }
  nn[i] = (uint32_t)best;
}

// per cluster: bit 0 (low word) = survives the step, bit 32 = creates a node
__global__ __launch_bounds__(kBlock) void k_ploc_flags(const uint32_t *nn, uint32_t c, unsigned long long *flags)
{
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= c)
    return;
  const uint32_t j = nn[i];
  const bool mutual = nn[j] == i;
  const bool merge = mutual && i < j, valid = !(mutual && j < i);
  flags[i] = (valid ? 1ull : 0ull) | (merge ? (1ull << 32) : 0ull);
}

__global__ __launch_bounds__(kBlock) void k_ploc_compact(const uint32_t *cluster, const uint32_t *nn, uint32_t c,
    const unsigned long long *flags, const unsigned long long *scan, uint32_t nodeBase, uint32_t n,
    uint32_t *clusterOut, float4 *lo, float4 *hi, uint32_t *left, uint32_t *right, uint32_t *count)
{
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= c)
    return;
  const unsigned long long f = flags[i];
  if (!(f & 1ull))
    return;
  const unsigned long long sc = scan[i];
  const uint32_t pos = (uint32_t)sc;
  const uint32_t a = cluster[i];
  if (f >> 32) {
    const uint32_t b = cluster[nn[i]];
    const uint32_t k = nodeBase + (uint32_t)(sc >> 32);
    const float4 a0 = lo[a], a1 = hi[a], b0 = lo[b], b1 = hi[b];
    lo[k] = make_float4(fminf(a0.x, b0.x), fminf(a0.y, b0.y), fminf(a0.z, b0.z), 0.f);
    hi[k] = make_float4(fmaxf(a1.x, b1.x), fmaxf(a1.y, b1.y), fmaxf(a1.z, b1.z), 0.f);
    left[k - n] = a;
    right[k - n] = b;
    count[k] = count[a] + count[b];
    clusterOut[pos] = k;
  } else {
    clusterOut[pos] = a;
  }
}

// ---- collapse into 4-wide nodes -------------------------------------------------------

struct Entry
{
  uint32_t node; // binary node (inner)
  int32_t out; // its 4-wide node
  int32_t need; // stack entries pushed above it
  uint32_t off; // first leaf position of its subtree
};

__device__ __forceinline__ float halfArea(float4 a, float4 b)
{
  const float dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
  return dx * dy + dy * dz + dz * dx;
}

__device__ void writePrim(const Geo *geos, int ng, uint32_t item, DPrim *out)
{
  const Geo &g = geos[findGeo(geos, ng, item)];
  const uint32_t local = item - g.firstItem, prim = local / g.halves, half = local - prim * g.halves;
  DPrim P;
  P.primID = prim;
  P.pad = 0;
  const uint32_t model = (uint32_t)(&g - geos);
  if (g.geomType == hpt::GEOM_SPHERES) {
    const float *s = g.spheres + 4 * (size_t)prim;
    P.a[0] = s[0], P.a[1] = s[1], P.a[2] = s[2];
    P.b[0] = s[3], P.b[1] = 0.f, P.b[2] = 0.f;
    P.c[0] = P.c[1] = P.c[2] = 0.f;
    P.modelKind = model | ((uint32_t)hpt::PRIM_SPHERE << 28);
  } else {
    uint32_t vi[3];
    itemVerts(g, prim, half, vi);
    const uint32_t kind = g.geomType == hpt::GEOM_TRIANGLES ? hpt::PRIM_TRIANGLE
                                                           : (half == 0 ? hpt::PRIM_QUAD0 : hpt::PRIM_QUAD1);
    for (int k = 0; k < 3; k++) {
      P.a[k] = g.vertex[3 * (size_t)vi[0] + k];
      P.b[k] = g.vertex[3 * (size_t)vi[1] + k];
      P.c[k] = g.vertex[3 * (size_t)vi[2] + k];
    }
    P.modelKind = model | (kind << 28);
  }
  *out = P;
}

struct CollapseArgs
{
  const float4 *lo, *hi;
  const uint32_t *left, *right, *count, *vals;
  uint32_t n;
  const Geo *geos;
  int ng;
  DNode4 *nodes;
  DPrim *prims;
  uint32_t *counters; // [0] next-level entries, [1] nodes allocated, [2] largest stack need
};

__global__ __launch_bounds__(kBlock) void k_collapse(CollapseArgs A, const Entry *qin, uint32_t nin, Entry *qout)
{
  const uint32_t e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= nin)
    return;
  const Entry E = qin[e];
  uint32_t kid[4], off[4];
  kid[0] = A.left[E.node - A.n];
  kid[1] = A.right[E.node - A.n];
  off[0] = E.off;
  off[1] = E.off + A.count[kid[0]];
  int nk = 2;
  while (nk < 4) {
    int best = -1;
    float bestArea = -1.f;
    for (int c = 0; c < nk; c++)
      if (A.count[kid[c]] > kMaxLeaf) {
        const float a = halfArea(A.lo[kid[c]], A.hi[kid[c]]);
        if (a > bestArea) {
          bestArea = a;
          best = c;
        }
      }
    if (best < 0)
      break;
    const uint32_t k = kid[best];
    kid[best] = A.left[k - A.n];
    kid[nk] = A.right[k - A.n];
    off[nk] = off[best] + A.count[kid[best]];
    nk++;
  }
  const int need = E.need + nk - 1;
  atomicMax(&A.counters[2], (uint32_t)need);
  DNode4 N;
  for (int c = 0; c < 4; c++) {
    N.lox[c] = N.hix[c] = N.loy[c] = N.hiy[c] = N.loz[c] = N.hiz[c] = kInfF;
    N.child[c] = (int32_t)~0u;
    N.pad[c] = 0;
  }
  for (int c = 0; c < nk; c++) {
    const uint32_t k = kid[c];
    const float4 a = A.lo[k], b = A.hi[k];
    N.lox[c] = a.x, N.loy[c] = a.y, N.loz[c] = a.z;
    N.hix[c] = b.x, N.hiy[c] = b.y, N.hiz[c] = b.z;
    const uint32_t cnt = A.count[k];
    if (cnt <= kMaxLeaf) {
      N.child[c] = (int32_t) ~((off[c] << 3) | (cnt - 1u));
      if (k < A.n) {
        writePrim(A.geos, A.ng, A.vals[k], A.prims + off[c]);
      } else { // two leaves below: depth-first order
        writePrim(A.geos, A.ng, A.vals[A.left[k - A.n]], A.prims + off[c]);
        writePrim(A.geos, A.ng, A.vals[A.right[k - A.n]], A.prims + off[c] + 1);
      }
    } else {
      const int32_t idx = (int32_t)atomicAdd(&A.counters[1], 1u);
      N.child[c] = idx;
      const uint32_t slot = atomicAdd(&A.counters[0], 1u);
      qout[slot] = {k, idx, need, off[c]};
    }
  }
  A.nodes[E.out] = N;
}

__global__ __launch_bounds__(kBlock) void k_relocate(DNode4 *dst, const DNode4 *src, size_t n, int32_t nodeOff,
    uint32_t primOff)
{
  const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n)
    return;
  DNode4 N = src[i];
  for (int c = 0; c < 4; c++) {
    if (N.lox[c] == kInfF) // unused slot: never entered
      continue;
    const int32_t ch = N.child[c];
    if (ch >= 0) {
      N.child[c] = ch + nodeOff;
    } else {
      const uint32_t enc = ~(uint32_t)ch;
      N.child[c] = (int32_t) ~((((enc >> 3) + primOff) << 3) | (enc & 7u));
    }
  }
  dst[i] = N;
}

// ---- workspace layout ---------------------------------------------------------------

struct Layout
{
  size_t geos, itemLo, itemHi, partial, bounds, keysA, valsA, keysB, valsB, hist, lo, hi, count, left, right,
      clusterA, clusterB, nn, flags, scan, scanScratch, scanScratch32, total, queueA, queueB, counters, end;
};

Layout layout(uint64_t n)
{
  Layout L;
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t at = o;
    o += (bytes + 255) & ~(size_t)255;
    return at;
  };
  const uint32_t nb = cdiv(n, kBlock), sortBlocks = cdiv(n, kSortTile);
  L.geos = take(sizeof(Geo) * 4096);
  L.itemLo = take(16 * n);
  L.itemHi = take(16 * n);
  L.partial = take(24 * (size_t)nb);
  L.bounds = take(32);
  L.keysA = take(4 * n);
  L.valsA = take(4 * n);
  L.keysB = take(4 * n);
  L.valsB = take(4 * n);
  L.hist = take(4 * 256 * (size_t)sortBlocks);
  L.lo = take(16 * (2 * n));
  L.hi = take(16 * (2 * n));
  L.count = take(4 * (2 * n));
  L.left = take(4 * n);
  L.right = take(4 * n);
  L.clusterA = take(4 * n);
  L.clusterB = take(4 * n);
  L.nn = take(4 * n);
  L.flags = take(8 * n);
  L.scan = take(8 * n);
  L.scanScratch = take(8 * scanScratchElems<unsigned long long>(n));
  L.scanScratch32 = take(4 * scanScratchElems<uint32_t>(256 * (uint64_t)sortBlocks));
  L.total = take(16);
  L.queueA = take(sizeof(Entry) * n);
  L.queueB = take(sizeof(Entry) * n);
  L.counters = take(16);
  L.end = o;
  return L;
}

#define HPT_TRY(expr)                                                                                                  \
  do {                                                                                                                 \
    const hipError_t e_ = (expr);                                                                                      \
    if (e_ != hipSuccess)                                                                                              \
      return e_;                                                                                                       \
  } while (0)

} // namespace

extern "C" {

HPT_API uint64_t hpt_bvh_items(const HptBvhGeometry *geoms, int numGeoms)
{
  uint64_t n = 0;
  for (int g = 0; g < numGeoms; g++)
    n += (uint64_t)geoms[g].numPrims * (geoms[g].geomType == hpt::GEOM_QUADS ? 2u : 1u);
  return n;
}

HPT_API size_t hpt_bvh_workspace_bytes(uint64_t numItems) { return layout(numItems < 1 ? 1 : numItems).end; }

HPT_API hipError_t hpt_bvh_build(const HptBvhGeometry *geoms, int numGeoms, int radius, void *workspace,
    size_t workspaceBytes, void *outNodes, void *outPrims, uint64_t *hostWords, HptBvhResult *result, hipStream_t s)
{
  const uint64_t items = hpt_bvh_items(geoms, numGeoms);
  if (items < 3 || items >= (1ull << 28) || numGeoms < 1 || numGeoms > 4096 || radius < 1 || radius > kMaxRadius ||
      !workspace || !outNodes || !outPrims || !hostWords || !result)
    return hipErrorInvalidValue;
  const uint32_t n = (uint32_t)items;
  const Layout L = layout(n);
  if (workspaceBytes < L.end)
    return hipErrorInvalidValue;
  char *ws = (char *)workspace;
  auto at = [&](size_t off) { return (void *)(ws + off); };
  std::vector<Geo> hg((size_t)numGeoms);
  uint32_t first = 0;
  for (int g = 0; g < numGeoms; g++) {
    hg[g] = {geoms[g].vertex, geoms[g].index, geoms[g].spheres, geoms[g].geomType, first,
        geoms[g].geomType == hpt::GEOM_QUADS ? 2u : 1u, 0u};
    first += geoms[g].numPrims * hg[g].halves;
  }
  Geo *dgeo = (Geo *)at(L.geos);
  HPT_TRY(hipMemcpyAsync(dgeo, hg.data(), sizeof(Geo) * numGeoms, hipMemcpyHostToDevice, s));
  float4 *itemLo = (float4 *)at(L.itemLo), *itemHi = (float4 *)at(L.itemHi);
  float *partial = (float *)at(L.partial), *bounds = (float *)at(L.bounds);
  const uint32_t nb = cdiv(n, kBlock);
  k_items<<<nb, kBlock, 0, s>>>(dgeo, numGeoms, n, itemLo, itemHi, partial);
  HPT_TRY(hipGetLastError());
  k_bounds<<<1, kBlock, 0, s>>>(partial, nb, bounds);
  HPT_TRY(hipGetLastError());
  uint32_t *keys = (uint32_t *)at(L.keysA), *vals = (uint32_t *)at(L.valsA);
  uint32_t *keys2 = (uint32_t *)at(L.keysB), *vals2 = (uint32_t *)at(L.valsB);
  k_morton<<<nb, kBlock, 0, s>>>(itemLo, itemHi, n, bounds, keys, vals);
  HPT_TRY(hipGetLastError());

  // stable LSD radix sort of the 30-bit codes, 8-bit digits
  const uint32_t sortBlocks = cdiv(n, kSortTile);
  uint32_t *hist = (uint32_t *)at(L.hist);
  for (int shift = 0; shift < 32; shift += 8) {
    k_sort_hist<<<sortBlocks, kBlock, 0, s>>>(keys, n, shift, hist);
    HPT_TRY(hipGetLastError());
    HPT_TRY(exclusiveScan<uint32_t>(hist, hist, 256 * (uint64_t)sortBlocks, (uint32_t *)at(L.scanScratch32), nullptr,
        s));
    k_sort_scatter<<<sortBlocks, kBlock, 0, s>>>(keys, vals, keys2, vals2, n, shift, hist);
    HPT_TRY(hipGetLastError());
    uint32_t *t = keys;
    keys = keys2;
    keys2 = t;
    t = vals;
    vals = vals2;
    vals2 = t;
  }

  // PLOC
  float4 *lo = (float4 *)at(L.lo), *hi = (float4 *)at(L.hi);
  uint32_t *count = (uint32_t *)at(L.count), *left = (uint32_t *)at(L.left), *right = (uint32_t *)at(L.right);
  uint32_t *cl = (uint32_t *)at(L.clusterA), *cl2 = (uint32_t *)at(L.clusterB), *nn = (uint32_t *)at(L.nn);
  unsigned long long *flags = (unsigned long long *)at(L.flags), *scan = (unsigned long long *)at(L.scan);
  unsigned long long *total = (unsigned long long *)at(L.total);
  k_leaves<<<nb, kBlock, 0, s>>>(vals, n, itemLo, itemHi, lo, hi, count, cl);
  HPT_TRY(hipGetLastError());
  uint32_t c = n, nodeNext = n;
  int steps = 0;
  while (c > 1) {
    const uint32_t cb = cdiv(c, kBlock);
    k_ploc_nn<<<cb, kBlock, 0, s>>>(cl, c, lo, hi, nn, radius);
    HPT_TRY(hipGetLastError());
    k_ploc_flags<<<cb, kBlock, 0, s>>>(nn, c, flags);
    HPT_TRY(hipGetLastError());
    HPT_TRY(exclusiveScan<unsigned long long>(flags, scan, c, (unsigned long long *)at(L.scanScratch), total, s));
    k_ploc_compact<<<cb, kBlock, 0, s>>>(cl, nn, c, flags, scan, nodeNext, n, cl2, lo, hi, left, right, count);
    HPT_TRY(hipGetLastError());
    HPT_TRY(hipMemcpyAsync(hostWords, total, 8, hipMemcpyDeviceToHost, s));
    HPT_TRY(hipStreamSynchronize(s));
    const uint32_t survivors = (uint32_t)hostWords[0], merged = (uint32_t)(hostWords[0] >> 32);
    if (merged == 0 || survivors >= c) // cannot happen: the closest link is mutual
      return hipErrorUnknown;
    c = survivors;
    nodeNext += merged;
    uint32_t *t = cl;
    cl = cl2;
    cl2 = t;
    steps++;
  }
  if (nodeNext != 2 * n - 1)
    return hipErrorUnknown;
  const uint32_t root = 2 * n - 2; // the last node created
  float4 rootBox[2];
  HPT_TRY(hipMemcpyAsync(&rootBox[0], lo + root, 16, hipMemcpyDeviceToHost, s));
  HPT_TRY(hipMemcpyAsync(&rootBox[1], hi + root, 16, hipMemcpyDeviceToHost, s));

  // collapse, level by level
  Entry *qa = (Entry *)at(L.queueA), *qb = (Entry *)at(L.queueB);
  uint32_t *counters = (uint32_t *)at(L.counters);
  const Entry rootEntry = {root, 0, 0, 0u};
  HPT_TRY(hipMemcpyAsync(qa, &rootEntry, sizeof(Entry), hipMemcpyHostToDevice, s));
  const uint32_t initCounters[4] = {0u, 1u, 0u, 0u};
  HPT_TRY(hipMemcpyAsync(counters, initCounters, 16, hipMemcpyHostToDevice, s));
  CollapseArgs A = {lo, hi, left, right, count, vals, n, dgeo, numGeoms, (DNode4 *)outNodes, (DPrim *)outPrims,
      counters};
  uint32_t nin = 1;
  int levels = 0;
  while (nin > 0) {
    HPT_TRY(hipMemsetAsync(counters, 0, 4, s));
    k_collapse<<<cdiv(nin, kBlock), kBlock, 0, s>>>(A, qa, nin, qb);
    HPT_TRY(hipGetLastError());
    HPT_TRY(hipMemcpyAsync(hostWords, counters, 12, hipMemcpyDeviceToHost, s));
    HPT_TRY(hipStreamSynchronize(s));
    nin = ((const uint32_t *)hostWords)[0];
    Entry *t = qa;
    qa = qb;
    qb = t;
    levels++;
  }
  const uint32_t *w = (const uint32_t *)hostWords;
  result->numNodes = w[1];
  result->numItems = n;
  result->stackNeed = (int32_t)w[2];
  result->levels = levels;
  result->clusterSteps = steps;
  result->lo[0] = rootBox[0].x, result->lo[1] = rootBox[0].y, result->lo[2] = rootBox[0].z;
  result->hi[0] = rootBox[1].x, result->hi[1] = rootBox[1].y, result->hi[2] = rootBox[1].z;
  return hipSuccess;
}

HPT_API hipError_t hpt_bvh_relocate(void *dst, const void *src, size_t n, int32_t nodeOffset, uint32_t primOffset,
    hipStream_t s)
{
  if (n == 0)
    return hipSuccess;
  k_relocate<<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, s>>>((DNode4 *)dst, (const DNode4 *)src, n,
      nodeOffset, primOffset);
  return hipGetLastError();
}

} // extern "C"
