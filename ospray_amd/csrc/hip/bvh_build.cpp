// Binned-SAH BVH builder (host) producing 4-wide hpt::DNode4 nodes.
#include "bvh_build.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <stdexcept>
#include <thread>

namespace hpt {

namespace {

constexpr int kBins = 16;
constexpr int kMedianDepth = 40; // beyond this depth split at the median
constexpr float kInfF = __builtin_inff();

struct TNode
{
  BuildBox box;
  int32_t left = -1, right = -1; // local indices; -1 = leaf
  uint32_t first = 0, count = 0;
};

inline BuildBox emptyBox() { return {{kInfF, kInfF, kInfF}, {-kInfF, -kInfF, -kInfF}}; }
inline void grow(BuildBox &b, const BuildBox &o)
{
  for (int k = 0; k < 3; k++) {
    b.lo[k] = std::min(b.lo[k], o.lo[k]);
    b.hi[k] = std::max(b.hi[k], o.hi[k]);
  }
}
inline void grow(BuildBox &b, const float *p)
{
  for (int k = 0; k < 3; k++) {
    b.lo[k] = std::min(b.lo[k], p[k]);
    b.hi[k] = std::max(b.hi[k], p[k]);
  }
}
inline float halfArea(const BuildBox &b)
{
  const float dx = b.hi[0] - b.lo[0], dy = b.hi[1] - b.lo[1], dz = b.hi[2] - b.lo[2];
  if (dx < 0.f)
    return 0.f;
  return dx * dy + dy * dz + dz * dx;
}

struct Builder
{
  const std::vector<BuildBox> &boxes;
  std::vector<float> cent; // 3 per item
  std::vector<uint32_t> &order;
  int maxLeaf;

  Builder(const std::vector<BuildBox> &b, std::vector<uint32_t> &o, int leaf) : boxes(b), order(o), maxLeaf(leaf)
  {
    cent.resize(3 * b.size());
    for (size_t i = 0; i < b.size(); i++)
      for (int k = 0; k < 3; k++)
        cent[3 * i + k] = 0.5f * (b[i].lo[k] + b[i].hi[k]);
  }

  // split [begin,end) and return mid, or begin if it should be a leaf
  uint32_t split(uint32_t begin, uint32_t end, int depth, BuildBox &bb)
  {
    BuildBox cb = emptyBox();
    bb = emptyBox();
    for (uint32_t i = begin; i < end; i++) {
      grow(bb, boxes[order[i]]);
      grow(cb, &cent[3 * order[i]]);
    }
    const uint32_t n = end - begin;
    if ((int)n <= maxLeaf)
      return begin;
    int axis = 0;
    float ext[3];
    for (int k = 0; k < 3; k++)
      ext[k] = cb.hi[k] - cb.lo[k];
    if (ext[1] > ext[axis])
      axis = 1;
    if (ext[2] > ext[axis])
      axis = 2;
    if (!(ext[axis] > 0.f))
      return begin + n / 2; // coincident centroids: split the list
    if (depth >= kMedianDepth) {
      const uint32_t mid = begin + n / 2;
      std::nth_element(order.begin() + begin, order.begin() + mid, order.begin() + end,
          [&](uint32_t a, uint32_t b) { return cent[3 * a + axis] < cent[3 * b + axis]; });
      return mid;
    }
    float bestCost = kInfF;
    int bestAxis = -1, bestBin = -1;
    for (int ax = 0; ax < 3; ax++) {
      if (!(ext[ax] > 0.f))
        continue;
      BuildBox bin[kBins];
      uint32_t cnt[kBins] = {0};
      for (int b = 0; b < kBins; b++)
        bin[b] = emptyBox();
      const float scale = (float)kBins / ext[ax];
      const float lo = cb.lo[ax];
      for (uint32_t i = begin; i < end; i++) {
        const uint32_t id = order[i];
        int b = (int)((cent[3 * id + ax] - lo) * scale);
        b = b < 0 ? 0 : (b >= kBins ? kBins - 1 : b);
        cnt[b]++;
        grow(bin[b], boxes[id]);
      }
      float rArea[kBins];
      uint32_t rCnt[kBins];
      BuildBox acc = emptyBox();
      uint32_t c = 0;
      for (int b = kBins - 1; b > 0; b--) {
        grow(acc, bin[b]);
        c += cnt[b];
        rArea[b] = halfArea(acc);
        rCnt[b] = c;
      }
      acc = emptyBox();
      c = 0;
      for (int b = 0; b < kBins - 1; b++) {
        grow(acc, bin[b]);
        c += cnt[b];
        if (c == 0 || rCnt[b + 1] == 0)
          continue;
        const float cost = halfArea(acc) * (float)c + rArea[b + 1] * (float)rCnt[b + 1];
        if (cost < bestCost) {
          bestCost = cost;
          bestAxis = ax;
          bestBin = b;
        }
      }
    }
    if (bestAxis < 0)
      return begin + n / 2;
    const float scale = (float)kBins / ext[bestAxis];
    const float lo = cb.lo[bestAxis];
    auto pm = std::partition(order.begin() + begin, order.begin() + end, [&](uint32_t id) {
      int b = (int)((cent[3 * id + bestAxis] - lo) * scale);
      b = b < 0 ? 0 : (b >= kBins ? kBins - 1 : b);
      return b <= bestBin;
    });
    uint32_t mid = (uint32_t)(pm - order.begin());
    if (mid == begin || mid == end)
      mid = begin + n / 2;
    return mid;
  }

  // serial build of a subtree into `out` (root at out[rootIdx])
  void buildSerial(std::vector<TNode> &out, uint32_t begin, uint32_t end, int depth0, int &maxDepth)
  {
    struct Task
    {
      int32_t node;
      uint32_t begin, end;
      int depth;
    };
    std::vector<Task> stack;
    out.push_back(TNode{});
    stack.push_back({(int32_t)out.size() - 1, begin, end, depth0});
    while (!stack.empty()) {
      const Task t = stack.back();
      stack.pop_back();
      maxDepth = std::max(maxDepth, t.depth);
      BuildBox bb;
      const uint32_t mid = split(t.begin, t.end, t.depth, bb);
      out[t.node].box = bb;
      if (mid == t.begin) {
        out[t.node].first = t.begin;
        out[t.node].count = t.end - t.begin;
        continue;
      }
      const int32_t l = (int32_t)out.size();
      out.push_back(TNode{});
      out.push_back(TNode{});
      out[t.node].left = l;
      out[t.node].right = l + 1;
      stack.push_back({l + 1, mid, t.end, t.depth + 1});
      stack.push_back({l, t.begin, mid, t.depth + 1});
    }
  }
};

} // namespace

BuildResult buildBVH(const std::vector<BuildBox> &boxes, int maxLeaf, int nthreads)
{
  if (maxLeaf < 1 || maxLeaf > 8)
    throw std::runtime_error("BVH leaf size must be 1..8");
  if (boxes.size() >= (1u << 28))
    throw std::runtime_error("too many primitives in one BVH");
  BuildResult R;
  R.order.resize(boxes.size());
  for (size_t i = 0; i < boxes.size(); i++)
    R.order[i] = (uint32_t)i;
  DNode4 emptyNode;
  for (int c = 0; c < 4; c++) {
    emptyNode.lox[c] = emptyNode.hix[c] = kInfF;
    emptyNode.loy[c] = emptyNode.hiy[c] = kInfF;
    emptyNode.loz[c] = emptyNode.hiz[c] = kInfF;
    emptyNode.child[c] = (int32_t)~0u;
    emptyNode.pad[c] = 0;
  }
  if (boxes.empty()) {
    R.nodes.push_back(emptyNode);
    R.bounds = emptyBox();
    return R;
  }
  Builder B(boxes, R.order, maxLeaf);

  // temporary binary tree; the top is split serially into subtree tasks
  // that are built in parallel
  std::vector<TNode> top;
  struct Pending
  {
    int32_t node;
    uint32_t begin, end;
    int depth;
  };
  std::vector<Pending> work, leavesToBuild;
  top.push_back(TNode{});
  work.push_back({0, 0, (uint32_t)boxes.size(), 0});
  const uint32_t grain = std::max<uint32_t>(4096u, (uint32_t)(boxes.size() / (size_t)(std::max(1, nthreads) * 4)));
  int maxDepth = 0;
  while (!work.empty()) {
    Pending t = work.back();
    work.pop_back();
    if (t.end - t.begin <= grain || nthreads <= 1) {
      leavesToBuild.push_back(t);
      continue;
    }
    maxDepth = std::max(maxDepth, t.depth);
    BuildBox bb;
    const uint32_t mid = B.split(t.begin, t.end, t.depth, bb);
    top[t.node].box = bb;
    if (mid == t.begin) {
      top[t.node].first = t.begin;
      top[t.node].count = t.end - t.begin;
      continue;
    }
    const int32_t l = (int32_t)top.size();
    top.push_back(TNode{});
    top.push_back(TNode{});
    top[t.node].left = l;
    top[t.node].right = l + 1;
    work.push_back({l + 1, mid, t.end, t.depth + 1});
    work.push_back({l, t.begin, mid, t.depth + 1});
  }
  std::vector<std::vector<TNode>> sub(leavesToBuild.size());
  std::vector<int> subDepth(leavesToBuild.size(), 0);
  {
    std::vector<std::thread> th;
    std::atomic<size_t> next{0};
    const int nt = std::max(1, std::min<int>(nthreads, (int)leavesToBuild.size()));
    for (int w = 0; w < nt; w++)
      th.emplace_back([&]() {
        for (;;) {
          const size_t k = next.fetch_add(1);
          if (k >= leavesToBuild.size())
            break;
          const Pending &t = leavesToBuild[k];
          B.buildSerial(sub[k], t.begin, t.end, t.depth, subDepth[k]);
        }
      });
    for (auto &t : th)
      t.join();
  }
  for (int d : subDepth)
    maxDepth = std::max(maxDepth, d);
  // stitch: top-level placeholder nodes become references to subtree roots
  std::vector<std::pair<int, int32_t>> ref(top.size(), {-1, -1}); // (subtree, local idx)
  for (size_t k = 0; k < leavesToBuild.size(); k++)
    ref[leavesToBuild[k].node] = {(int)k, 0};
  auto get = [&](int tree, int32_t idx) -> const TNode & { return tree < 0 ? top[idx] : sub[tree][idx]; };
  // resolve a (tree, idx) that is a top placeholder into its subtree root
  auto resolve = [&](int tree, int32_t idx) -> std::pair<int, int32_t> {
    if (tree < 0 && ref[idx].first >= 0)
      return ref[idx];
    return {tree, idx};
  };

  // collapse into 4-wide nodes, emitted depth-first
  const auto root = resolve(-1, 0);
  R.bounds = get(root.first, root.second).box;
  using Ref = std::pair<int, int32_t>;
  auto setChild = [&](DNode4 &n, int c, const TNode &t) {
    n.lox[c] = t.box.lo[0];
    n.loy[c] = t.box.lo[1];
    n.loz[c] = t.box.lo[2];
    n.hix[c] = t.box.hi[0];
    n.hiy[c] = t.box.hi[1];
    n.hiz[c] = t.box.hi[2];
  };
  auto leafCode = [](uint32_t first, uint32_t count) -> int32_t {
    return (int32_t) ~((first << 3) | (count - 1u));
  };
  auto area = [](const BuildBox &b) {
    const float dx = b.hi[0] - b.lo[0], dy = b.hi[1] - b.lo[1], dz = b.hi[2] - b.lo[2];
    return dx * dy + dy * dz + dz * dx;
  };
  const TNode &r0 = get(root.first, root.second);
  if (r0.left < 0) {
    DNode4 n = emptyNode;
    setChild(n, 0, r0);
    n.child[0] = leafCode(r0.first, r0.count);
    R.nodes.push_back(n);
    R.maxDepth = 1;
    return R;
  }
  struct Emit
  {
    Ref ref;
    int32_t out;
    int need; // stack entries pushed above this node
  };
  R.nodes.reserve(boxes.size() / 2 + 1);
  std::vector<Emit> stack;
  R.nodes.push_back(emptyNode);
  stack.push_back({root, 0, 0});
  while (!stack.empty()) {
    const Emit e = stack.back();
    stack.pop_back();
    const TNode &tn = get(e.ref.first, e.ref.second);
    Ref kids[4] = {resolve(e.ref.first, tn.left), resolve(e.ref.first, tn.right), {-1, -1}, {-1, -1}};
    int nk = 2;
    while (nk < 4) {
      int best = -1;
      float bestArea = -1.f;
      for (int c = 0; c < nk; c++) {
        const TNode &t = get(kids[c].first, kids[c].second);
        if (t.left >= 0 && area(t.box) > bestArea) {
          bestArea = area(t.box);
          best = c;
        }
      }
      if (best < 0)
        break;
      const TNode &t = get(kids[best].first, kids[best].second);
      const int tree = kids[best].first;
      kids[best] = resolve(tree, t.left);
      kids[nk++] = resolve(tree, t.right);
    }
    const int need = e.need + (nk - 1);
    R.stackNeed = std::max(R.stackNeed, need);
    DNode4 n = emptyNode;
    int32_t slot[4] = {-1, -1, -1, -1};
    for (int c = 0; c < nk; c++) {
      const TNode &ch = get(kids[c].first, kids[c].second);
      setChild(n, c, ch);
      if (ch.left < 0) {
        n.child[c] = leafCode(ch.first, ch.count);
      } else {
        slot[c] = (int32_t)R.nodes.size();
        R.nodes.push_back(emptyNode);
        n.child[c] = slot[c];
      }
    }
    R.nodes[e.out] = n;
    for (int c = nk - 1; c >= 0; c--)
      if (slot[c] >= 0)
        stack.push_back({kids[c], slot[c], need});
  }
  R.maxDepth = maxDepth + 1;
  return R;
}

} // namespace hpt
