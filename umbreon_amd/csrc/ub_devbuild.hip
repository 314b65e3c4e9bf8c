// ub_devbuild.hip -- BVH8 construction on the GPU.
//
// 1. Morton order: centroid bounds (block reduction + ordered-int atomics),
//    63-bit Morton codes (21 bits per axis), rocPRIM's device radix sort.
// 2. Binary tree by PLOC (Meister and Bittner, "Parallel Locally-Ordered
//    Clustering for Bounding Volume Hierarchy Construction", TVCG 2018):
//    clusters start as the sorted leaves; each round every cluster finds the
//    neighbour within kWindow positions whose merged box has the smallest
//    surface area (ties by a strict order on pairs), mutual pairs merge into a new
//    node that takes the lower position, and a prefix sum compacts the
//    cluster list and numbers the new nodes, so the tree does not depend on
//    thread timing. Rounds repeat until one cluster is left.
// 3. Wide collapse costs: when a node is created both children are final, so
//    the merge kernel also fills the node's SAH-optimal 8-wide collapse table
//    (the same dynamic programming as the host builder: cost[i] = least SAH
//    cost of the subtree spread over at most i slots of its parent, a subtree
//    of at most 3 primitives may be one leaf slot).
// 4. Emission, one kernel per level of the 8-wide tree: each node expands its
//    binary subtree into <= 8 slots by the collapse table, places children in
//    octant slots, quantizes their boxes conservatively in FP64 (floor / ceil
//    against a power-of-two scale, so every 8-bit box contains its FP32 box)
//    and reserves its inner children and its leaf primitives as contiguous
//    ranges (one atomic per wave). Node and primitive placement in memory may
//    vary from run to run; what a ray visits and in which order does not.
#include "ub_devbuild.h"

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "../../include/umbreon_gpu.h"
#include "ub_layout.h"

namespace ubg {
namespace {

constexpr int kBlock = 256;
constexpr int kWindow = 16;  // PLOC search radius
constexpr float kNodeCost = 1.0f;
constexpr float kPrimCost = 1.0f;

struct Tree {
  float4* blo;
  float4* bhi;
  int32_t* left;
  int32_t* right;
  int32_t* count;
  float* cost;       // [8 * node + i - 1], i = 1..8
  uint8_t* distK;    // [8 * node + j - 1], j = 2..8
  uint16_t* useDist; // bit i (i = 2..8)
  uint8_t* leafSel;  // subtree is one leaf slot
};

__device__ __forceinline__ float areaOf(float4 lo, float4 hi) {
  const float dx = hi.x - lo.x, dy = hi.y - lo.y, dz = hi.z - lo.z;
  return 2.0f * (dx * dy + dy * dz + dz * dx);
}
__device__ __forceinline__ float4 min4(float4 a, float4 b) {
  return make_float4(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z), 0.0f);
}
__device__ __forceinline__ float4 max4(float4 a, float4 b) {
  return make_float4(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z), 0.0f);
}

// float <-> int with the same order, for atomicMin / atomicMax on floats
__device__ __forceinline__ int orderedInt(float f) {
  const int i = __float_as_int(f);
  return i >= 0 ? i : i ^ 0x7FFFFFFF;
}
__device__ __forceinline__ float orderedFloat(int i) { return __int_as_float(i >= 0 ? i : i ^ 0x7FFFFFFF); }

__global__ void k_bounds_init(int* b) {
  if (threadIdx.x < 3) b[threadIdx.x] = 0x7FFFFFFF;
  else if (threadIdx.x < 6) b[threadIdx.x] = static_cast<int>(0x80000000u);
}

// Centroid bounds: grid-stride, block reduction in LDS, 6 atomics per block.
__global__ __launch_bounds__(kBlock) void k_bounds(const float4* boxes, int64_t n, int* b) {
  __shared__ float red[6][kBlock];
  float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t i = blockIdx.x * static_cast<int64_t>(kBlock) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * kBlock) {
    const float4 lo = boxes[2 * i], hi = boxes[2 * i + 1];
    const float c[3] = {0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z)};
    for (int k = 0; k < 3; ++k) {
      mn[k] = fminf(mn[k], c[k]);
      mx[k] = fmaxf(mx[k], c[k]);
    }
  }
  for (int k = 0; k < 3; ++k) {
    red[k][threadIdx.x] = mn[k];
    red[3 + k][threadIdx.x] = mx[k];
  }
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (static_cast<int>(threadIdx.x) < s)
      for (int k = 0; k < 3; ++k) {
        red[k][threadIdx.x] = fminf(red[k][threadIdx.x], red[k][threadIdx.x + s]);
        red[3 + k][threadIdx.x] = fmaxf(red[3 + k][threadIdx.x], red[3 + k][threadIdx.x + s]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 3) atomicMin(&b[threadIdx.x], orderedInt(red[threadIdx.x][0]));
  else if (threadIdx.x < 6) atomicMax(&b[threadIdx.x], orderedInt(red[threadIdx.x][0]));
}

__device__ __forceinline__ uint64_t spread21(uint64_t v) {
  v &= 0x1FFFFFull;
  v = (v | (v << 32)) & 0x1F00000000FFFFull;
  v = (v | (v << 16)) & 0x1F0000FF0000FFull;
  v = (v | (v << 8)) & 0x100F00F00F00F00Full;
  v = (v | (v << 4)) & 0x10C30C30C30C30C3ull;
  v = (v | (v << 2)) & 0x1249249249249249ull;
  return v;
}

__global__ __launch_bounds__(kBlock) void k_morton(const float4* boxes, int64_t n, const int* b, uint64_t* keys,
                                                   int32_t* vals) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(kBlock) + threadIdx.x;
  if (i >= n) return;
  const float4 lo = boxes[2 * i], hi = boxes[2 * i + 1];
  const float c[3] = {0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z)};
  uint64_t q[3];
  for (int k = 0; k < 3; ++k) {
    const float a = orderedFloat(b[k]), e = orderedFloat(b[3 + k]) - a;
    const float u = e > 0.0f ? (c[k] - a) / e : 0.0f;
    const float s = fminf(fmaxf(u, 0.0f), 1.0f) * 2097152.0f;  // 2^21
    q[k] = static_cast<uint64_t>(fminf(s, 2097151.0f));
  }
  keys[i] = (spread21(q[0]) << 2) | (spread21(q[1]) << 1) | spread21(q[2]);
  vals[i] = static_cast<int32_t>(i);
}

// Leaf node i = sorted primitive i.
__global__ __launch_bounds__(kBlock) void k_leaves(const float4* boxes, const int32_t* perm, int64_t n, Tree t,
                                                   int32_t* cl) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(kBlock) + threadIdx.x;
  if (i >= n) return;
  const int32_t p = perm[i];
  const float4 lo = boxes[2 * p], hi = boxes[2 * p + 1];
  t.blo[i] = lo;
  t.bhi[i] = hi;
  t.count[i] = 1;
  t.left[i] = -1;
  t.right[i] = -1;
  const float c = areaOf(lo, hi) * kPrimCost;
  for (int k = 0; k < 8; ++k) t.cost[8 * static_cast<size_t>(i) + k] = c;
  t.useDist[i] = 0;
  t.leafSel[i] = 1;
  cl[i] = static_cast<int32_t>(i);
}

// Nearest neighbour (smallest merged area) of each cluster within kWindow.
__global__ __launch_bounds__(kBlock) void k_nearest(const int32_t* cl, int m, Tree t, int32_t* nn) {
  __shared__ float4 slo[kBlock + 2 * kWindow];
  __shared__ float4 shi[kBlock + 2 * kWindow];
  const int b0 = blockIdx.x * kBlock;
  for (int k = threadIdx.x; k < kBlock + 2 * kWindow; k += kBlock) {
    const int idx = b0 - kWindow + k;
    if (idx >= 0 && idx < m) {
      const int32_t c = cl[idx];
      slo[k] = t.blo[c];
      shi[k] = t.bhi[c];
    }
  }
  __syncthreads();
  const int i = b0 + threadIdx.x;
  if (i >= m) return;
  const float4 lo = slo[threadIdx.x + kWindow], hi = shi[threadIdx.x + kWindow];
  // Candidates are ranked by a key of the unordered pair: merged area, then
  // "position buddies" (2k, 2k+1) first, then the closer position, then the
  // lower one. A strict order on pairs makes the best pair of the whole list
  // mutual, so every round merges at least one pair, and equal boxes pair up
  // evenly instead of chaining.
  float best = INFINITY;
  int bj = -1, bBuddy = 2, bDist = 0;
  const int j0 = max(0, i - kWindow), j1 = min(m - 1, i + kWindow);
  for (int j = j0; j <= j1; ++j) {
    if (j == i) continue;
    const int k = j - b0 + kWindow;
    const float a = areaOf(min4(lo, slo[k]), max4(hi, shi[k]));
    const int buddy = j == (i ^ 1) ? 0 : 1;
    const int dist = abs(j - i);
    // j runs upwards, so on a full tie the earlier (lower) j stays
    if (a < best || (a == best && (buddy < bBuddy || (buddy == bBuddy && dist < bDist)))) {
      best = a;
      bj = j;
      bBuddy = buddy;
      bDist = dist;
    }
  }
  nn[i] = bj;
}

// packed[i] = (creates a node) << 32 | (keeps a cluster slot)
__global__ __launch_bounds__(kBlock) void k_flags(const int32_t* nn, int m, unsigned long long* packed) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= m) return;
  const int j = nn[i];
  const bool mutual = j >= 0 && nn[j] == i;
  const unsigned long long create = (mutual && i < j) ? 1ull : 0ull;
  const unsigned long long keep = (mutual && i > j) ? 0ull : 1ull;
  packed[i] = (create << 32) | keep;
}

__global__ __launch_bounds__(kBlock) void k_merge(const int32_t* cl, const int32_t* nn, int m,
                                                  const unsigned long long* packed, const unsigned long long* pre,
                                                  int32_t nodeBase, Tree t, int32_t* cl2, int32_t* totals,
                                                  int leafMax) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= m) return;
  const unsigned long long f = packed[i], p = pre[i];
  const uint32_t keep = static_cast<uint32_t>(f & 0xFFFFFFFFull), create = static_cast<uint32_t>(f >> 32);
  const uint32_t keepAt = static_cast<uint32_t>(p & 0xFFFFFFFFull), createAt = static_cast<uint32_t>(p >> 32);
  if (i == m - 1) {
    totals[0] = static_cast<int32_t>(keepAt + keep);
    totals[1] = static_cast<int32_t>(createAt + create);
  }
  if (!keep) return;
  if (!create) {
    cl2[keepAt] = cl[i];
    return;
  }
  const int32_t L = cl[i], R = cl[nn[i]];
  const int32_t node = nodeBase + static_cast<int32_t>(createAt);
  const float4 lo = min4(t.blo[L], t.blo[R]), hi = max4(t.bhi[L], t.bhi[R]);
  t.blo[node] = lo;
  t.bhi[node] = hi;
  t.left[node] = L;
  t.right[node] = R;
  const int32_t cnt = t.count[L] + t.count[R];
  t.count[node] = cnt;
  // 8-wide collapse table
  float cL[8], cR[8];
  for (int k = 0; k < 8; ++k) {
    cL[k] = t.cost[8 * static_cast<size_t>(L) + k];
    cR[k] = t.cost[8 * static_cast<size_t>(R) + k];
  }
  const float A = areaOf(lo, hi);
  float dist[9];
  uint8_t dk[9];
  for (int j = 2; j <= 8; ++j) {
    float best = INFINITY;
    int bk = 1;
    for (int k = 1; k < j; ++k) {
      const float c = cL[k - 1] + cR[j - k - 1];
      if (c < best) {
        best = c;
        bk = k;
      }
    }
    dist[j] = best;
    dk[j] = static_cast<uint8_t>(bk);
  }
  const float asNode = A * kNodeCost + dist[8];
  const float asLeaf = cnt <= leafMax ? A * static_cast<float>(cnt) * kPrimCost : INFINITY;
  t.leafSel[node] = asLeaf <= asNode ? 1 : 0;
  float prev = fminf(asLeaf, asNode);
  const size_t row = 8 * static_cast<size_t>(node);
  t.cost[row] = prev;
  uint16_t use = 0;
  t.distK[row] = 0;
  for (int i2 = 2; i2 <= 8; ++i2) {
    const bool u = dist[i2] < prev;
    if (u) use |= static_cast<uint16_t>(1u << i2);
    prev = u ? dist[i2] : prev;
    t.cost[row + i2 - 1] = prev;
    t.distK[row + i2 - 1] = dk[i2];
  }
  t.useDist[node] = use;
  cl2[keepAt] = node;
}

// SAH of the binary tree: sum of node areas / root area (every leaf holds one primitive).
__global__ __launch_bounds__(kBlock) void k_sah(const Tree t, int64_t nNodes, double* out) {
  __shared__ double red[kBlock];
  double s = 0.0;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(kBlock) + threadIdx.x; i < nNodes;
       i += static_cast<int64_t>(gridDim.x) * kBlock)
    s += static_cast<double>(areaOf(t.blo[i], t.bhi[i]));
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = kBlock / 2; k > 0; k >>= 1) {
    if (static_cast<int>(threadIdx.x) < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(out, red[0]);
}

__device__ __forceinline__ uint32_t waveInclusiveSum(uint32_t v) {
  const int lane = static_cast<int>(__lane_id());
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t u = __shfl_up(v, off, 64);
    if (lane >= off) v += u;
  }
  return v;
}

// Reserves `v` consecutive entries of counter `c` for this lane: one atomic
// per wave. Every lane of the wave must call it.
__device__ __forceinline__ uint32_t waveReserve(unsigned int* c, uint32_t v) {
  const uint32_t incl = waveInclusiveSum(v);
  const uint32_t total = __shfl(incl, 63, 64);
  uint32_t base = 0;
  if (__lane_id() == 0 && total) base = atomicAdd(c, total);
  base = __shfl(base, 0, 64);
  return base + incl - v;
}

// One level of 8-wide nodes [begin, end).
__global__ __launch_bounds__(64) void k_emit(int32_t begin, int32_t end, int32_t* srcOf, const Tree t, int32_t n,
                                             const int32_t* perm, const float4* recs, Node8* nodes, float4* prims,
                                             uint2* kinds, unsigned int* counters) {
  const int32_t k = begin + static_cast<int32_t>(blockIdx.x * 64 + threadIdx.x);
  const bool valid = k < end;
  int32_t ch[8];
  int nc = 0;
  if (valid) {
    const int32_t b = srcOf[k];
    if (b < n) {
      ch[nc++] = b;  // a single-primitive scene: the root holds one leaf slot
    } else {
      int32_t sn[16];
      int si[16];
      int sp = 0;
      const int kk = t.distK[8 * static_cast<size_t>(b) + 7];
      sn[sp] = t.right[b];
      si[sp++] = 8 - kk;
      sn[sp] = t.left[b];
      si[sp++] = kk;
      while (sp > 0) {
        --sp;
        const int32_t c = sn[sp];
        int i = si[sp];
        if (c >= n)
          while (i > 1 && !((t.useDist[c] >> i) & 1u)) --i;
        if (c < n || i <= 1) {
          ch[nc++] = c;
          continue;
        }
        const int k2 = t.distK[8 * static_cast<size_t>(c) + i - 1];
        sn[sp] = t.right[c];
        si[sp++] = i - k2;
        sn[sp] = t.left[c];
        si[sp++] = k2;
      }
    }
  }
  // node box and octant slots (FP64, as the host builder)
  int slotChild[8];
  for (int& v : slotChild) v = -1;
  float4 nlo = make_float4(INFINITY, INFINITY, INFINITY, 0.0f), nhi = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.0f);
  uint32_t nInner = 0, nLeafPrims = 0;
  if (valid) {
    for (int j = 0; j < nc; ++j) {
      nlo = min4(nlo, t.blo[ch[j]]);
      nhi = max4(nhi, t.bhi[ch[j]]);
    }
    const double cc[3] = {0.5 * (static_cast<double>(nlo.x) + nhi.x), 0.5 * (static_cast<double>(nlo.y) + nhi.y),
                          0.5 * (static_cast<double>(nlo.z) + nhi.z)};
    double rel[8][3];
    for (int j = 0; j < nc; ++j) {
      const float4 lo = t.blo[ch[j]], hi = t.bhi[ch[j]];
      rel[j][0] = 0.5 * (static_cast<double>(lo.x) + hi.x) - cc[0];
      rel[j][1] = 0.5 * (static_cast<double>(lo.y) + hi.y) - cc[1];
      rel[j][2] = 0.5 * (static_cast<double>(lo.z) + hi.z) - cc[2];
    }
    uint32_t used = 0;
    for (int r = 0; r < nc; ++r) {
      double bestS = -1e300;
      int bj = -1, bs = -1;
      for (int j = 0; j < nc; ++j) {
        if ((used >> j) & 1u) continue;
        for (int sl = 0; sl < 8; ++sl) {
          if (slotChild[sl] >= 0) continue;
          const double sc = rel[j][0] * ((sl & 1) ? 1.0 : -1.0) + rel[j][1] * ((sl & 2) ? 1.0 : -1.0) +
                            rel[j][2] * ((sl & 4) ? 1.0 : -1.0);
          if (sc > bestS) {
            bestS = sc;
            bj = j;
            bs = sl;
          }
        }
      }
      used |= 1u << bj;
      slotChild[bs] = bj;
    }
    for (int j = 0; j < nc; ++j) {
      const int32_t c = ch[j];
      if (c < n || t.leafSel[c]) nLeafPrims += static_cast<uint32_t>(t.count[c]);
      else ++nInner;
    }
  }
  const uint32_t childBase = waveReserve(&counters[0], nInner);
  const uint32_t primBase = waveReserve(&counters[1], nLeafPrims);
  if (!valid) return;
  Node8 nd;
  const float nl[3] = {nlo.x, nlo.y, nlo.z}, nh[3] = {nhi.x, nhi.y, nhi.z};
  double scale[3];
  for (int a = 0; a < 3; ++a) {
    nd.p[a] = nl[a];
    const double ext = static_cast<double>(nh[a]) - static_cast<double>(nl[a]);
    int e = -126;
    if (ext > 0.0) {
      (void)frexp(ext / 255.0, &e);  // 2^e >= ext / 255
      e = max(-126, min(127, e));
    }
    nd.e[a] = static_cast<uint8_t>(e + 127);
    scale[a] = ldexp(1.0, e);
  }
  nd.imask = 0;
  nd.childBase = childBase;
  nd.primBase = primBase;
  uint32_t inner = 0, off = 0;
  uint2 kw = make_uint2(0u, 0u);
  for (int sl = 0; sl < 8; ++sl) {
    const int j = slotChild[sl];
    nd.meta[sl] = 0;
    for (int a = 0; a < 3; ++a) {
      nd.qlo[a][sl] = 0;
      nd.qhi[a][sl] = 0;
    }
    if (j < 0) continue;
    const int32_t c = ch[j];
    const float4 lo = t.blo[c], hi = t.bhi[c];
    const float cl[3] = {lo.x, lo.y, lo.z}, chh[3] = {hi.x, hi.y, hi.z};
    for (int a = 0; a < 3; ++a) {
      const double ql = floor((static_cast<double>(cl[a]) - nd.p[a]) / scale[a]);
      const double qh = ceil((static_cast<double>(chh[a]) - nd.p[a]) / scale[a]);
      nd.qlo[a][sl] = static_cast<uint8_t>(fmax(0.0, fmin(255.0, ql)));
      nd.qhi[a][sl] = static_cast<uint8_t>(fmax(0.0, fmin(255.0, qh)));
    }
    if (c < n || t.leafSel[c]) {
      const uint32_t cnt = static_cast<uint32_t>(t.count[c]);
      nd.meta[sl] = static_cast<uint8_t>((((1u << cnt) - 1u) << 5) | off);
      // the subtree's primitives, left to right
      int32_t st[8];
      int sp = 0;
      st[sp++] = c;
      uint32_t q = 0;
      while (sp > 0) {
        const int32_t x = st[--sp];
        if (x < n) {
          const int32_t pr = perm[x];
          float4* dst = prims + 3 * static_cast<size_t>(primBase + off + q);
          const float4 r2 = recs[3 * static_cast<size_t>(pr) + 2];
          dst[0] = recs[3 * static_cast<size_t>(pr)];
          dst[1] = recs[3 * static_cast<size_t>(pr) + 1];
          dst[2] = r2;
          const uint32_t kind = __float_as_uint(r2.w) >> kKindShift;
          kw.x |= (kind & 1u) << (off + q);
          kw.y |= (kind >> 1) << (off + q);
          ++q;
        } else {
          st[sp++] = t.right[x];
          st[sp++] = t.left[x];
        }
      }
      off += cnt;
    } else {
      nd.imask |= static_cast<uint8_t>(1u << sl);
      nd.meta[sl] = static_cast<uint8_t>((1u << 5) | (24u + static_cast<uint32_t>(sl)));
      srcOf[childBase + inner] = c;
      ++inner;
    }
  }
  nodes[k] = nd;
  kinds[k] = kw;
}

struct DevMem {
  std::vector<void*> ptrs;
  ~DevMem() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <typename T>
  hipError_t alloc(T** p, size_t count) {
    void* v = nullptr;
    const hipError_t e = hipMalloc(&v, std::max<size_t>(1, count) * sizeof(T));
    if (e == hipSuccess) ptrs.push_back(v);
    *p = static_cast<T*>(v);
    return e;
  }
};

inline unsigned gridFor(int64_t n, int b) { return static_cast<unsigned>((n + b - 1) / b); }

}  // namespace

int deviceBuildBvh8(const float4* recs, const float4* boxes, int64_t n, hipStream_t st, DeviceBvh8& out,
                    std::string& err, int leafMax) {
  leafMax = std::max(1, std::min(kMaxLeaf8, leafMax));
  out = DeviceBvh8();
  if (n < 1 || n >= (int64_t(1) << 28)) {
    err = "device BVH build: primitive count out of range";
    return UBG_ERR_INVALID_ARGUMENT;
  }
  const auto t0 = std::chrono::steady_clock::now();
  DevMem mem;
  hipError_t e = hipSuccess;
  auto ok = [&](hipError_t r, const char* what) {
    if (r != hipSuccess && e == hipSuccess) {
      e = r;
      err = std::string("device BVH build: ") + what + ": " + hipGetErrorString(r);
    }
    return e == hipSuccess;
  };
  const int64_t nNodes = 2 * n - 1;
  Tree t{};
  int* bounds = nullptr;
  uint64_t *keys = nullptr, *keys2 = nullptr;
  int32_t *vals = nullptr, *perm = nullptr, *cl = nullptr, *cl2 = nullptr, *nn = nullptr, *totals = nullptr;
  unsigned long long *packed = nullptr, *pre = nullptr;
  double* sah = nullptr;
  ok(mem.alloc(&t.blo, nNodes), "alloc");
  ok(mem.alloc(&t.bhi, nNodes), "alloc");
  ok(mem.alloc(&t.left, nNodes), "alloc");
  ok(mem.alloc(&t.right, nNodes), "alloc");
  ok(mem.alloc(&t.count, nNodes), "alloc");
  ok(mem.alloc(&t.cost, 8 * nNodes), "alloc");
  ok(mem.alloc(&t.distK, 8 * nNodes), "alloc");
  ok(mem.alloc(&t.useDist, nNodes), "alloc");
  ok(mem.alloc(&t.leafSel, nNodes), "alloc");
  ok(mem.alloc(&bounds, 6), "alloc");
  ok(mem.alloc(&keys, n), "alloc");
  ok(mem.alloc(&keys2, n), "alloc");
  ok(mem.alloc(&vals, n), "alloc");
  ok(mem.alloc(&perm, n), "alloc");
  ok(mem.alloc(&cl, n), "alloc");
  ok(mem.alloc(&cl2, n), "alloc");
  ok(mem.alloc(&nn, n), "alloc");
  ok(mem.alloc(&packed, n), "alloc");
  ok(mem.alloc(&pre, n), "alloc");
  ok(mem.alloc(&totals, 4), "alloc");
  ok(mem.alloc(&sah, 1), "alloc");
  if (e != hipSuccess) return UBG_ERR_OUT_OF_MEMORY;

  // 1. Morton order
  hipLaunchKernelGGL(k_bounds_init, dim3(1), dim3(64), 0, st, bounds);
  hipLaunchKernelGGL(k_bounds, dim3(std::min<unsigned>(1024u, gridFor(n, kBlock))), dim3(kBlock), 0, st, boxes, n,
                     bounds);
  hipLaunchKernelGGL(k_morton, dim3(gridFor(n, kBlock)), dim3(kBlock), 0, st, boxes, n, bounds, keys, vals);
  size_t tmpBytes = 0, scanBytes = 0;
  ok(rocprim::radix_sort_pairs(nullptr, tmpBytes, keys, keys2, vals, perm, static_cast<size_t>(n), 0u, 63u, st),
     "sort size");
  ok(rocprim::exclusive_scan(nullptr, scanBytes, packed, pre, 0ull, static_cast<size_t>(n),
                             rocprim::plus<unsigned long long>(), st),
     "scan size");
  void* tmp = nullptr;
  ok(mem.alloc(reinterpret_cast<uint8_t**>(&tmp), std::max(tmpBytes, scanBytes)), "alloc");
  if (e != hipSuccess) return UBG_ERR_OUT_OF_MEMORY;
  ok(rocprim::radix_sort_pairs(tmp, tmpBytes, keys, keys2, vals, perm, static_cast<size_t>(n), 0u, 63u, st),
     "sort");
  hipLaunchKernelGGL(k_leaves, dim3(gridFor(n, kBlock)), dim3(kBlock), 0, st, boxes, perm, n, t, cl);
  ok(hipGetLastError(), "launch");

  // 2./3. PLOC rounds with the collapse table
  int32_t m = static_cast<int32_t>(n);
  int32_t nodeBase = static_cast<int32_t>(n);
  int32_t hostTot[2] = {0, 0};
  int rounds = 0;
  while (m > 1 && e == hipSuccess) {
    hipLaunchKernelGGL(k_nearest, dim3(gridFor(m, kBlock)), dim3(kBlock), 0, st, cl, m, t, nn);
    hipLaunchKernelGGL(k_flags, dim3(gridFor(m, kBlock)), dim3(kBlock), 0, st, nn, m, packed);
    size_t sb = scanBytes;
    ok(rocprim::exclusive_scan(tmp, sb, packed, pre, 0ull, static_cast<size_t>(m),
                               rocprim::plus<unsigned long long>(), st),
       "scan");
    hipLaunchKernelGGL(k_merge, dim3(gridFor(m, kBlock)), dim3(kBlock), 0, st, cl, nn, m, packed, pre, nodeBase, t,
                       cl2, totals, leafMax);
    ok(hipGetLastError(), "launch");
    ok(hipMemcpyAsync(hostTot, totals, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, st), "copy");
    ok(hipStreamSynchronize(st), "round");
    if (e != hipSuccess) break;
    if (hostTot[1] <= 0 || hostTot[0] >= m) {
      err = "device BVH build: a clustering round merged nothing";
      return UBG_ERR_INTERNAL;
    }
    m = hostTot[0];
    nodeBase += hostTot[1];
    std::swap(cl, cl2);
    ++rounds;
  }
  if (e != hipSuccess) return UBG_ERR_DEVICE;
  int32_t root = 0;
  ok(hipMemcpyAsync(&root, cl, sizeof(int32_t), hipMemcpyDeviceToHost, st), "copy");
  ok(hipMemsetAsync(sah, 0, sizeof(double), st), "memset");
  hipLaunchKernelGGL(k_sah, dim3(std::min<unsigned>(1024u, gridFor(nNodes, kBlock))), dim3(kBlock), 0, st, t,
                     static_cast<int64_t>(nodeBase), sah);
  double sahSum = 0.0;
  ok(hipMemcpyAsync(&sahSum, sah, sizeof(double), hipMemcpyDeviceToHost, st), "copy");
  ok(hipStreamSynchronize(st), "tree");
  if (e != hipSuccess) return UBG_ERR_DEVICE;

  // 4. emission, level by level
  Node8* nodes = nullptr;
  float4* prims = nullptr;
  int32_t* srcOf = nullptr;
  unsigned int* counters = nullptr;
  void* rawNodes = nullptr;
  void* rawPrims = nullptr;
  void* rawKinds = nullptr;
  auto failOut = [&](int status) {
    if (rawNodes) (void)hipFree(rawNodes);
    if (rawPrims) (void)hipFree(rawPrims);
    if (rawKinds) (void)hipFree(rawKinds);
    return status;
  };
  if (!ok(hipMalloc(&rawNodes, std::max<int64_t>(1, n) * sizeof(Node8)), "alloc nodes") ||
      !ok(hipMalloc(&rawPrims, 3 * n * sizeof(float4)), "alloc prims") ||
      !ok(hipMalloc(&rawKinds, std::max<int64_t>(1, n) * sizeof(uint2)), "alloc kinds"))
    return failOut(UBG_ERR_OUT_OF_MEMORY);
  nodes = static_cast<Node8*>(rawNodes);
  prims = static_cast<float4*>(rawPrims);
  if (!ok(mem.alloc(&srcOf, std::max<int64_t>(1, n)), "alloc") || !ok(mem.alloc(&counters, 2), "alloc"))
    return failOut(UBG_ERR_OUT_OF_MEMORY);
  const unsigned int init[2] = {1u, 0u};
  ok(hipMemcpyAsync(srcOf, &root, sizeof(int32_t), hipMemcpyHostToDevice, st), "copy");
  ok(hipMemcpyAsync(counters, init, sizeof(init), hipMemcpyHostToDevice, st), "copy");
  int32_t begin = 0, end = 1;
  int depth = 0;
  unsigned int hostCnt[2] = {1u, 0u};
  while (begin < end && e == hipSuccess) {
    hipLaunchKernelGGL(k_emit, dim3(gridFor(end - begin, 64)), dim3(64), 0, st, begin, end, srcOf, t,
                       static_cast<int32_t>(n), perm, recs, nodes, prims, static_cast<uint2*>(rawKinds), counters);
    ok(hipGetLastError(), "launch");
    ok(hipMemcpyAsync(hostCnt, counters, sizeof(hostCnt), hipMemcpyDeviceToHost, st), "copy");
    ok(hipStreamSynchronize(st), "level");
    ++depth;
    begin = end;
    end = static_cast<int32_t>(hostCnt[0]);
    if (depth > 64) {
      err = "device BVH build: tree deeper than the traversal stack";
      return failOut(UBG_ERR_INTERNAL);
    }
  }
  if (e != hipSuccess) return failOut(UBG_ERR_DEVICE);
  if (static_cast<int64_t>(hostCnt[1]) != n) {
    err = "device BVH build: leaf primitives do not add up";
    return failOut(UBG_ERR_INTERNAL);
  }
  const auto t1 = std::chrono::steady_clock::now();
  out.nodes = rawNodes;
  out.prims = rawPrims;
  out.kinds = rawKinds;
  out.nNodes = end;
  out.depth = depth;
  // the host builder's figure: sum over binary nodes of area / root area
  float4 rlo, rhi;
  (void)hipMemcpy(&rlo, t.blo + root, sizeof(float4), hipMemcpyDeviceToHost);
  (void)hipMemcpy(&rhi, t.bhi + root, sizeof(float4), hipMemcpyDeviceToHost);
  const double dx = rhi.x - rlo.x, dy = rhi.y - rlo.y, dz = rhi.z - rlo.z;
  const double rootArea = std::max(1e-30, 2.0 * (dx * dy + dy * dz + dz * dx));
  out.sahCost = sahSum / rootArea;
  out.seconds = std::chrono::duration<double>(t1 - t0).count();
  (void)rounds;
  return UBG_OK;
}

}  // namespace ubg
