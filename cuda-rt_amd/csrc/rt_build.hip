// LBVH build for gfx950: fused triangle AABB + scene bounds, SoA Morton
// keys with an all-digit histogram, onesweep radix sort, Karras hierarchy,
// bottom-up refit with per-node atomic flags, and the export into the
// reference's node layout.
//
// Reference path: kernels/scene.cu:48-120 (AABBs, bounds),
// kernels/bvh.cu:13-61 (Morton), :267-273 (Thrust sort_by_key),
// :71-207 (Karras hierarchy), :209-258 (refit).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <new>

#include "rt_internal.h"
#include "rt_math.h"

namespace {

using rt::f2ord;
using rt::ord2f;

constexpr int kWave = 64;

__device__ __forceinline__ uint64_t lanemask_lt() {
  const uint32_t lane = __lane_id();
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// ------------------------------------------------------------ init
// Zero the histogram, look-back status and climb arrival counters (one
// contiguous range) and the per-build counters; if init_bounds, set the
// scene-bound accumulators to the reference's initial values (+1e9 for
// the minimum, -1e9 for the maximum, kernels/scene.cu:78,83).
__device__ __forceinline__ void init_body(uint32_t bid, uint32_t nblk, uint32_t* __restrict__ misc, uint32_t* __restrict__ zero, size_t words, int init_bounds) {
  size_t i = bid * (size_t)blockDim.x + threadIdx.x;
  size_t stride = (size_t)nblk * blockDim.x;
  for (size_t k = i; k < words; k += stride) zero[k] = 0u;
  if (i < kMiscWords) {
    if (i < kMiscBounds + 6 * kBoundsReplicas) {
      if (init_bounds) misc[i] = (i % 6) < 3 ? f2ord(1e9f) : f2ord(-1e9f);
    } else if (i != kMiscOverflow && i != kMiscBuildError) {
      misc[i] = 0u;
    }
  }
}
__global__ void k_init(uint32_t* __restrict__ misc, uint32_t* __restrict__ zero, size_t words, int init_bounds) {
  init_body(blockIdx.x, gridDim.x, misc, zero, words, init_bounds);
}

// ------------------------------------------------------------ AABB + bounds
// Per-triangle box (kernels/scene.cu:88-120): min/max of the three vertices
// with fminf/fmaxf. If with_bounds, each workgroup also reduces its boxes
// (wave shuffles, then LDS) and adds one ordered-integer atomic per axis to
// one of kBoundsReplicas accumulator sets (kernels/scene.cu:72-85). NaN
// coordinates are ignored, as fminf/fmaxf ignore them. trirec receives each
// triangle's leaf record in triangle order (the climb then copies one
// record per sorted leaf instead of gathering index and vertices).
__device__ __forceinline__ void aabb_bounds_body(uint32_t bid, uint32_t nblk, const float* __restrict__ vx, const float* __restrict__ vy,
                                                     const float* __restrict__ vz, const int32_t* __restrict__ idx,
                                                     int n, float* __restrict__ tri_box, float4* __restrict__ trirec,
                                                     uint32_t* __restrict__ misc, int with_bounds) {
  float acc[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (int i = bid * 256 + threadIdx.x; i < n; i += nblk * 256) {
    int a = idx[3 * i + 0], b = idx[3 * i + 1], c = idx[3 * i + 2];
    float ax = vx[a], ay = vy[a], az = vz[a];
    float bx = vx[b], by = vy[b], bz = vz[b];
    float cx = vx[c], cy = vy[c], cz = vz[c];
    float bb[6] = {fminf(cx, fminf(ax, bx)), fminf(cy, fminf(ay, by)), fminf(cz, fminf(az, bz)),
                   fmaxf(cx, fmaxf(ax, bx)), fmaxf(cy, fmaxf(ay, by)), fmaxf(cz, fmaxf(az, bz))};
    float* o = tri_box + 6 * (size_t)i;
#pragma unroll
    for (int k = 0; k < 6; ++k) o[k] = bb[k];
    float4* L = trirec + 3 * (size_t)i;
    L[0] = make_float4(ax, ay, az, bx);
    L[1] = make_float4(by, bz, cx, cy);
    L[2] = make_float4(cz, __int_as_float(i), 0.f, 0.f);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      acc[k] = fminf(acc[k], bb[k]);
      acc[3 + k] = fmaxf(acc[3 + k], bb[3 + k]);
    }
  }
  if (!with_bounds) return;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int k = 0; k < 3; ++k) acc[k] = fminf(acc[k], __shfl_xor(acc[k], off));
#pragma unroll
    for (int k = 3; k < 6; ++k) acc[k] = fmaxf(acc[k], __shfl_xor(acc[k], off));
  }
  __shared__ float s_acc[4][6];
  int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int k = 0; k < 6; ++k) s_acc[w][k] = acc[k];
  __syncthreads();
  if (threadIdx.x < 6) {
    int k = threadIdx.x;
    float v = s_acc[0][k];
    for (int j = 1; j < 4; ++j) v = (k < 3) ? fminf(v, s_acc[j][k]) : fmaxf(v, s_acc[j][k]);
    uint32_t* dst = misc + kMiscBounds + 6 * (bid % kBoundsReplicas) + k;
    if (k < 3) {
      if (v < INFINITY) atomicMin(dst, f2ord(v));
    } else {
      if (v > -INFINITY) atomicMax(dst, f2ord(v));
    }
  }
}
__global__ __launch_bounds__(256) void k_aabb_bounds(const float* __restrict__ vx, const float* __restrict__ vy,
                                                     const float* __restrict__ vz, const int32_t* __restrict__ idx,
                                                     int n, float* __restrict__ tri_box, float4* __restrict__ trirec,
                                                     uint32_t* __restrict__ misc, int with_bounds) {
  aabb_bounds_body(blockIdx.x, gridDim.x, vx, vy, vz, idx, n, tri_box, trirec, misc, with_bounds);
}

// Scene bounds from the accumulator replicas.
__device__ __forceinline__ void load_bounds(const uint32_t* __restrict__ misc, double* mn, double* mx) {
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    uint32_t a = misc[kMiscBounds + k], b = misc[kMiscBounds + 3 + k];
    for (int r = 1; r < kBoundsReplicas; ++r) {
      a = min(a, misc[kMiscBounds + 6 * r + k]);
      b = max(b, misc[kMiscBounds + 6 * r + 3 + k]);
    }
    mn[k] = (double)ord2f(a);
    mx[k] = (double)ord2f(b);
  }
}

// ------------------------------------------------------------ Morton
__device__ __forceinline__ uint32_t spread10(uint32_t x) {
  x &= 0x3ffu;
  x = (x | (x << 16)) & 0x030000ffu;
  x = (x | (x << 8)) & 0x0300f00fu;
  x = (x | (x << 4)) & 0x030c30c3u;
  x = (x | (x << 2)) & 0x09249249u;
  return x;
}
__device__ __forceinline__ uint64_t spread21(uint64_t x) {
  x &= 0x1fffffull;
  x = (x | (x << 32)) & 0x001f00000000ffffull;
  x = (x | (x << 16)) & 0x001f0000ff0000ffull;
  x = (x | (x << 8)) & 0x100f00f00f00f00full;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}

template <typename K>
struct KeyTraits;
template <>
struct KeyTraits<uint32_t> {
  static constexpr int kPasses = 4;
  static constexpr double kScale = 1024.0, kMax = 1023.0;
  __device__ static uint32_t make(uint32_t qx, uint32_t qy, uint32_t qz) {
    return spread10(qx) * 4u + spread10(qy) * 2u + spread10(qz);
  }
  __device__ static int delta(uint32_t a, uint32_t b, int i, int j) {
    return a != b ? __clz(a ^ b) : 32 + __clz((uint32_t)(i ^ j));
  }
};
template <>
struct KeyTraits<uint64_t> {
  static constexpr int kPasses = 8;
  static constexpr double kScale = 2097152.0, kMax = 2097151.0;
  __device__ static uint64_t make(uint32_t qx, uint32_t qy, uint32_t qz) {
    return (spread21(qx) << 2) | (spread21(qy) << 1) | spread21(qz);
  }
  __device__ static int delta(uint64_t a, uint64_t b, int i, int j) {
    return a != b ? __clzll(a ^ b) : 64 + __clz((uint32_t)(i ^ j));
  }
};

// kernels/bvh.cu:22-32,52-58: centroid in FP64, (c - min)/(max - min),
// scaled, clamped with fmax/fmin (NaN -> 0), rounded to float, truncated.
// Each workgroup also builds the digit histograms of every sort pass in
// LDS and adds them to the global histogram once.
// Bounds handed to BVH_d::computeMortonCodes(mMin, mMax) by the caller
// (kernels/bvh.cu:260-265); when !use the keys are normalised by the scene
// bounds the build reduced into misc[].
struct GivenBounds {
  double mn[3], mx[3];
  int use;
};

template <typename K>
__device__ __forceinline__ void morton_body(uint32_t bid, uint32_t nblk, const float* __restrict__ tri_box, int n,
                                                const uint32_t* __restrict__ misc, K* __restrict__ keys,
                                                uint32_t* __restrict__ hist, int per_block,
                                                const GivenBounds& gb) {
  constexpr int P = KeyTraits<K>::kPasses;
  __shared__ uint32_t s_hist[P * 256];
  for (int k = threadIdx.x; k < P * 256; k += 256) s_hist[k] = 0u;
  double mn[3], mx[3], ext[3];
  if (gb.use) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      mn[k] = gb.mn[k];
      mx[k] = gb.mx[k];
    }
  } else {
    load_bounds(misc, mn, mx);
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) ext[k] = mx[k] - mn[k];
  __syncthreads();
  int begin = bid * per_block;
  int end = min(n, begin + per_block);
  for (int i = begin + threadIdx.x; i < end; i += 256) {
    const float* bx = tri_box + 6 * (size_t)i;
    uint32_t q[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      double c = ((double)bx[k] + (double)bx[3 + k]) * 0.5;
      double u = (c - mn[k]) / ext[k];
      float f = (float)fmin(fmax(u * KeyTraits<K>::kScale, 0.0), KeyTraits<K>::kMax);
      q[k] = (uint32_t)f;
    }
    K key = KeyTraits<K>::make(q[0], q[1], q[2]);
    keys[i] = key;
#pragma unroll
    for (int p = 0; p < P; ++p) atomicAdd(&s_hist[p * 256 + (int)((key >> (8 * p)) & 255u)], 1u);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < P * 256; k += 256) {
    uint32_t v = s_hist[k];
    if (v) atomicAdd(&hist[k], v);
  }
}
template <typename K>
__global__ __launch_bounds__(256) void k_morton(const float* __restrict__ tri_box, int n,
                                                const uint32_t* __restrict__ misc, K* __restrict__ keys,
                                                uint32_t* __restrict__ hist, int per_block, GivenBounds gb) {
  morton_body(blockIdx.x, gridDim.x, tri_box, n, misc, keys, hist, per_block, gb);
}

// ------------------------------------------------------------ onesweep sort
// One pass of a stable LSD radix sort over 8-bit digits (Adinets & Merrill,
// "Onesweep", 2022): tiles are taken in launch order from an atomic
// counter; each tile ranks its keys with wave64 ballot matching, publishes
// its per-digit counts, and finds its global offsets by decoupled look-back
// over the preceding tiles. Keys are staged in LDS in digit order so the
// global scatter is mostly contiguous. Ties keep input order (stable), as
// thrust::sort_by_key does (kernels/bvh.cu:272).
constexpr uint32_t kFlagAgg = 1u << 30;
constexpr uint32_t kFlagPrefix = 2u << 30;
constexpr uint32_t kValMask = (1u << 30) - 1u;

// Exclusive scan of one value per thread across a 256-thread workgroup.
__device__ __forceinline__ uint32_t block_excl_scan256(uint32_t v, uint32_t* s_wsum) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t y = __shfl_up(x, off);
    if (lane >= off) x += y;
  }
  if (lane == 63) s_wsum[w] = x;
  __syncthreads();
  uint32_t pre = 0;
  for (int j = 0; j < w; ++j) pre += s_wsum[j];
  __syncthreads();
  return pre + x - v;
}

template <typename K>
__device__ __forceinline__ void onesweep_body(uint32_t bid, uint32_t nblk, const K* __restrict__ kin, const uint32_t* __restrict__ iin,
                                                           K* __restrict__ kout, uint32_t* __restrict__ iout, int n,
                                                           int shift, const uint32_t* __restrict__ hist,
                                                           uint32_t* status, uint32_t* tile_ctr) {
  constexpr int KPT = kSortKeysPerThread;
  constexpr int NW = kSortThreads / kWave;
  __shared__ uint32_t s_tile;
  __shared__ uint32_t s_wcnt[NW][256];
  __shared__ uint32_t s_lstart[256];
  __shared__ uint32_t s_gbase[256];
  __shared__ uint32_t s_wsum[NW];
  __shared__ K s_keys[kSortTile];
  __shared__ uint32_t s_ids[kSortTile];

  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  if (tid == 0) s_tile = atomicAdd(tile_ctr, 1u);
#pragma unroll
  for (int j = 0; j < NW; ++j) s_wcnt[j][tid] = 0u;
  __syncthreads();
  const int tile = (int)s_tile;
  const int base = tile * kSortTile;
  const int wbase = base + w * kWave * KPT;

  K key[KPT];
  uint32_t id[KPT];
  uint32_t rank[KPT];
#pragma unroll
  for (int k = 0; k < KPT; ++k) {
    int i = wbase + k * kWave + lane;
    bool valid = i < n;
    key[k] = valid ? kin[i] : (K)0;
    id[k] = valid ? (iin ? iin[i] : (uint32_t)i) : 0u;
  }
  const uint64_t mlt = lanemask_lt();
#pragma unroll
  for (int k = 0; k < KPT; ++k) {
    int i = wbase + k * kWave + lane;
    bool valid = i < n;
    uint32_t d = (uint32_t)(key[k] >> shift) & 255u;
    uint64_t m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      bool bit = (d >> b) & 1u;
      uint64_t bal = __ballot(bit);
      m &= bit ? bal : ~bal;
    }
    uint32_t r = (uint32_t)__popcll(m & mlt);
    uint32_t cnt = valid ? s_wcnt[w][d] : 0u;
    rank[k] = cnt + r;
    if (valid && r == 0) s_wcnt[w][d] = cnt + (uint32_t)__popcll(m);
  }
  __syncthreads();

  // Per digit (thread = digit): wave-exclusive prefixes and tile total.
  uint32_t tot = 0;
#pragma unroll
  for (int j = 0; j < NW; ++j) {
    uint32_t c = s_wcnt[j][tid];
    s_wcnt[j][tid] = tot;
    tot += c;
  }
  uint32_t* my = status + (size_t)tile * 256 + tid;
  __hip_atomic_store(my, (tile == 0 ? kFlagPrefix : kFlagAgg) | tot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  uint32_t lstart = block_excl_scan256(tot, s_wsum);
  uint32_t gex = block_excl_scan256(hist[tid], s_wsum);
  uint32_t excl = 0;
  if (tile > 0) {
    int j = tile - 1;
    while (true) {
      uint32_t v = __hip_atomic_load(status + (size_t)j * 256 + tid, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if ((v & ~kValMask) == 0u) {
        __builtin_amdgcn_s_sleep(1);
        continue;
      }
      excl += v & kValMask;
      if (v & kFlagPrefix) break;
      --j;
    }
    __hip_atomic_store(my, kFlagPrefix | (excl + tot), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  s_lstart[tid] = lstart;
  s_gbase[tid] = gex + excl;
  __syncthreads();

#pragma unroll
  for (int k = 0; k < KPT; ++k) {
    int i = wbase + k * kWave + lane;
    if (i < n) {
      uint32_t d = (uint32_t)(key[k] >> shift) & 255u;
      uint32_t pos = s_lstart[d] + s_wcnt[w][d] + rank[k];
      s_keys[pos] = key[k];
      s_ids[pos] = id[k];
    }
  }
  __syncthreads();
  const int cnt = min(kSortTile, n - base);
  for (int j = tid; j < cnt; j += kSortThreads) {
    K kk = s_keys[j];
    uint32_t d = (uint32_t)(kk >> shift) & 255u;
    uint32_t g = s_gbase[d] + (uint32_t)j - s_lstart[d];
    kout[g] = kk;
    iout[g] = s_ids[j];
  }
}
template <typename K>
__global__ __launch_bounds__(kSortThreads) void k_onesweep(const K* __restrict__ kin, const uint32_t* __restrict__ iin,
                                                           K* __restrict__ kout, uint32_t* __restrict__ iout, int n,
                                                           int shift, const uint32_t* __restrict__ hist,
                                                           uint32_t* status, uint32_t* tile_ctr) {
  onesweep_body(blockIdx.x, gridDim.x, kin, iin, kout, iout, n, shift, hist, status, tile_ctr);
}

// ------------------------------------------------------------ hierarchy
// kernels/bvh.cu:71-207. One thread per sorted position: leaf setup
// (node = triangle id, box = the triangle's box, traversal record = its
// vertices) and, for positions < n-1, the internal node's span (doubling
// then binary search) and split (bisection), children and parent links.
template <typename K>
__device__ __forceinline__ int kdelta(const K* __restrict__ keys, int n, int i, K ki, int j) {
  if (j < 0 || j >= n) return -1;
  return KeyTraits<K>::delta(ki, keys[j], i, j);
}
template <typename K>
__device__ __forceinline__ int kdelta2(const K* __restrict__ keys, int n, int i, int j) {
  int l = min(i, j), r = max(i, j);
  if (l < 0 || r >= n) return -1;
  return KeyTraits<K>::delta(keys[l], keys[r], l, r);
}

template <typename K>
__global__ __launch_bounds__(256) void k_hierarchy(const K* __restrict__ keys, const uint32_t* __restrict__ ids, int n,
                                                   const float* __restrict__ tri_box, const float* __restrict__ vx,
                                                   const float* __restrict__ vy, const float* __restrict__ vz,
                                                   const int32_t* __restrict__ idx, int32_t* __restrict__ parent,
                                                   int32_t* __restrict__ left, int32_t* __restrict__ right,
                                                   uint32_t* __restrict__ flags, float* __restrict__ node_box,
                                                   float4* __restrict__ ltri, int32_t* __restrict__ leaf_of) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  {
    int tri = (int)ids[i];
    const float* src = tri_box + 6 * (size_t)tri;
    float* dst = node_box + 6 * (size_t)(n - 1 + i);
#pragma unroll
    for (int k = 0; k < 6; ++k) dst[k] = src[k];
    int a = idx[3 * tri + 0], b = idx[3 * tri + 1], c = idx[3 * tri + 2];
    float4* L = ltri + kLeafStride * (size_t)i;
    L[0] = make_float4(vx[a], vy[a], vz[a], vx[b]);
    L[1] = make_float4(vy[b], vz[b], vx[c], vy[c]);
    L[2] = make_float4(vz[c], __int_as_float(tri), 0.f, 0.f);
    leaf_of[tri] = i;
  }
  if (i == 0) parent[0] = -1;
  if (i >= n - 1) return;
  flags[i] = 0u;
  const K ki = keys[i];
  // FindSpan (kernels/bvh.cu:100-133)
  int dr = kdelta<K>(keys, n, i, ki, i + 1), dl = kdelta<K>(keys, n, i, ki, i - 1);
  int d = (dr - dl) > 0 ? 1 : ((dr - dl) < 0 ? -1 : 0);
  int dmin = kdelta<K>(keys, n, i, ki, i - d);
  int lmax = 2;
  while (kdelta<K>(keys, n, i, ki, i + lmax * d) > dmin) lmax *= 2;
  int l = 0, t = lmax;
  do {
    t /= 2;
    if (kdelta<K>(keys, n, i, ki, i + (l + t) * d) > dmin) l += t;
  } while (t > 1);
  int first = min(i, i + l * d), last = max(i, i + l * d);
  // FindSplit (kernels/bvh.cu:134-162)
  int common = kdelta2<K>(keys, n, first, last);
  int lo = first, hi = last;
  do {
    int mid = (hi + lo) / 2;
    if (kdelta2<K>(keys, n, lo, mid) > common)
      lo = mid;
    else
      hi = mid;
  } while (hi > lo + 1);
  int s = lo;
  int c1 = (s == first) ? (n - 1 + s) : s;
  int c2 = (s + 1 == last) ? (n - 1 + s + 1) : (s + 1);
  left[i] = c1;
  right[i] = c2;
  parent[c1] = i;
  parent[c2] = i;
}

// ------------------------------------------------------------ refit
__device__ __forceinline__ int child_ref(int c, int n) { return c >= n - 1 ? ~(c - (n - 1)) : c; }
// kernels/bvh.cu:209-258 with the missing release/acquire made explicit:
// each leaf climbs; the first arrival at a node leaves, the second unions
// the two child boxes (fminf/fmaxf), writes the node box and the node's
// 64-byte traversal record, and continues until the root.
__global__ __launch_bounds__(256) void k_refit(int n, const int32_t* __restrict__ parent,
                                               const int32_t* __restrict__ left, const int32_t* __restrict__ right,
                                               uint32_t* flags, float* node_box, float4* __restrict__ tnode) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n || n < 2) return;
  int idx = n - 1 + i;
  float b[6];
  {
    const float* src = node_box + 6 * (size_t)idx;
#pragma unroll
    for (int k = 0; k < 6; ++k) b[k] = src[k];
  }
  while (true) {
    int p = parent[idx];
    uint32_t old = __hip_atomic_fetch_add(&flags[p], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (old == 0u) break;
    int l = left[p], r = right[p];
    int sib = (l == idx) ? r : l;
    float o[6];
    const float* src = node_box + 6 * (size_t)sib;
#pragma unroll
    for (int k = 0; k < 6; ++k) o[k] = src[k];
    const float* lb = (l == idx) ? b : o;
    const float* rb = (l == idx) ? o : b;
    float nb[6];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      nb[k] = fminf(lb[k], rb[k]);
      nb[3 + k] = fmaxf(lb[3 + k], rb[3 + k]);
    }
    float* dst = node_box + 6 * (size_t)p;
#pragma unroll
    for (int k = 0; k < 6; ++k) dst[k] = nb[k];
    float4* T = tnode + 4 * (size_t)p;
    T[0] = make_float4(lb[0], lb[1], lb[2], lb[3]);
    T[1] = make_float4(lb[4], lb[5], rb[0], rb[1]);
    T[2] = make_float4(rb[2], rb[3], rb[4], rb[5]);
    T[3] = make_float4(__int_as_float(child_ref(l, n)), __int_as_float(child_ref(r, n)), 0.f, 0.f);
#pragma unroll
    for (int k = 0; k < 6; ++k) b[k] = nb[k];
    idx = p;
    if (p == 0) break;
  }
}

// Triangle boxes for a refit: same arithmetic as k_aabb_bounds without the
// scene bounds; the leaf copies and leaf records are refreshed in place.
__global__ __launch_bounds__(256) void k_refit_leaves(const float* __restrict__ vx, const float* __restrict__ vy,
                                                      const float* __restrict__ vz, const int32_t* __restrict__ idx,
                                                      const uint32_t* __restrict__ ids, int n,
                                                      float* __restrict__ tri_box, float* __restrict__ node_box,
                                                      float4* __restrict__ ltri, uint32_t* __restrict__ flags) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  if (i < n - 1) flags[i] = 0u;
  int tri = (int)ids[i];
  int a = idx[3 * tri + 0], b = idx[3 * tri + 1], c = idx[3 * tri + 2];
  float ax = vx[a], ay = vy[a], az = vz[a];
  float bx = vx[b], by = vy[b], bz = vz[b];
  float cx = vx[c], cy = vy[c], cz = vz[c];
  float bb[6] = {fminf(cx, fminf(ax, bx)), fminf(cy, fminf(ay, by)), fminf(cz, fminf(az, bz)),
                 fmaxf(cx, fmaxf(ax, bx)), fmaxf(cy, fmaxf(ay, by)), fmaxf(cz, fmaxf(az, bz))};
  float* t = tri_box + 6 * (size_t)tri;
  float* d = node_box + 6 * (size_t)(n - 1 + i);
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    t[k] = bb[k];
    d[k] = bb[k];
  }
  float4* L = ltri + kLeafStride * (size_t)i;
  L[0] = make_float4(ax, ay, az, bx);
  L[1] = make_float4(by, bz, cx, cy);
  L[2] = make_float4(cz, __int_as_float(tri), 0.f, 0.f);
}

// The quantized planes of a 4-wide node (q[0..2]): corner = the node's box
// minimum nb[0..2]; per axis the smallest power-of-two step with 255 steps
// covering the box (stored as a float, 0 for a flat axis); per child slot j
// < k the 8-bit planes rounded outward, so the decoded box contains the
// child's box; empty slots decode as empty (lo 255 above hi 0).
__device__ __forceinline__ void quantize4(const float (&nb)[6], const float (&lo)[3][4], const float (&hi)[3][4],
                                          int k, float4* __restrict__ q) {
  uint32_t ebits = 0, qlo[3] = {0, 0, 0}, qhi[3] = {0, 0, 0};
  for (int a = 0; a < 3; ++a) {
    const double o = nb[a], ext = (double)nb[3 + a] - o;
    int e = 0;  // biased exponent of the step; 0 = zero extent
    if (ext > 0) {
      int ex2;
      frexp(ext / 255.0, &ex2);  // ext/255 < 2^ex2
      ex2 = ex2 < -126 ? -126 : (ex2 > 127 ? 127 : ex2);  // an infinite extent leaves ex2 unspecified
      e = ex2 + 127;
    }
    ebits |= (uint32_t)e << (8 * a);
    const double step = e ? ldexp(1.0, e - 127) : 0.0;
    // step is a power of two (2^-126 at least, far below DBL_MAX), so
    // x / step == x * inv exactly
    const double inv = e ? ldexp(1.0, 127 - e) : 0.0;
    for (int j = 0; j < 4; ++j) {
      uint32_t ql = 255, qh = 0;  // empty slot: lo above hi
      if (j < k) {
        if (step == 0) {
          ql = 0;
          qh = 0;
        } else {
          double l = floor((lo[a][j] - o) * inv), h = ceil((hi[a][j] - o) * inv);
          // within [0, 255] before the adjustment (NaN from non-finite
          // coordinates spans the whole node), so each loop runs at most
          // 255 times whatever the box
          l = l >= 0 ? (l > 255 ? 255 : l) : 0;
          h = h <= 255 ? (h < 0 ? 0 : h) : 255;
          while (l > 0 && o + l * step > lo[a][j]) l -= 1;
          while (h < 255 && o + h * step < hi[a][j]) h += 1;
          ql = (uint32_t)(l < 0 ? 0 : (l > 255 ? 255 : l));
          qh = (uint32_t)(h < 0 ? 0 : (h > 255 ? 255 : h));
        }
      }
      qlo[a] |= ql << (8 * j);
      qhi[a] |= qh << (8 * j);
    }
  }
  // the steps as floats (2^(e-127), or 0), one in each spare word
  q[0] = make_float4(nb[0], nb[1], nb[2], __uint_as_float((ebits & 0xffu) << 23));
  q[1] = make_float4(__uint_as_float(qlo[0]), __uint_as_float(qlo[1]), __uint_as_float(qlo[2]),
                     __uint_as_float(((ebits >> 8) & 0xffu) << 23));
  q[2] = make_float4(__uint_as_float(qhi[0]), __uint_as_float(qhi[1]), __uint_as_float(qhi[2]),
                     __uint_as_float(((ebits >> 16) & 0xffu) << 23));
}

// ------------------------------------------------------------ fused climb
// Hierarchy and refit in one bottom-up pass (the agglomerative LBVH
// construction of Apetrei, "Fast and Simple Agglomerative LBVH
// Construction", 2014), numbered exactly like the reference's Karras tree
// (kernels/bvh.cu:164-207): node [l, r]'s parent splits at r (the node is
// its left child) when delta(r, r+1) > delta(l-1, l), else at l-1 (right
// child); delta is the reference's prefix length with the index fallback
// (kernels/bvh.cu:72-88). A left child takes index r, a right child index l,
// the root [0, n-1] index 0 -- the numbering Karras's span/split search
// produces. Each split position is shared by exactly two arrivals; the
// second one unions the boxes (fminf/fmaxf, kernels/bvh.cu:63-69) and goes
// on. The per-split arrival counter plays the role of the reference's
// per-node atomicCAS flag (kernels/bvh.cu:228).
//
// One workgroup handles a chunk of kClimbChunk sorted leaves. Each arriving
// child first writes a 32-byte deposit (Karras index, far end of its range,
// box) for its side of the split, then counts its arrival; for splits inside
// the chunk both live in LDS (workgroup scope, no cache maintenance). A node
// whose parent split lies on or past the chunk's edge is handed to the
// top-level passes (k_climb_pass, then k_climb_top) with a global deposit,
// and so is every in-chunk split still waiting for its sibling when the
// chunk is done. Those passes count arrivals with global counters, each
// reading only deposits written by earlier launches, so no agent-scope
// release/acquire is needed anywhere.
template <typename K>
__device__ __forceinline__ int climb_delta(const K* __restrict__ keys, int n, int i, int j) {
  if (i < 0 || j >= n) return -1;
  return KeyTraits<K>::delta(keys[i], keys[j], i, j);
}

struct ClimbNode {
  int l, r, idx;        // sorted-leaf range and Karras index (idx < 0: not yet written)
  int lidx, ridx;       // children (internal nodes only)
  float b[6], lb[6], rb[6];
};

__device__ __forceinline__ void write_node(const ClimbNode& c, int idx, int n, int32_t* __restrict__ parent,
                                           int32_t* __restrict__ left, int32_t* __restrict__ right,
                                           float* __restrict__ node_box, float4* __restrict__ tnode) {
  float* dst = node_box + 6 * (size_t)idx;
#pragma unroll
  for (int k = 0; k < 6; ++k) dst[k] = c.b[k];
  if (tnode) {
    float4* T = tnode + 4 * (size_t)idx;
    T[0] = make_float4(c.lb[0], c.lb[1], c.lb[2], c.lb[3]);
    T[1] = make_float4(c.lb[4], c.lb[5], c.rb[0], c.rb[1]);
    T[2] = make_float4(c.rb[2], c.rb[3], c.rb[4], c.rb[5]);
    T[3] = make_float4(__int_as_float(child_ref(c.lidx, n)), __int_as_float(child_ref(c.ridx, n)), 0.f, 0.f);
  }
  if (left) {  // nullptr: a refit, whose topology is already in place
    left[idx] = c.lidx;
    right[idx] = c.ridx;
    parent[c.lidx] = idx;
    parent[c.ridx] = idx;
  }
}

// Merge the arriving child c with the sibling's deposit; c becomes the parent.
__device__ __forceinline__ void merge_deposit(ClimbNode& c, bool is_left, const float4* __restrict__ dep) {
  float4 o0 = dep[0], o1 = dep[1];
  int oidx = __float_as_int(o0.x), ofar = __float_as_int(o0.y);
  float ob[6] = {o0.z, o0.w, o1.x, o1.y, o1.z, o1.w};
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    c.lb[k] = is_left ? c.b[k] : ob[k];
    c.rb[k] = is_left ? ob[k] : c.b[k];
  }
  c.lidx = is_left ? c.idx : oidx;
  c.ridx = is_left ? oidx : c.idx;
  if (is_left)
    c.r = ofar;
  else
    c.l = ofar;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    c.b[k] = fminf(c.lb[k], c.rb[k]);
    c.b[3 + k] = fmaxf(c.lb[3 + k], c.rb[3 + k]);
  }
  c.idx = -1;
}

// Decide the parent split of c; write c's node record if it is a new
// internal node; write c's deposit. Returns the split; *side = 0 if c is
// the left child.
template <typename K>
__device__ __forceinline__ int climb_step(ClimbNode& c, const K* __restrict__ keys, int n, int* side,
                                          int32_t* __restrict__ parent, int32_t* __restrict__ left,
                                          int32_t* __restrict__ right, float* __restrict__ node_box,
                                          float4* __restrict__ tnode, float4* __restrict__ deposits) {
  int dl = climb_delta<K>(keys, n, c.l - 1, c.l);
  int dr = climb_delta<K>(keys, n, c.r, c.r + 1);
  bool is_left = dr > dl;
  int split = is_left ? c.r : c.l - 1;
  if (split < 0 || split >= n - 1) return -1;  // cannot happen for a consistent climb
  if (c.idx < 0) {
    c.idx = is_left ? c.r : c.l;
    write_node(c, c.idx, n, parent, left, right, node_box, tnode);
  }
  float4* d = deposits + 2 * (2 * (size_t)split + (is_left ? 0 : 1));
  d[0] = make_float4(__int_as_float(c.idx), __int_as_float(is_left ? c.l : c.r), c.b[0], c.b[1]);
  d[1] = make_float4(c.b[2], c.b[3], c.b[4], c.b[5]);
  *side = is_left ? 0 : 1;
  return split;
}

// Leaf record and box of sorted leaf i: copied from the triangle-order
// record (trirec, a build) or gathered through the indices (trirec ==
// nullptr, a refit, whose vertices moved). The box is formed from the three
// vertices exactly as k_aabb_bounds forms it.
__device__ __forceinline__ void leaf_record(int i, int tri, const float4* __restrict__ trirec,
                                            const float* __restrict__ vx, const float* __restrict__ vy,
                                            const float* __restrict__ vz, const int32_t* __restrict__ idx,
                                            float4 (&q)[3], float (&bb)[6]) {
  if (trirec) {
    const float4* R = trirec + 3 * (size_t)tri;
    q[0] = R[0];
    q[1] = R[1];
    q[2] = R[2];
  } else {
    const int a = idx[3 * (size_t)tri], b = idx[3 * (size_t)tri + 1], c = idx[3 * (size_t)tri + 2];
    q[0] = make_float4(vx[a], vy[a], vz[a], vx[b]);
    q[1] = make_float4(vy[b], vz[b], vx[c], vy[c]);
    q[2] = make_float4(vz[c], __int_as_float(tri), 0.f, 0.f);
  }
  const float ax = q[0].x, ay = q[0].y, az = q[0].z, bx = q[0].w, by = q[1].x, bz = q[1].y, cx = q[1].z,
              cy = q[1].w, cz = q[2].x;
  bb[0] = fminf(cx, fminf(ax, bx));
  bb[1] = fminf(cy, fminf(ay, by));
  bb[2] = fminf(cz, fminf(az, bz));
  bb[3] = fmaxf(cx, fmaxf(ax, bx));
  bb[4] = fmaxf(cy, fmaxf(ay, by));
  bb[5] = fmaxf(cz, fmaxf(az, bz));
  (void)i;
}

// The chunk kernel keeps everything its climbers exchange in LDS: the
// chunk's keys (plus one on each side) for the split decisions, and for
// every split inside the chunk an arrival count and one 32-byte deposit per
// side. Only a split left with a single arrival when the chunk is done
// publishes its deposit and count to global memory, for the top-level
// passes; the rest of the climb touches global memory only for its outputs.
template <typename K>
struct ClimbLds {
  K key[kClimbChunk + 2];            // keys[s0 - 1 .. e0]
  uint32_t arr[kClimbChunk];         // split s0 + k: arrivals (low byte) + 0x100 per right-child arrival
  float4 dep[4 * kClimbChunk];       // split s0 + k, side: dep[2 * (2k + side)], dep[2 * (2k + side) + 1]
};

template <typename K>
__device__ __forceinline__ int lds_delta(const ClimbLds<K>& S, int s0, int n, int i, int j) {
  if (i < 0 || j >= n) return -1;
  return KeyTraits<K>::delta(S.key[i - s0 + 1], S.key[j - s0 + 1], i, j);
}

template <typename K>
__device__ __forceinline__ void climb_chunks_body(uint32_t bid, uint32_t nblk, 
    const K* __restrict__ keys, const uint32_t* __restrict__ ids, int n, const float4* __restrict__ trirec,
    const float* __restrict__ vx, const float* __restrict__ vy, const float* __restrict__ vz,
    const int32_t* __restrict__ idx, int32_t* __restrict__ parent, int32_t* __restrict__ left,
    int32_t* __restrict__ right, float* __restrict__ node_box, float4* __restrict__ tnode, float4* __restrict__ ltri,
    float4* __restrict__ deposits, uint32_t* __restrict__ arrivals, uint32_t* __restrict__ deferred,
    uint32_t* __restrict__ misc, int32_t* __restrict__ leaf_of) {
  __shared__ ClimbLds<K> S;
  const int s0 = bid * kClimbChunk;
  const int e0 = min(n, s0 + kClimbChunk);
  for (int k = threadIdx.x; k < kClimbChunk; k += kClimbThreads) S.arr[k] = 0u;
  for (int k = threadIdx.x; k < e0 - s0 + 2; k += kClimbThreads) {
    const int g = s0 - 1 + k;
    if (g >= 0 && g < n) S.key[k] = keys[g];
  }
  if (s0 == 0 && threadIdx.x == 0 && parent) parent[0] = -1;
  __syncthreads();
  for (int i = s0 + threadIdx.x; i < e0; i += kClimbThreads) {
    // leaf n-1+i: its record in sorted order and its box
    const int tri = (int)ids[i];
    float4 q[3];
    ClimbNode c;
    leaf_record(i, tri, trirec, vx, vy, vz, idx, q, c.b);
    float4* L = ltri + kLeafStride * (size_t)i;
    L[0] = q[0];
    L[1] = q[1];
    L[2] = q[2];
    float* dst = node_box + 6 * (size_t)(n - 1 + i);
#pragma unroll
    for (int k = 0; k < 6; ++k) dst[k] = c.b[k];
    if (leaf_of) leaf_of[tri] = i;  // only a build reorders the leaves
    if (n == 1) continue;
    c.l = c.r = i;
    c.idx = n - 1 + i;
    c.lidx = c.ridx = 0;
    for (int guard = 0;; ++guard) {
      const bool is_left = lds_delta(S, s0, n, c.r, c.r + 1) > lds_delta(S, s0, n, c.l - 1, c.l);
      const int split = is_left ? c.r : c.l - 1;
      if (split < 0 || split >= n - 1 || guard >= n) {  // cannot happen for a consistent climb
        atomicAdd(&misc[kMiscBuildError], 1u);
        break;
      }
      if (c.idx < 0) {
        c.idx = is_left ? c.r : c.l;
        write_node(c, c.idx, n, parent, left, right, node_box, tnode);
      }
      const int side = is_left ? 0 : 1;
      const float4 d0 = make_float4(__int_as_float(c.idx), __int_as_float(is_left ? c.l : c.r), c.b[0], c.b[1]);
      const float4 d1 = make_float4(c.b[2], c.b[3], c.b[4], c.b[5]);
      if (split >= s0 && split + 1 < e0) {
        const int k = split - s0;
        S.dep[2 * (2 * k + side)] = d0;
        S.dep[2 * (2 * k + side) + 1] = d1;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        const uint32_t old = atomicAdd(&S.arr[k], 1u + ((uint32_t)side << 8));
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        if ((old & 0xffu) == 0u) break;  // the sibling arrives later (here or, after publishing, above)
        merge_deposit(c, side == 0, &S.dep[2 * (2 * k + (side ^ 1))]);
        if (c.l == 0 && c.r == n - 1) {
          write_node(c, 0, n, parent, left, right, node_box, tnode);
          break;
        }
      } else {
        // the split lies on or beyond the chunk's edge: hand off to the top-level passes
        float4* d = deposits + 2 * (2 * (size_t)split + side);
        d[0] = d0;
        d[1] = d1;
        const uint32_t slot = atomicAdd(&misc[kMiscDeferred], 1u);
        if (slot < (uint32_t)n) deferred[slot] = 2u * (uint32_t)split + (uint32_t)side;
        break;
      }
    }
  }
  __syncthreads();
  // splits still waiting for a sibling from outside the chunk: publish the
  // arrival count and the deposit for the top-level passes
  for (int k = threadIdx.x; k + 1 < e0 - s0; k += kClimbThreads) {
    const uint32_t a = S.arr[k];
    if ((a & 0xffu) != 1u) continue;
    const int side = (int)(a >> 8);
    const int split = s0 + k;
    arrivals[split] = 1u;
    float4* d = deposits + 2 * (2 * (size_t)split + side);
    d[0] = S.dep[2 * (2 * k + side)];
    d[1] = S.dep[2 * (2 * k + side) + 1];
  }
}
template <typename K>
__global__ __launch_bounds__(kClimbThreads) void k_climb_chunks(
    const K* __restrict__ keys, const uint32_t* __restrict__ ids, int n, const float4* __restrict__ trirec,
    const float* __restrict__ vx, const float* __restrict__ vy, const float* __restrict__ vz,
    const int32_t* __restrict__ idx, int32_t* __restrict__ parent, int32_t* __restrict__ left,
    int32_t* __restrict__ right, float* __restrict__ node_box, float4* __restrict__ tnode, float4* __restrict__ ltri,
    float4* __restrict__ deposits, uint32_t* __restrict__ arrivals, uint32_t* __restrict__ deferred,
    uint32_t* __restrict__ misc, int32_t* __restrict__ leaf_of) {
  climb_chunks_body(blockIdx.x, gridDim.x, keys, ids, n, trirec, vx, vy, vz, idx, parent, left, right, node_box,
                    tnode, ltri, deposits, arrivals, deferred, misc, leaf_of);
}

// The climber of hand-off `code` (split * 2 + side) from its deposit.
__device__ __forceinline__ ClimbNode climber_of(uint32_t code, const float4* __restrict__ deposits) {
  const int split = (int)(code >> 1), side = (int)(code & 1u);
  ClimbNode c;
  const float4* d = deposits + 2 * (size_t)code;
  float4 q0 = d[0], q1 = d[1];
  c.idx = __float_as_int(q0.x);
  int far = __float_as_int(q0.y);
  c.l = side == 0 ? far : split + 1;
  c.r = side == 0 ? split : far;
  c.b[0] = q0.z; c.b[1] = q0.w; c.b[2] = q1.x; c.b[3] = q1.y; c.b[4] = q1.z; c.b[5] = q1.w;
  c.lidx = c.ridx = 0;
  return c;
}

// One level of the top of the tree per launch, over all workgroups: each
// hand-off counts its arrival at its split; the second arrival merges the
// two deposits and hands the parent on to the next launch. Every deposit
// such a launch reads was written by an earlier launch, and the arrival
// counters are device atomics, so no cache maintenance is needed. A
// balanced top halves the hand-offs per launch; what a bounded number of
// launches leaves (a deep, unbalanced top) k_climb_top finishes.
template <typename K>
__device__ __forceinline__ void climb_pass_body(uint32_t bid, uint32_t nblk, 
    const K* __restrict__ keys, int n, int32_t* __restrict__ parent, int32_t* __restrict__ left,
    int32_t* __restrict__ right, float* __restrict__ node_box, float4* __restrict__ tnode,
    float4* __restrict__ deposits, uint32_t* __restrict__ arrivals, const uint32_t* __restrict__ in,
    uint32_t* __restrict__ out, uint32_t* __restrict__ misc, int in_slot, int out_slot, int clear_slot) {
  if (bid == 0 && threadIdx.x == 0) misc[clear_slot] = 0u;
  const uint32_t count = min(misc[in_slot], (uint32_t)n);
  for (uint32_t e = bid * 256 + threadIdx.x; e < count; e += nblk * 256) {
    const uint32_t code = in[e];
    const int split = (int)(code >> 1), side = (int)(code & 1u);
    if (split >= n - 1) {
      atomicAdd(&misc[kMiscBuildError], 1u);
      continue;
    }
    if (atomicAdd(&arrivals[split], 1u) == 0u) continue;
    ClimbNode c = climber_of(code, deposits);
    merge_deposit(c, side == 0, deposits + 2 * (size_t)(code ^ 1u));
    if (c.l == 0 && c.r == n - 1) {
      write_node(c, 0, n, parent, left, right, node_box, tnode);
      continue;
    }
    int side2;
    const int split2 = climb_step<K>(c, keys, n, &side2, parent, left, right, node_box, tnode, deposits);
    if (split2 < 0) {
      atomicAdd(&misc[kMiscBuildError], 1u);
      continue;
    }
    const uint32_t slot = atomicAdd(&misc[out_slot], 1u);
    if (slot < (uint32_t)n)
      out[slot] = 2u * (uint32_t)split2 + (uint32_t)side2;
    else
      atomicAdd(&misc[kMiscBuildError], 1u);  // cannot happen: hand-offs have disjoint leaf ranges
  }
}
template <typename K>
__global__ __launch_bounds__(256) void k_climb_pass(
    const K* __restrict__ keys, int n, int32_t* __restrict__ parent, int32_t* __restrict__ left,
    int32_t* __restrict__ right, float* __restrict__ node_box, float4* __restrict__ tnode,
    float4* __restrict__ deposits, uint32_t* __restrict__ arrivals, const uint32_t* __restrict__ in,
    uint32_t* __restrict__ out, uint32_t* __restrict__ misc, int in_slot, int out_slot, int clear_slot) {
  climb_pass_body(blockIdx.x, gridDim.x, keys, n, parent, left, right, node_box, tnode, deposits, arrivals, in, out,
                  misc, in_slot, out_slot, clear_slot);
}

template <typename K>
__device__ __forceinline__ void climb_top_body(uint32_t bid, uint32_t nblk, 
    const K* __restrict__ keys, int n, int32_t* __restrict__ parent, int32_t* __restrict__ left,
    int32_t* __restrict__ right, float* __restrict__ node_box, float4* __restrict__ tnode,
    float4* __restrict__ deposits, uint32_t* __restrict__ arrivals, const uint32_t* __restrict__ deferred,
    uint32_t* __restrict__ misc, int in_slot) {
  const uint32_t count = min(misc[in_slot], (uint32_t)n);
  for (uint32_t e = threadIdx.x; e < count; e += kClimbTopThreads) {
    uint32_t code = deferred[e];
    int split = (int)(code >> 1), side = (int)(code & 1u);
    ClimbNode c = climber_of(code, deposits);
    for (int guard = 0;; ++guard) {
      if (split < 0 || guard >= n) {
        atomicAdd((uint32_t*)&misc[kMiscBuildError], 1u);
        break;
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      uint32_t old = atomicAdd(&arrivals[split], 1u);
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      if (old == 0u) break;
      merge_deposit(c, side == 0, deposits + 2 * (2 * (size_t)split + (side ^ 1)));
      if (c.l == 0 && c.r == n - 1) {
        write_node(c, 0, n, parent, left, right, node_box, tnode);
        break;
      }
      split = climb_step<K>(c, keys, n, &side, parent, left, right, node_box, tnode, deposits);
    }
  }
}
template <typename K>
__global__ __launch_bounds__(kClimbTopThreads) void k_climb_top(
    const K* __restrict__ keys, int n, int32_t* __restrict__ parent, int32_t* __restrict__ left,
    int32_t* __restrict__ right, float* __restrict__ node_box, float4* __restrict__ tnode,
    float4* __restrict__ deposits, uint32_t* __restrict__ arrivals, const uint32_t* __restrict__ deferred,
    uint32_t* __restrict__ misc, int in_slot) {
  climb_top_body(blockIdx.x, gridDim.x, keys, n, parent, left, right, node_box, tnode, deposits, arrivals, deferred,
                 misc, in_slot);
}

// ------------------------------------------------------------ export
template <typename K>
__global__ __launch_bounds__(256) void k_export(int n, const K* __restrict__ keys, const uint32_t* __restrict__ ids,
                                                const int32_t* __restrict__ parent, const int32_t* __restrict__ left,
                                                const int32_t* __restrict__ right, const float* __restrict__ node_box,
                                                const float4* __restrict__ ltri, K* codes_out, uint32_t* ids_out,
                                                int32_t* parent_out, int32_t* left_out, int32_t* right_out,
                                                float* lo_out, float* hi_out, float* tlo_out, float* thi_out) {
  int i = blockIdx.x * 256 + threadIdx.x;
  int total = 2 * n - 1;
  if (i >= total) return;
  if (i < n) {
    if (codes_out) codes_out[i] = keys[i];
    if (ids_out) ids_out[i] = ids[i];
    if (tlo_out || thi_out) {
      // the triangle of sorted leaf i: its box from the leaf record's
      // vertices, formed as k_aabb_bounds forms it (kernels/scene.cu:110-120)
      const float4* R = ltri + kLeafStride * (size_t)i;
      const float4 q0 = R[0], q1 = R[1], q2 = R[2];
      const size_t t = (size_t)__float_as_int(q2.y);
      if (tlo_out) {
        tlo_out[3 * t + 0] = fminf(q1.z, fminf(q0.x, q0.w));
        tlo_out[3 * t + 1] = fminf(q1.w, fminf(q0.y, q1.x));
        tlo_out[3 * t + 2] = fminf(q2.x, fminf(q0.z, q1.y));
      }
      if (thi_out) {
        thi_out[3 * t + 0] = fmaxf(q1.z, fmaxf(q0.x, q0.w));
        thi_out[3 * t + 1] = fmaxf(q1.w, fmaxf(q0.y, q1.x));
        thi_out[3 * t + 2] = fmaxf(q2.x, fmaxf(q0.z, q1.y));
      }
    }
  }
  if (parent_out) parent_out[i] = parent[i];
  int l, r;
  if (i < n - 1) {
    l = left[i];
    r = right[i];
  } else {
    l = r = (int)ids[i - (n - 1)];
  }
  if (left_out) left_out[i] = l;
  if (right_out) right_out[i] = r;
  if (lo_out)
    for (int k = 0; k < 3; ++k) lo_out[3 * i + k] = node_box[6 * (size_t)i + k];
  if (hi_out)
    for (int k = 0; k < 3; ++k) hi_out[3 * i + k] = node_box[6 * (size_t)i + 3 + k];
}

// ---- 4-wide traversal nodes
// Depth parity of each internal node (walk to the root through the parent
// links): the even-depth internal nodes become 4-wide nodes.
__device__ __forceinline__ void wide_parity_body(uint32_t bid, uint32_t nblk, int n, const int32_t* __restrict__ parent,
                                                     uint8_t* __restrict__ flag) {
  const int i = bid * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  int x = i, d = 0;
  for (int guard = 0; x != 0 && guard < n; ++guard) {
    x = parent[x];
    ++d;
  }
  flag[i] = ((d & 1) == 0 ? 1 : 0) | (d % 3 == 0 ? 2 : 0);
}
__global__ __launch_bounds__(256) void k_wide_parity(int n, const int32_t* __restrict__ parent,
                                                     uint8_t* __restrict__ flag) {
  wide_parity_body(blockIdx.x, gridDim.x, n, parent, flag);
}

// ---- packing the 4-wide nodes: node i (even depth, flag bit 0) goes to
// slot wcomp[i] = number of even-depth nodes before it, so the nodes the
// traversal reads fill whole cache lines. A three-step scan over blocks of
// 1024 flags: block counts, their exclusive scan, then in-block ranks.
__device__ __forceinline__ uint32_t block_excl_scan_1024(uint32_t v, uint32_t* s_w, uint32_t* total) {
  const uint32_t lane = __lane_id(), w = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t y = __shfl_up(x, off);
    if (lane >= (uint32_t)off) x += y;
  }
  if (lane == 63) s_w[w] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t acc = 0;
    for (int k = 0; k < 16; ++k) {
      uint32_t t = s_w[k];
      s_w[k] = acc;
      acc += t;
    }
    s_w[16] = acc;
  }
  __syncthreads();
  if (total) *total = s_w[16];
  return s_w[w] + x - v;
}

__device__ __forceinline__ void pack_count_body(uint32_t bid, uint32_t nblk, int m, const uint8_t* __restrict__ flag, uint32_t* sums) {
  __shared__ uint32_t s_w[17];
  const int i = bid * 1024 + threadIdx.x;
  const uint32_t v = i < m ? (flag[i] & 1u) : 0u;
  uint32_t total;
  block_excl_scan_1024(v, s_w, &total);
  if (threadIdx.x == 0) sums[bid] = total;
}
__global__ __launch_bounds__(1024) void k_pack_count(int m, const uint8_t* __restrict__ flag, uint32_t* sums) {
  pack_count_body(blockIdx.x, gridDim.x, m, flag, sums);
}

// Exclusive scan of the per-block counts in place; sums[blocks] receives the
// total (the number of 4-wide nodes).
__device__ __forceinline__ void pack_offsets_body(uint32_t bid, uint32_t nblk, int blocks, uint32_t* sums) {
  __shared__ uint32_t s_w[17];
  __shared__ uint32_t s_carry;
  if (threadIdx.x == 0) s_carry = 0;
  __syncthreads();
  for (int base = 0; base < blocks; base += 1024) {
    const int k = base + threadIdx.x;
    const uint32_t v = k < blocks ? sums[k] : 0u;
    uint32_t total;
    const uint32_t ex = block_excl_scan_1024(v, s_w, &total);
    const uint32_t carry = s_carry;
    if (k < blocks) sums[k] = carry + ex;
    __syncthreads();
    if (threadIdx.x == 0) s_carry = carry + total;
    __syncthreads();
  }
  if (threadIdx.x == 0) sums[blocks] = s_carry;
}
__global__ __launch_bounds__(1024) void k_pack_offsets(int blocks, uint32_t* sums) {
  pack_offsets_body(blockIdx.x, gridDim.x, blocks, sums);
}

__device__ __forceinline__ void pack_index_body(uint32_t bid, uint32_t nblk, int m, const uint8_t* __restrict__ flag,
                                                     const uint32_t* __restrict__ offsets, int32_t* __restrict__ wcomp) {
  __shared__ uint32_t s_w[17];
  const int i = bid * 1024 + threadIdx.x;
  const uint32_t v = i < m ? (flag[i] & 1u) : 0u;
  const uint32_t r = block_excl_scan_1024(v, s_w, nullptr);
  if (i < m) wcomp[i] = v ? (int32_t)(offsets[bid] + r) : -1;
}
__global__ __launch_bounds__(1024) void k_pack_index(int m, const uint8_t* __restrict__ flag,
                                                     const uint32_t* __restrict__ offsets, int32_t* __restrict__ wcomp) {
  pack_index_body(blockIdx.x, gridDim.x, m, flag, offsets, wcomp);
}

// Fill the 4-wide node of every even-depth internal node from the binary
// topology and the refit boxes (node numbering of the reference: internal
// [0, n-2], leaf of sorted position p at n-1+p). wsrc (if not null)
// receives, per packed node, the binary nodes of its child slots (-1:
// empty), which a refit reads instead of walking the topology again.
__device__ __forceinline__ void wide_fill_body(uint32_t bid, uint32_t nblk, int n, const int32_t* __restrict__ left,
                                                   const int32_t* __restrict__ right,
                                                   const float* __restrict__ node_box,
                                                   const uint8_t* __restrict__ flag, float4* __restrict__ wide,
                                                   float4* __restrict__ wideq, const int32_t* __restrict__ wcomp,
                                                   int4* __restrict__ wsrc) {
  const int i = bid * blockDim.x + threadIdx.x;
  if (i >= n - 1 || !(flag[i] & 1)) return;
  int slot[4];
  int k = 0;
  const int c[2] = {left[i], right[i]};
  for (int j = 0; j < 2; ++j) {
    if (c[j] >= n - 1) {
      slot[k++] = c[j];
    } else {
      slot[k++] = left[c[j]];
      slot[k++] = right[c[j]];
    }
  }
  float lo[3][4], hi[3][4];
  int ref[4];
  for (int j = 0; j < 4; ++j) {
    if (j < k) {
      const float* b = node_box + 6 * (size_t)slot[j];
      for (int a = 0; a < 3; ++a) {
        lo[a][j] = b[a];
        hi[a][j] = b[3 + a];
      }
      ref[j] = slot[j] >= n - 1 ? ~(slot[j] - (n - 1)) : slot[j];
    } else {
      for (int a = 0; a < 3; ++a) {
        lo[a][j] = __int_as_float(0x7f800000);   // +inf
        hi[a][j] = __int_as_float(0xff800000);   // -inf
      }
      ref[j] = kNullLeaf(n);
    }
  }
  const float4 refs =
      make_float4(__int_as_float(ref[0]), __int_as_float(ref[1]), __int_as_float(ref[2]), __int_as_float(ref[3]));
  if (wide) {
    float4* w = wide + 8 * (size_t)i;
    for (int a = 0; a < 3; ++a) {
      w[a] = make_float4(lo[a][0], lo[a][1], lo[a][2], lo[a][3]);
      w[3 + a] = make_float4(hi[a][0], hi[a][1], hi[a][2], hi[a][3]);
    }
    w[6] = refs;
    w[7] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
  if (!wideq) return;
  const float* nbp = node_box + 6 * (size_t)i;
  const float nb[6] = {nbp[0], nbp[1], nbp[2], nbp[3], nbp[4], nbp[5]};
  // packed: this node's slot, and its internal children's slots as refs
  const int w = wcomp[i];
  float4* q = wideq + 4 * (size_t)w;
  quantize4(nb, lo, hi, k, q);
  int cref[4];
  for (int j = 0; j < 4; ++j) cref[j] = ref[j] >= 0 ? wcomp[ref[j]] : ref[j];
  q[3] = make_float4(__int_as_float(cref[0]), __int_as_float(cref[1]), __int_as_float(cref[2]),
                     __int_as_float(cref[3]));
  if (wsrc) wsrc[w] = make_int4(k > 0 ? slot[0] : -1, k > 1 ? slot[1] : -1, k > 2 ? slot[2] : -1, k > 3 ? slot[3] : -1);
}

// A refit's 4-wide nodes: the topology is unchanged, so each packed node
// keeps its child refs (q[3]) and takes its planes from the child slots'
// new boxes, found through wsrc (written by the build's wide fill). The
// node's own box is the union of its child slots' boxes (fminf/fmaxf, exact
// and order-independent), so no topology is read.
__global__ __launch_bounds__(256) void k_wide_refit(int n, const uint32_t* __restrict__ wscan,
                                                    const int4* __restrict__ wsrc,
                                                    const float* __restrict__ node_box, float4* __restrict__ wideq) {
  const uint32_t count = wscan[(n - 1 + 1023) / 1024];
  const uint32_t w = blockIdx.x * 256 + threadIdx.x;
  if (w >= count) return;
  const int4 src = wsrc[w];
  const int sl[4] = {src.x, src.y, src.z, src.w};
  float lo[3][4], hi[3][4];
  float nb[6] = {__int_as_float(0x7f800000), __int_as_float(0x7f800000), __int_as_float(0x7f800000),
                 __int_as_float(0xff800000), __int_as_float(0xff800000), __int_as_float(0xff800000)};
  int k = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    if (sl[j] >= 0) {
      const float* b = node_box + 6 * (size_t)sl[j];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        lo[a][j] = b[a];
        hi[a][j] = b[3 + a];
        nb[a] = fminf(nb[a], lo[a][j]);
        nb[3 + a] = fmaxf(nb[3 + a], hi[a][j]);
      }
      ++k;
    } else {
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        lo[a][j] = __int_as_float(0x7f800000);
        hi[a][j] = __int_as_float(0xff800000);
      }
    }
  }
  float4 q[3];
  quantize4(nb, lo, hi, k, q);
  float4* dst = wideq + 4 * (size_t)w;
  dst[0] = q[0];
  dst[1] = q[1];
  dst[2] = q[2];
}

__global__ __launch_bounds__(256) void k_wide_fill(int n, const int32_t* __restrict__ left,
                                                   const int32_t* __restrict__ right,
                                                   const float* __restrict__ node_box,
                                                   const uint8_t* __restrict__ flag, float4* __restrict__ wide,
                                                   float4* __restrict__ wideq, const int32_t* __restrict__ wcomp,
                                                   int4* __restrict__ wsrc) {
  wide_fill_body(blockIdx.x, gridDim.x, n, left, right, node_box, flag, wide, wideq, wcomp, wsrc);
}

// Quantize up to W child boxes against the node box nb: per axis the
// smallest power-of-two step with 255 steps covering the box; planes are
// rounded outward so the decoded boxes contain the exact ones. Slots k..W-1
// are empty (qlo 255 > qhi 0). Byte j of word (axis, j / 4) holds child j.
template <int W>
__device__ void quantize_children(const float* nb, int k, const float (&lo)[3][W], const float (&hi)[3][W],
                                  uint32_t& ebits, uint32_t (&qlo)[3][W / 4], uint32_t (&qhi)[3][W / 4]) {
  ebits = 0;
  for (int a = 0; a < 3; ++a) {
    const double o = nb[a], ext = (double)nb[3 + a] - o;
    int e = 0;
    if (ext > 0) {
      int ex2;
      frexp(ext / 255.0, &ex2);
      ex2 = ex2 < -126 ? -126 : (ex2 > 127 ? 127 : ex2);
      e = ex2 + 127;
    }
    ebits |= (uint32_t)e << (8 * a);
    const double step = e ? ldexp(1.0, e - 127) : 0.0;
    const double inv = e ? ldexp(1.0, 127 - e) : 0.0;  // exact: step is a power of two
    for (int w = 0; w < W / 4; ++w) qlo[a][w] = qhi[a][w] = 0;
    for (int j = 0; j < W; ++j) {
      uint32_t ql = 255, qh = 0;
      if (j < k) {
        if (step == 0) {
          ql = qh = 0;
        } else {
          double l = floor((lo[a][j] - o) * inv), h = ceil((hi[a][j] - o) * inv);
          // within [0, 255] before the adjustment (NaN from non-finite
          // coordinates spans the whole node), so each loop runs at most
          // 255 times whatever the box
          l = l >= 0 ? (l > 255 ? 255 : l) : 0;
          h = h <= 255 ? (h < 0 ? 0 : h) : 255;
          while (l > 0 && o + l * step > lo[a][j]) l -= 1;
          while (h < 255 && o + h * step < hi[a][j]) h += 1;
          ql = (uint32_t)(l < 0 ? 0 : (l > 255 ? 255 : l));
          qh = (uint32_t)(h < 0 ? 0 : (h > 255 ? 255 : h));
        }
      }
      qlo[a][j / 4] |= ql << (8 * (j & 3));
      qhi[a][j / 4] |= qh << (8 * (j & 3));
    }
  }
}

__global__ __launch_bounds__(256) void k_wide8_fill(int n, const int32_t* __restrict__ left,
                                                    const int32_t* __restrict__ right,
                                                    const float* __restrict__ node_box,
                                                    const uint8_t* __restrict__ flag, float4* __restrict__ wide8) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1 || !(flag[i] & 2)) return;
  int slot[8];
  int k = 0;
  const int c[2] = {left[i], right[i]};
  for (int j = 0; j < 2; ++j) {
    if (c[j] >= n - 1) {
      slot[k++] = c[j];
      continue;
    }
    const int g[2] = {left[c[j]], right[c[j]]};
    for (int h = 0; h < 2; ++h) {
      if (g[h] >= n - 1) {
        slot[k++] = g[h];
      } else {
        slot[k++] = left[g[h]];
        slot[k++] = right[g[h]];
      }
    }
  }
  float lo[3][8], hi[3][8];
  int ref[8];
  for (int j = 0; j < 8; ++j) {
    if (j < k) {
      const float* b = node_box + 6 * (size_t)slot[j];
      for (int a = 0; a < 3; ++a) {
        lo[a][j] = b[a];
        hi[a][j] = b[3 + a];
      }
      ref[j] = slot[j] >= n - 1 ? ~(slot[j] - (n - 1)) : slot[j];
    } else {
      for (int a = 0; a < 3; ++a) {
        lo[a][j] = 0.f;
        hi[a][j] = 0.f;
      }
      ref[j] = kNullLeaf(n);
    }
  }
  uint32_t ebits, qlo[3][2], qhi[3][2];
  const float* nb = node_box + 6 * (size_t)i;
  quantize_children<8>(nb, k, lo, hi, ebits, qlo, qhi);
  auto u = [](uint32_t v) { return __uint_as_float(v); };
  float4* w = wide8 + 8 * (size_t)i;
  w[0] = make_float4(nb[0], nb[1], nb[2], u(ebits));
  w[1] = make_float4(u(qlo[0][0]), u(qlo[0][1]), u(qlo[1][0]), u(qlo[1][1]));
  w[2] = make_float4(u(qlo[2][0]), u(qlo[2][1]), u(qhi[0][0]), u(qhi[0][1]));
  w[3] = make_float4(u(qhi[1][0]), u(qhi[1][1]), u(qhi[2][0]), u(qhi[2][1]));
  w[4] = make_float4(__int_as_float(ref[0]), __int_as_float(ref[1]), __int_as_float(ref[2]), __int_as_float(ref[3]));
  w[5] = make_float4(__int_as_float(ref[4]), __int_as_float(ref[5]), __int_as_float(ref[6]), __int_as_float(ref[7]));
  w[6] = make_float4(0.f, 0.f, 0.f, 0.f);
  w[7] = make_float4(0.f, 0.f, 0.f, 0.f);
}


// ---- segmented build: the BVHs of a batch share every launch
// A launch's workgroups are split into consecutive ranges, one per BVH
// (off[s] .. off[s+1]-1; empty for a BVH with no work in that stage), and
// each workgroup runs the stage's body for its BVH with its block index
// inside the range. Every stage runs once over all BVHs, so a batch build
// is about 30 launches that each fill the GPU, instead of 30 small ones per
// BVH. The onesweep tiles of a BVH are still taken in launch order from its
// own tile counter, so a tile waits only on tiles of workgroups that have
// already started.
struct SegBvh {
  int n, ntiles;
  size_t zero_words;
  const float *vx, *vy, *vz;
  const int32_t* idx;
  float* tri_box;
  float4* trirec;
  uint32_t *misc, *hist, *status, *arrivals, *deferred, *deferred2, *wscan;
  void* keys[2];
  uint32_t* ids[2];
  int32_t *parent, *left, *right, *leaf_of, *wcomp;
  float* node_box;
  float4 *tnode, *ltri, *deposits, *wideq;
  int4* wsrc;
  uint8_t* wide_flag;
};

struct SegBlock {
  const SegBvh* g;
  uint32_t bid, nblk;
};

// The BVH of this workgroup: the last s with off[s] <= blockIdx.x.
__device__ __forceinline__ SegBlock seg_block(const SegBvh* __restrict__ segs, const uint32_t* __restrict__ off,
                                              int ns) {
  int lo = 0, hi = ns - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= blockIdx.x)
      lo = mid;
    else
      hi = mid - 1;
  }
  return SegBlock{segs + lo, blockIdx.x - off[lo], off[lo + 1] - off[lo]};
}

constexpr int kMortonPerBlock = 4096;
// Hand-off count slot k % 3 of the top-level passes (pass p reads slot p,
// writes slot p + 1 and clears slot p + 2).
__host__ __device__ constexpr int hand_off_slot(int k) {
  return k % 3 == 0 ? kMiscDeferred : (k % 3 == 1 ? kMiscDeferredB : kMiscDeferredC);
}

__global__ __launch_bounds__(256) void k_seg_init(const SegBvh* segs, const uint32_t* off, int ns) {
  const SegBlock b = seg_block(segs, off, ns);
  init_body(b.bid, b.nblk, b.g->misc, b.g->hist, b.g->zero_words, 1);
}

__global__ __launch_bounds__(256) void k_seg_aabb_bounds(const SegBvh* segs, const uint32_t* off, int ns) {
  const SegBlock b = seg_block(segs, off, ns);
  const SegBvh& g = *b.g;
  aabb_bounds_body(b.bid, b.nblk, g.vx, g.vy, g.vz, g.idx, g.n, g.tri_box, g.trirec, g.misc, 1);
}

template <typename K>
__global__ __launch_bounds__(256) void k_seg_morton(const SegBvh* segs, const uint32_t* off, int ns) {
  const SegBlock b = seg_block(segs, off, ns);
  const SegBvh& g = *b.g;
  morton_body<K>(b.bid, b.nblk, g.tri_box, g.n, g.misc, (K*)g.keys[0], g.hist, kMortonPerBlock, GivenBounds{});
}

template <typename K>
__global__ __launch_bounds__(kSortThreads) void k_seg_onesweep(const SegBvh* segs, const uint32_t* off, int ns,
                                                               int p) {
  const SegBlock b = seg_block(segs, off, ns);
  const SegBvh& g = *b.g;
  const int src = p & 1, dst = src ^ 1;
  onesweep_body<K>(b.bid, b.nblk, (const K*)g.keys[src], p == 0 ? nullptr : g.ids[src], (K*)g.keys[dst],
                   g.ids[dst], g.n, 8 * p, g.hist + 256 * p, g.status + (size_t)p * g.ntiles * 256,
                   g.misc + kMiscTileCtr + p);
}

template <typename K>
__global__ __launch_bounds__(kClimbThreads) void k_seg_climb_chunks(const SegBvh* segs, const uint32_t* off,
                                                                    int ns) {
  const SegBlock b = seg_block(segs, off, ns);
  const SegBvh& g = *b.g;
  climb_chunks_body<K>(b.bid, b.nblk, (const K*)g.keys[0], g.ids[0], g.n, g.trirec, g.vx, g.vy, g.vz, g.idx,
                       g.parent, g.left, g.right, g.node_box, g.tnode, g.ltri, g.deposits, g.arrivals, g.deferred,
                       g.misc, g.leaf_of);
}

template <typename K>
__global__ __launch_bounds__(256) void k_seg_climb_pass(const SegBvh* segs, const uint32_t* off, int ns, int p) {
  const SegBlock b = seg_block(segs, off, ns);
  const SegBvh& g = *b.g;
  uint32_t* lists[2] = {g.deferred, g.deferred2};
  climb_pass_body<K>(b.bid, b.nblk, (const K*)g.keys[0], g.n, g.parent, g.left, g.right, g.node_box, g.tnode,
                     g.deposits, g.arrivals, lists[p & 1], lists[(p + 1) & 1], g.misc, hand_off_slot(p),
                     hand_off_slot(p + 1), hand_off_slot(p + 2));
}

template <typename K>
__global__ __launch_bounds__(kClimbTopThreads) void k_seg_climb_top(const SegBvh* segs, const uint32_t* off, int ns,
                                                                    int passes) {
  const SegBlock b = seg_block(segs, off, ns);
  const SegBvh& g = *b.g;
  climb_top_body<K>(b.bid, b.nblk, (const K*)g.keys[0], g.n, g.parent, g.left, g.right, g.node_box, g.tnode,
                    g.deposits, g.arrivals, (passes & 1) ? g.deferred2 : g.deferred, g.misc, hand_off_slot(passes));
}

__global__ __launch_bounds__(256) void k_seg_wide_parity(const SegBvh* segs, const uint32_t* off, int ns) {
  const SegBlock b = seg_block(segs, off, ns);
  wide_parity_body(b.bid, b.nblk, b.g->n, b.g->parent, b.g->wide_flag);
}

__global__ __launch_bounds__(1024) void k_seg_pack_count(const SegBvh* segs, const uint32_t* off, int ns) {
  const SegBlock b = seg_block(segs, off, ns);
  pack_count_body(b.bid, b.nblk, b.g->n - 1, b.g->wide_flag, b.g->wscan);
}

__global__ __launch_bounds__(1024) void k_seg_pack_offsets(const SegBvh* segs, const uint32_t* off, int ns) {
  const SegBlock b = seg_block(segs, off, ns);
  pack_offsets_body(b.bid, b.nblk, (b.g->n - 1 + 1023) / 1024, b.g->wscan);
}

__global__ __launch_bounds__(1024) void k_seg_pack_index(const SegBvh* segs, const uint32_t* off, int ns) {
  const SegBlock b = seg_block(segs, off, ns);
  pack_index_body(b.bid, b.nblk, b.g->n - 1, b.g->wide_flag, b.g->wscan, b.g->wcomp);
}

__global__ __launch_bounds__(256) void k_seg_wide_fill(const SegBvh* segs, const uint32_t* off, int ns) {
  const SegBlock b = seg_block(segs, off, ns);
  const SegBvh& g = *b.g;
  wide_fill_body(b.bid, b.nblk, g.n, g.left, g.right, g.node_box, g.wide_flag, nullptr, g.wideq, g.wcomp, g.wsrc);
}

// Binary traversal records of every internal node from the topology and
// the refit boxes (the layout-2 traversal and the megakernel read them; the
// climb writes them itself while a BVH wants them).
__global__ __launch_bounds__(256) void k_tnode_fill(int n, const int32_t* __restrict__ left,
                                                    const int32_t* __restrict__ right,
                                                    const float* __restrict__ node_box, float4* __restrict__ tnode) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n - 1) return;
  const int l = left[i], r = right[i];
  const float* lb = node_box + 6 * (size_t)l;
  const float* rb = node_box + 6 * (size_t)r;
  float4* T = tnode + 4 * (size_t)i;
  T[0] = make_float4(lb[0], lb[1], lb[2], lb[3]);
  T[1] = make_float4(lb[4], lb[5], rb[0], rb[1]);
  T[2] = make_float4(rb[2], rb[3], rb[4], rb[5]);
  T[3] = make_float4(__int_as_float(child_ref(l, n)), __int_as_float(child_ref(r, n)), 0.f, 0.f);
}

// The climb writes binary records only while a traversal wants them.
inline float4* tnode_out(const rt_bvh* b) { return b->want_tnode ? b->tnode : nullptr; }

inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

// The wide nodes after a build (topology_changed) or a refit. A refit keeps
// the packed 4-wide nodes' refs and requantizes their planes (k_wide_refit)
// unless the unquantized 4-wide nodes are wanted too.
void launch_wide(rt_bvh* b, bool topology_changed, hipStream_t s) {
  const int n = b->n;
  if (n < 2) return;
  if (topology_changed) {
    const int m = n - 1, blocks = cdiv(m, 1024);
    k_wide_parity<<<cdiv(m, 256), 256, 0, s>>>(n, b->parent, b->wide_flag);
    k_pack_count<<<blocks, 1024, 0, s>>>(m, b->wide_flag, b->wscan);
    k_pack_offsets<<<1, 1024, 0, s>>>(blocks, b->wscan);
    k_pack_index<<<blocks, 1024, 0, s>>>(m, b->wide_flag, b->wscan, b->wcomp);
  }
  if (!topology_changed && !b->want_wide4)
    k_wide_refit<<<cdiv(n - 1, 256), 256, 0, s>>>(n, b->wscan, b->wsrc, b->node_box, b->wideq);
  else
    k_wide_fill<<<cdiv(n - 1, 256), 256, 0, s>>>(n, b->left, b->right, b->node_box, b->wide_flag,
                                                 b->want_wide4 ? b->wide : nullptr, b->wideq, b->wcomp, b->wsrc);
  if (b->want_wide8)
    k_wide8_fill<<<cdiv(n - 1, 256), 256, 0, s>>>(n, b->left, b->right, b->node_box, b->wide_flag, b->wide8);
}

// Number of top-level passes for n leaves (0 when one chunk holds them all)
// and their grid. A pass over an empty hand-off list changes nothing, so a
// segmented build runs the largest count for every BVH.
int top_passes(int n) {
  const int chunks = cdiv(n, kClimbChunk);
  int passes = 0;
  if (chunks > 1) {
    passes = 2;
    while ((1 << (passes - 2)) < chunks && passes < 26) ++passes;
  }
  return passes;
}
int top_pass_grid(int n) { return std::max(1, std::min(cdiv(3L * cdiv(n, kClimbChunk), 256), 1024)); }

// The top of the tree above the chunks: about log2(chunks) + 2 one-level
// launches over the whole GPU, then one workgroup for whatever remains.
template <typename K>
void launch_climb_top(rt_bvh* b, const K* keys, hipStream_t s, bool topology = true) {
  const int n = b->n;
  if (n <= 1) return;
  int32_t *par = topology ? b->parent : nullptr, *lft = topology ? b->left : nullptr,
          *rgt = topology ? b->right : nullptr;
  const int passes = top_passes(n);
  uint32_t* lists[2] = {b->deferred, b->deferred2};
  const int grid = top_pass_grid(n);
  for (int p = 0; p < passes; ++p)
    k_climb_pass<K><<<grid, 256, 0, s>>>(keys, n, par, lft, rgt, b->node_box, tnode_out(b), b->deposits,
                                         b->arrivals, lists[p & 1], lists[(p + 1) & 1], b->misc, hand_off_slot(p),
                                         hand_off_slot(p + 1), hand_off_slot(p + 2));
  k_climb_top<K><<<1, kClimbTopThreads, 0, s>>>(keys, n, par, lft, rgt, b->node_box, tnode_out(b),
                                                b->deposits, b->arrivals, lists[passes & 1], b->misc,
                                                hand_off_slot(passes));
}

// hist, look-back status and arrival counters: one contiguous range
size_t counter_words(const rt_bvh* b) {
  return (size_t)b->passes * 256 + (size_t)b->passes * b->ntiles * 256 + (size_t)(b->n > 1 ? b->n - 1 : 1);
}

template <typename K>
void launch_morton(rt_bvh* b, const GivenBounds& gb, hipStream_t s) {
  const int per_block = kMortonPerBlock;
  k_morton<K><<<cdiv(b->n, per_block), 256, 0, s>>>(b->tri_box, b->n, b->misc, (K*)b->keys[0], b->hist, per_block,
                                                      gb);
}

template <typename K>
void launch_sort(rt_bvh* b, hipStream_t s) {
  constexpr int P = KeyTraits<K>::kPasses;
  for (int p = 0; p < P; ++p) {
    int src = p & 1, dst = src ^ 1;
    k_onesweep<K><<<b->ntiles, kSortThreads, 0, s>>>(
        (const K*)b->keys[src], p == 0 ? nullptr : b->ids[src], (K*)b->keys[dst], b->ids[dst], b->n, 8 * p,
        b->hist + 256 * p, b->status + (size_t)p * b->ntiles * 256, b->misc + kMiscTileCtr + p);
  }
}

// Hierarchy, node boxes and the traversal layouts from the sorted keys (P is
// even: the sorted keys and ids are back in buffer 0).
template <typename K>
void launch_tree(rt_bvh* b, const rt_mesh_view* m, hipStream_t s) {
  const int n = b->n;
  const K* keys = (const K*)b->keys[0];
  if (b->build_flags & RT_BUILD_KARRAS_SEARCH) {
    k_hierarchy<K><<<cdiv(n, 256), 256, 0, s>>>(keys, b->ids[0], n, b->tri_box, m->vx_dev, m->vy_dev, m->vz_dev,
                                                m->idx_dev, b->parent, b->left, b->right, b->flags, b->node_box,
                                                b->ltri, b->leaf_of);
    if (n > 1)
      k_refit<<<cdiv(n, 256), 256, 0, s>>>(n, b->parent, b->left, b->right, b->flags, b->node_box, b->tnode);
  } else {
    k_climb_chunks<K><<<cdiv(n, kClimbChunk), kClimbThreads, 0, s>>>(
        keys, b->ids[0], n, b->trirec, m->vx_dev, m->vy_dev, m->vz_dev, m->idx_dev, b->parent, b->left, b->right,
        b->node_box, tnode_out(b), b->ltri, b->deposits, b->arrivals, b->deferred, b->misc, b->leaf_of);
    launch_climb_top<K>(b, keys, s);
  }
  launch_wide(b, true, s);
}

template <typename K>
hipError_t build_typed(rt_bvh* b, const rt_mesh_view* m, hipStream_t s) {
  const int n = b->n;
  k_init<<<min(cdiv((long)counter_words(b), 256), 2048), 256, 0, s>>>(b->misc, b->hist, counter_words(b), 1);
  k_aabb_bounds<<<min(cdiv(n, 256), 1024), 256, 0, s>>>(m->vx_dev, m->vy_dev, m->vz_dev, m->idx_dev, n, b->tri_box,
                                                        b->trirec, b->misc, 1);
  launch_morton<K>(b, GivenBounds{}, s);
  launch_sort<K>(b, s);
  launch_tree<K>(b, m, s);
  return hipGetLastError();
}

// Triangle boxes in triangle order as two n*3 arrays.
__global__ void k_split_boxes(int n, const float* __restrict__ tri_box, float* __restrict__ lo, float* __restrict__ hi) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float* b = tri_box + 6 * (size_t)i;
  for (int k = 0; k < 3; ++k) {
    if (lo) lo[3 * (size_t)i + k] = b[k];
    if (hi) hi[3 * (size_t)i + k] = b[3 + k];
  }
}

template <typename K>
hipError_t refit_typed(rt_bvh* b, const rt_mesh_view* m, hipStream_t s) {
  const int n = b->n;
  if (b->build_flags & RT_BUILD_KARRAS_SEARCH) {
    k_refit_leaves<<<cdiv(n, 256), 256, 0, s>>>(m->vx_dev, m->vy_dev, m->vz_dev, m->idx_dev, b->ids[0], n,
                                                  b->tri_box, b->node_box, b->ltri, b->flags);
    if (n > 1)
      k_refit<<<cdiv(n, 256), 256, 0, s>>>(n, b->parent, b->left, b->right, b->flags, b->node_box, b->tnode);
    launch_wide(b, false, s);
    return hipGetLastError();
  }
  // The keys are unchanged, so the climb re-derives the same topology (and
  // leaves the parent and child links in place) and recomputes every box
  // from the moved vertices.
  size_t zero_words = (size_t)(n > 1 ? n - 1 : 1);
  k_init<<<min(cdiv((long)zero_words, 256), 2048), 256, 0, s>>>(b->misc, b->arrivals, zero_words, 0);
  // leaf records and boxes come straight from the moved vertices (no
  // triangle-order pass); the scene bounds of the build are kept
  const K* keys = (const K*)b->keys[0];
  k_climb_chunks<K><<<cdiv(n, kClimbChunk), kClimbThreads, 0, s>>>(
      keys, b->ids[0], n, nullptr, m->vx_dev, m->vy_dev, m->vz_dev, m->idx_dev, nullptr, nullptr, nullptr,
      b->node_box, tnode_out(b), b->ltri, b->deposits, b->arrivals, b->deferred, b->misc, nullptr);
  launch_climb_top<K>(b, keys, s, false);
  launch_wide(b, false, s);
  return hipGetLastError();
}

}  // namespace

hipError_t rt_launch_stage_bounds(rt_bvh* b, const rt_mesh_view* m, float* lo, float* hi, hipStream_t s) {
  const int n = b->n;
  k_init<<<1, 256, 0, s>>>(b->misc, nullptr, 0, 1);
  k_aabb_bounds<<<min(cdiv(n, 256), 1024), 256, 0, s>>>(m->vx_dev, m->vy_dev, m->vz_dev, m->idx_dev, n, b->tri_box,
                                                        b->trirec, b->misc, 1);
  if (lo || hi) k_split_boxes<<<cdiv(n, 256), 256, 0, s>>>(n, b->tri_box, lo, hi);
  return hipGetLastError();
}

hipError_t rt_launch_stage_morton(rt_bvh* b, const double* mn, const double* mx, hipStream_t s) {
  GivenBounds gb{};
  if (mn && mx) {
    for (int k = 0; k < 3; ++k) {
      gb.mn[k] = mn[k];
      gb.mx[k] = mx[k];
    }
    gb.use = 1;
  }
  k_init<<<min(cdiv((long)counter_words(b), 256), 2048), 256, 0, s>>>(b->misc, b->hist, counter_words(b), 0);
  if (b->bits == 30)
    launch_morton<uint32_t>(b, gb, s);
  else
    launch_morton<uint64_t>(b, gb, s);
  return hipGetLastError();
}

hipError_t rt_launch_stage_sort(rt_bvh* b, hipStream_t s) {
  if (b->bits == 30)
    launch_sort<uint32_t>(b, s);
  else
    launch_sort<uint64_t>(b, s);
  return hipGetLastError();
}

hipError_t rt_launch_stage_tree(rt_bvh* b, const rt_mesh_view* m, hipStream_t s) {
  if (b->bits == 30)
    launch_tree<uint32_t>(b, m, s);
  else
    launch_tree<uint64_t>(b, m, s);
  return hipGetLastError();
}

hipError_t rt_note_capture(const rt_bvh* b, hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  hipError_t e = hipStreamIsCapturing(s, &st);
  if (e != hipSuccess || st != hipStreamCaptureStatusActive) return e;
  b->cap_tnode = (b->captured ? b->cap_tnode : true) && b->want_tnode;
  b->cap_wide4 = (b->captured ? b->cap_wide4 : true) && b->want_wide4;
  b->cap_wide8 = (b->captured ? b->cap_wide8 : true) && b->want_wide8;
  b->captured = true;
  return hipSuccess;
}

hipError_t rt_ensure_layout(const rt_bvh* b, int layout, hipStream_t s) {
  const int n = b->n;
  if (layout == 2 && (!b->want_tnode || (b->captured && !b->cap_tnode))) {
    b->want_tnode = true;
    if (n >= 2) k_tnode_fill<<<cdiv(n - 1, 256), 256, 0, s>>>(n, b->left, b->right, b->node_box, b->tnode);
  } else if (layout == 4 && (!b->want_wide4 || (b->captured && !b->cap_wide4))) {
    b->want_wide4 = true;
    if (n >= 2)
      k_wide_fill<<<cdiv(n - 1, 256), 256, 0, s>>>(n, b->left, b->right, b->node_box, b->wide_flag, b->wide, nullptr,
                                                   b->wcomp, nullptr);
  } else if (layout == 8 && (!b->want_wide8 || (b->captured && !b->cap_wide8))) {
    b->want_wide8 = true;
    if (n >= 2)
      k_wide8_fill<<<cdiv(n - 1, 256), 256, 0, s>>>(n, b->left, b->right, b->node_box, b->wide_flag, b->wide8);
  }
  return hipGetLastError();
}

void rt_keep_only_layout(const rt_bvh* b, int layout) {
  b->want_tnode = b->want_tnode && layout == 2;
  b->want_wide4 = b->want_wide4 && layout == 4;
  b->want_wide8 = b->want_wide8 && layout == 8;
}

// ---- segmented batch build (host side)
namespace {
enum SegStage {
  kStInit, kStAabb, kStMorton, kStSort, kStChunks, kStPass, kStTop, kStParity, kStPackCount, kStPackOffsets,
  kStPackIndex, kStWideFill, kSegStages
};
}  // namespace

struct rt_seg_build {
  int ns = 0, bits = 0, passes = 0;
  void* dev = nullptr;  // SegBvh[ns], then kSegStages tables of ns + 1 block offsets
  size_t dev_bytes = 0;
  const SegBvh* segs = nullptr;
  const uint32_t* off[kSegStages] = {};
  uint32_t total[kSegStages] = {};
};

void rt_seg_build_destroy(rt_ctx* ctx, rt_seg_build* p) {
  if (!p) return;
  rt_dev_free(ctx, p->dev, p->dev_bytes);
  delete p;
}

hipError_t rt_seg_build_create(rt_ctx* ctx, const rt_batch_item* items, int count, rt_seg_build** out) {
  *out = nullptr;
  if (count < 1) return hipSuccess;
  const int bits = items[0].bvh->bits;
  for (int i = 0; i < count; ++i) {
    const rt_bvh* b = items[i].bvh;
    if (b->bits != bits || b->n < 1 || (b->build_flags & RT_BUILD_KARRAS_SEARCH)) return hipSuccess;
  }
  rt_seg_build* p = new (std::nothrow) rt_seg_build();
  if (!p) return hipErrorOutOfMemory;
  p->ns = count;
  p->bits = bits;
  std::vector<SegBvh> segs(count);
  std::vector<uint32_t> off((size_t)kSegStages * (count + 1), 0u);
  for (int i = 0; i < count; ++i) {
    const rt_bvh* b = items[i].bvh;
    const rt_mesh_view& m = items[i].mesh;
    const int n = b->n;
    SegBvh& g = segs[i];
    g.n = n;
    g.ntiles = b->ntiles;
    g.zero_words = (size_t)b->passes * 256 + (size_t)b->passes * b->ntiles * 256 + (size_t)(n > 1 ? n - 1 : 1);
    g.vx = m.vx_dev;
    g.vy = m.vy_dev;
    g.vz = m.vz_dev;
    g.idx = m.idx_dev;
    g.tri_box = b->tri_box;
    g.trirec = b->trirec;
    g.misc = b->misc;
    g.hist = b->hist;
    g.status = b->status;
    g.arrivals = b->arrivals;
    g.deferred = b->deferred;
    g.deferred2 = b->deferred2;
    g.wscan = b->wscan;
    g.keys[0] = b->keys[0];
    g.keys[1] = b->keys[1];
    g.ids[0] = b->ids[0];
    g.ids[1] = b->ids[1];
    g.parent = b->parent;
    g.left = b->left;
    g.right = b->right;
    g.leaf_of = b->leaf_of;
    g.wcomp = b->wcomp;
    g.node_box = b->node_box;
    g.tnode = nullptr;  // the segmented build runs only while no BVH wants binary records
    g.ltri = b->ltri;
    g.deposits = b->deposits;
    g.wideq = b->wideq;
    g.wsrc = b->wsrc;
    g.wide_flag = b->wide_flag;
    // workgroups of each stage, as the single-BVH launches use them
    const int wide = n >= 2;
    const uint32_t blocks[kSegStages] = {
        (uint32_t)std::min(cdiv((long)g.zero_words, 256), 2048),
        (uint32_t)std::min(cdiv(n, 256), 1024),
        (uint32_t)cdiv(n, kMortonPerBlock),
        (uint32_t)b->ntiles,
        (uint32_t)cdiv(n, kClimbChunk),
        n > 1 ? (uint32_t)top_pass_grid(n) : 0u,
        n > 1 ? 1u : 0u,
        wide ? (uint32_t)cdiv(n - 1, 256) : 0u,
        wide ? (uint32_t)cdiv(n - 1, 1024) : 0u,
        wide ? 1u : 0u,
        wide ? (uint32_t)cdiv(n - 1, 1024) : 0u,
        wide ? (uint32_t)cdiv(n - 1, 256) : 0u};
    for (int st = 0; st < kSegStages; ++st) {
      uint32_t* o = off.data() + (size_t)st * (count + 1);
      if ((uint64_t)o[i] + blocks[st] >= (1ull << 31) / 1024) {  // grid too large for one launch
        delete p;
        return hipSuccess;
      }
      o[i + 1] = o[i] + blocks[st];
    }
    p->passes = std::max(p->passes, top_passes(n));
  }
  const size_t seg_bytes = rt_align256(sizeof(SegBvh) * count);
  hipError_t e = rt_dev_alloc(ctx, &p->dev, seg_bytes + off.size() * sizeof(uint32_t));
  if (e == hipSuccess) p->dev_bytes = seg_bytes + off.size() * sizeof(uint32_t);
  if (e == hipSuccess) e = hipMemcpy(p->dev, segs.data(), sizeof(SegBvh) * count, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy((char*)p->dev + seg_bytes, off.data(), off.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    rt_seg_build_destroy(ctx, p);
    return e;
  }
  p->segs = (const SegBvh*)p->dev;
  for (int st = 0; st < kSegStages; ++st) {
    p->off[st] = (const uint32_t*)((char*)p->dev + seg_bytes) + (size_t)st * (count + 1);
    p->total[st] = off[(size_t)st * (count + 1) + count];
  }
  *out = p;
  return hipSuccess;
}

namespace {
template <typename K>
void seg_build_typed(const rt_seg_build* p, hipStream_t s) {
  constexpr int P = KeyTraits<K>::kPasses;
  const SegBvh* g = p->segs;
  const int ns = p->ns;
#define RT_SEG(ST, KERNEL, THREADS, ...) \
  if (p->total[ST]) KERNEL<<<p->total[ST], THREADS, 0, s>>>(g, p->off[ST], ns, ##__VA_ARGS__)
  RT_SEG(kStInit, k_seg_init, 256);
  RT_SEG(kStAabb, k_seg_aabb_bounds, 256);
  RT_SEG(kStMorton, k_seg_morton<K>, 256);
  for (int q = 0; q < P; ++q) RT_SEG(kStSort, k_seg_onesweep<K>, kSortThreads, q);
  RT_SEG(kStChunks, k_seg_climb_chunks<K>, kClimbThreads);
  for (int q = 0; q < p->passes; ++q) RT_SEG(kStPass, k_seg_climb_pass<K>, 256, q);
  RT_SEG(kStTop, k_seg_climb_top<K>, kClimbTopThreads, p->passes);
  RT_SEG(kStParity, k_seg_wide_parity, 256);
  RT_SEG(kStPackCount, k_seg_pack_count, 1024);
  RT_SEG(kStPackOffsets, k_seg_pack_offsets, 1024);
  RT_SEG(kStPackIndex, k_seg_pack_index, 1024);
  RT_SEG(kStWideFill, k_seg_wide_fill, 256);
#undef RT_SEG
}
}  // namespace

hipError_t rt_seg_build_launch(const rt_seg_build* p, hipStream_t s) {
  if (p->bits == 30)
    seg_build_typed<uint32_t>(p, s);
  else
    seg_build_typed<uint64_t>(p, s);
  return hipGetLastError();
}

hipError_t rt_launch_build(rt_bvh* b, const rt_mesh_view* m, hipStream_t s) {
  return b->bits == 30 ? build_typed<uint32_t>(b, m, s) : build_typed<uint64_t>(b, m, s);
}

hipError_t rt_launch_refit(rt_bvh* b, const rt_mesh_view* m, hipStream_t s) {
  return b->bits == 30 ? refit_typed<uint32_t>(b, m, s) : refit_typed<uint64_t>(b, m, s);
}

hipError_t rt_launch_export(const rt_bvh* b, const rt_bvh_export_bufs* d, hipStream_t s) {
  const int n = b->n;
  if (b->bits == 30)
    k_export<uint32_t><<<cdiv(2L * n - 1, 256), 256, 0, s>>>(
        n, (const uint32_t*)b->keys[0], b->ids[0], b->parent, b->left, b->right, b->node_box, b->ltri,
        (uint32_t*)d->codes_dev, d->ids_dev, d->parent_dev, d->left_dev, d->right_dev, d->lo_dev, d->hi_dev,
        d->tri_lo_dev, d->tri_hi_dev);
  else
    k_export<uint64_t><<<cdiv(2L * n - 1, 256), 256, 0, s>>>(
        n, (const uint64_t*)b->keys[0], b->ids[0], b->parent, b->left, b->right, b->node_box, b->ltri,
        (uint64_t*)d->codes_dev, d->ids_dev, d->parent_dev, d->left_dev, d->right_dev, d->lo_dev, d->hi_dev,
        d->tri_lo_dev, d->tri_hi_dev);
  return hipGetLastError();
}
