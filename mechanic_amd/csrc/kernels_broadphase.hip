// Broadphase: Morton keys and AABBs (with analytic ground contacts fused in),
// an LSD radix sort (or an LDS-staged bitonic sort), Karras LBVH topology with
// the refit of each workgroup's subtrees in LDS and of the nodes spanning
// workgroups from prefix/suffix boxes, and a packet pair traversal with
// wavefront-ballot-compacted output.
#include "internal.h"
#include "mh_geometry.h"
#include "mh_math.h"

namespace mh {

__device__ __forceinline__ uint64_t tick_now(const Dev& d) {
  return *reinterpret_cast<const volatile uint64_t*>(&d.tick->tick_index);
}

// expand_morton_bits / morton_code (lbvh.wgsl:84-98).
__device__ __forceinline__ uint32_t expand_bits(uint32_t v) {
  v &= 0x000003ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}
__device__ __forceinline__ uint32_t quantize(float c) {
  return (uint32_t)fmn(fmx(floorf(c + 512.0f), 0.0f), 1023.0f);
}
__device__ __forceinline__ uint32_t morton_code(v3 c) {
  return expand_bits(quantize(c.x)) | (expand_bits(quantize(c.y)) << 1) | (expand_bits(quantize(c.z)) << 2);
}

// Digit of a sort key in one LSD radix pass (8 bits above the collider index).
__device__ __forceinline__ uint32_t radix_digit(const Dev& d, uint64_t key, uint32_t pass) {
  return (uint32_t)(key >> (d.key_index_bits + 8u * pass)) & 0xffu;
}

// compute_morton (lbvh.wgsl:108-130) + generate_ground_contacts
// (collision.wgsl:622-664), the clear of this tick's manifold index and the
// radix sort's digit counts. One thread per sort slot; slots past the collider
// count get the padding key (0xffffffff, slot).
__global__ __launch_bounds__(256) void k_morton(Dev d) {
  __shared__ uint32_t lds_append[kBlock / 64 + 1];
  __shared__ uint32_t lds_hist[kRadixMaxPasses * 256];  // this workgroup's digit counts per radix pass
  const uint64_t tick = tick_now(d);
  uint32_t* diag = d.diag[tick & 1];
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < kCounterWords) d.counters[i] = 0u;  // sort_count >= 256 threads always run
  if (i < (d.collider_count + 63u) / 64u) d.work_count[i] = 0u;  // packets' item counts, unpublished
  {
    // This tick's manifold index starts empty (prepare_contacts inserts into it;
    // the previous tick's index, read by the lookups, is the other one).
    ulonglong2* table = reinterpret_cast<ulonglong2*>(d.htab[tick & 1]);
    const uint32_t words = (d.hash_mask + 1u) / 2u, stride = gridDim.x * blockDim.x;
    for (uint32_t j = i; j < words; j += stride) table[j] = make_ulonglong2(0ull, 0ull);
  }
  const bool radix = d.t_sort_kind == 0u;
  if (radix) {
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t status_words = d.rs_passes * d.rs_blocks * 256u;
    for (uint32_t j = i; j < status_words; j += stride) d.rs_status[j] = 0u;
    for (uint32_t j = threadIdx.x; j < d.rs_passes * 256u; j += blockDim.x) lds_hist[j] = 0u;
  }
  __syncthreads();
  bool emit = false;
  ColliderPose p;
  v3 support;
  float bottom = 0.0f;
  if (i < d.sort_count && i >= d.collider_count) d.keys[i] = (0xffffffffull << 32) | i;
  if (i < d.collider_count) {
    p = collider_pose(d, i);
    store_pose(d, i, p);
    v3 we = (vabs(p.axes[0]) * p.half.x + vabs(p.axes[1]) * p.half.y) + vabs(p.axes[2]) * p.half.z;
    d.aabbs[2 * (size_t)i] = f4(p.center - we, 0.0f);
    d.aabbs[2 * (size_t)i + 1] = f4(p.center + we, 0.0f);
    // Sort key: (scene, Morton code, collider); a single scene has no scene field
    // and a 32-bit collider field, the reference's (code << 32 | collider).
    const uint64_t code = morton_code(p.center);
    // The radix sort starts from the buffer its last pass does not write.
    uint64_t* keys = radix && (d.rs_passes & 1u) ? d.keys_alt : d.keys;
    const uint64_t key = d.scene_count > 1
                             ? ((uint64_t)d.cscene[i] << (30u + d.key_index_bits)) | (code << d.key_index_bits) | i
                             : (code << 32) | i;
    keys[i] = key;
    if (radix)
      for (uint32_t pass = 0; pass < d.rs_passes; ++pass)
        atomicAdd(&lds_hist[pass * 256u + radix_digit(d, key, pass)], 1u);
    // Ground plane y = 0.
    const v3 down = mk(0.0f, -1.0f, 0.0f);
    support = support_point(p, down);
    emit = true;
    if (p.role != 0u) {
      uint32_t count = full_cylinder_contact_count(p, down);
      if (p.role > count) emit = false;
      else support = full_cylinder_support_point(p, down, p.role);
    }
    bottom = support.y;
    if (bottom > 1.0e-5f) emit = false;
  }
  // The grid covers sort_count (a multiple of 256) exactly, so every thread of
  // every block reaches the block-wide append.
  uint32_t row = block_append(&diag[kDiagContacts], emit, lds_append);
  // The append's barriers also complete the digit counts: add them to the
  // radix sort's histogram (zeroed by k_integrate at the start of the tick).
  if (radix)
    for (uint32_t j = threadIdx.x; j < d.rs_passes * 256u; j += blockDim.x) {
      const uint32_t n = lds_hist[j];
      if (n) atomicAdd(&d.rs_hist[j], n);
    }
  if (!emit) return;
  if (row < d.pair_capacity) {
    float4 bp = d.pos[p.body];
    store_contact(d, row, make_uint4(p.body, kInvalid, i, kInvalid),
                  make_float4(0.0f, -1.0f, 0.0f, fmx(-bottom, 0.0f)),
                  f4(support - xyz(bp), 0.0f),
                  make_float4(0.0f, 0.0f, 0.0f, p.role != 0u ? -1.0f : 0.0f));
  } else {
    atomicOr(&diag[kDiagFlags], 1u);
  }
}

// ---------------------------------------------------------------------------
// Bitonic sort of sort_count u64 keys, ascending. Keys are unique, so the result
// equals the reference's order (lbvh.wgsl:100-106, 132-217).
//
// A tile of up to kTile keys lives in one workgroup (256 threads, 16 KiB of
// LDS): 8 keys per thread in registers
// (blocked: thread t holds tile positions 8t..8t+7). Exchanges with j < 8 stay
// in registers, 8 <= j < 512 go through wavefront shuffles (partner lane
// t ^ j/8 is in the same 64-lane wavefront), and j >= 512 through LDS.
// Distances of a tile or more are global multi-level butterfly kernels.
// ---------------------------------------------------------------------------
constexpr uint32_t kPer = 8;
constexpr uint32_t kTile = 2048;

__device__ __forceinline__ void cas(uint64_t& a, uint64_t& b, bool ascending) {
  bool swap = ascending ? (a > b) : (a < b);
  uint64_t lo = swap ? b : a;
  uint64_t hi = swap ? a : b;
  a = lo;
  b = hi;
}

__device__ __forceinline__ uint64_t shfl_xor64(uint64_t v, int m) {
  uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
  lo = (uint32_t)__shfl_xor((int)lo, m);
  hi = (uint32_t)__shfl_xor((int)hi, m);
  return ((uint64_t)hi << 32) | lo;
}

// One merge level at distance j inside a tile, keys in registers.
// `base` is the global position of this thread's first key.
__device__ __forceinline__ void tile_level(uint64_t (&r)[kPer], uint64_t* lds, uint32_t tid, uint32_t nthreads,
                                           uint32_t tile_base, uint32_t base, uint32_t k, uint32_t j) {
  if (j < kPer) {
#pragma unroll
    for (uint32_t e = 0; e < kPer; ++e) {
      uint32_t f = e ^ j;
      if (f > e) cas(r[e], r[f], ((base + e) & k) == 0);
    }
  } else if (j < kPer * 64) {
    int m = (int)(j / kPer);
    bool lower = (tid & (uint32_t)m) == 0;
#pragma unroll
    for (uint32_t e = 0; e < kPer; ++e) {
      uint64_t other = shfl_xor64(r[e], m);
      bool ascending = ((base + e) & k) == 0;
      uint64_t mn = r[e] < other ? r[e] : other;
      uint64_t mx = r[e] < other ? other : r[e];
      r[e] = (lower == ascending) ? mn : mx;
    }
  } else {
#pragma unroll
    for (uint32_t e = 0; e < kPer; ++e) lds[tid * kPer + e] = r[e];
    __syncthreads();
    uint32_t pairs = nthreads * kPer / 2;
    for (uint32_t q = tid; q < pairs; q += nthreads) {
      uint32_t lo = ((q / j) * 2 * j) + (q % j);
      uint32_t hi = lo + j;
      uint64_t a = lds[lo], b = lds[hi];
      cas(a, b, ((tile_base + lo) & k) == 0);
      lds[lo] = a;
      lds[hi] = b;
    }
    __syncthreads();
#pragma unroll
    for (uint32_t e = 0; e < kPer; ++e) r[e] = lds[tid * kPer + e];
    __syncthreads();
  }
}

// Full bitonic sort of each tile (stages k = 2 .. tile), alternating directions by
// the global position so neighbouring tiles form bitonic sequences.
__global__ __launch_bounds__(256) void k_sort_tile(uint64_t* keys, uint32_t tile) {
  __shared__ uint64_t lds[kTile];
  const uint32_t tid = threadIdx.x;
  const uint32_t nthreads = blockDim.x;
  const uint32_t tile_base = blockIdx.x * tile;
  const uint32_t base = tile_base + tid * kPer;
  uint64_t r[kPer];
#pragma unroll
  for (uint32_t e = 0; e < kPer; ++e) r[e] = keys[base + e];
  for (uint32_t k = 2; k <= tile; k <<= 1)
    for (uint32_t j = k >> 1; j > 0; j >>= 1) tile_level(r, lds, tid, nthreads, tile_base, base, k, j);
#pragma unroll
  for (uint32_t e = 0; e < kPer; ++e) keys[base + e] = r[e];
}

// Finishing levels j = tile/2 .. 1 of stage k inside each tile.
__global__ __launch_bounds__(256) void k_sort_merge(uint64_t* keys, uint32_t tile, uint32_t k) {
  __shared__ uint64_t lds[kTile];
  const uint32_t tid = threadIdx.x;
  const uint32_t nthreads = blockDim.x;
  const uint32_t tile_base = blockIdx.x * tile;
  const uint32_t base = tile_base + tid * kPer;
  uint64_t r[kPer];
#pragma unroll
  for (uint32_t e = 0; e < kPer; ++e) r[e] = keys[base + e];
  for (uint32_t j = tile >> 1; j > 0; j >>= 1) tile_level(r, lds, tid, nthreads, tile_base, base, k, j);
#pragma unroll
  for (uint32_t e = 0; e < kPer; ++e) keys[base + e] = r[e];
}

// `levels` consecutive levels j_hi, j_hi/2, ... of stage k: each thread owns the
// 2^levels keys i0 + l * j_lo and runs the butterflies in registers.
template <int L>
__global__ __launch_bounds__(256) void k_sort_global(uint64_t* keys, uint32_t count, uint32_t k, uint32_t j_hi) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t n = 1u << L;
  if (t >= (count >> L)) return;
  const uint32_t j_lo = j_hi >> (L - 1);
  const uint32_t lj = 31u - __clz(j_lo);
  const uint32_t i0 = ((t >> lj) << (lj + L)) | (t & (j_lo - 1));
  const bool ascending = (i0 & k) == 0;
  uint64_t r[1 << L];
#pragma unroll
  for (uint32_t l = 0; l < n; ++l) r[l] = keys[i0 + l * j_lo];
#pragma unroll
  for (int lev = L - 1; lev >= 0; --lev) {
    const uint32_t s = 1u << lev;
#pragma unroll
    for (uint32_t l = 0; l < n; ++l)
      if ((l & s) == 0) cas(r[l], r[l | s], ascending);
  }
#pragma unroll
  for (uint32_t l = 0; l < n; ++l) keys[i0 + l * j_lo] = r[l];
}

// ---------------------------------------------------------------------------
// Radix sort of the collider_count real keys (LSD, 8-bit digits) over the bits
// above the collider field. k_morton writes the keys in collider order and every
// pass is stable, so equal (scene, code) fields keep collider order: the result
// is the order of the full keys, and the padding slots past collider_count are
// already in place.
//
// One launch counts every pass's digits (a permutation leaves them unchanged),
// then each pass is one launch with decoupled look-back (Merrill & Garland,
// "Single-pass Parallel Prefix Scan with Decoupled Look-back", 2016): a
// workgroup takes the next tile ticket, ranks its 1024 keys stably with
// wavefront ballots, publishes its per-digit counts, and sums the counts of the
// tiles before it from their published words. A tile only waits for tiles with
// smaller tickets, which started before it, so every wait ends.
// ---------------------------------------------------------------------------
constexpr uint32_t kRsAggregate = 1u << 30;  // status word: the tile's own count
constexpr uint32_t kRsPrefix = 2u << 30;     // status word: count over tiles 0..this one
constexpr uint32_t kRsCount = (1u << 30) - 1u;
constexpr uint32_t kRsSpinLimit = 1u << 22;  // reads of an unpublished word before giving up
// A packet's published item count: the count in the low bits and a published bit.
constexpr uint32_t kWalkCountMask = 0xffu;
constexpr uint32_t kWalkPublished = 0x100u;

// Exclusive prefix sum over a 256-thread workgroup.
__device__ __forceinline__ uint32_t exclusive_scan256(uint32_t v, uint32_t* wave_sums) {
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (uint32_t off = 1; off < 64u; off <<= 1) {
    const uint32_t y = (uint32_t)__shfl_up((int)x, off);
    if (lane >= off) x += y;
  }
  if (lane == 63u) wave_sums[wave] = x;
  __syncthreads();
  uint32_t before = 0;
  for (uint32_t w = 0; w < wave; ++w) before += wave_sums[w];
  return before + x - v;
}

__global__ __launch_bounds__(256) void k_radix_pass(Dev d, const uint64_t* src, uint64_t* dst, uint32_t pass) {
  constexpr uint32_t kRows = kRadixTile / 256;
  __shared__ uint32_t offset[kRows][4][256];  // keys of digit t before (row, wave) within the tile
  __shared__ uint32_t base[256];
  __shared__ uint32_t wave_sums[4];
  __shared__ uint32_t ticket;
  const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
  if (t == 0) ticket = atomicAdd(&d.rs_hist[kRadixTicket + pass], 1u);
#pragma unroll
  for (uint32_t r = 0; r < kRows; ++r)
#pragma unroll
    for (uint32_t w = 0; w < 4; ++w) offset[r][w][t] = 0u;
  __syncthreads();
  const uint32_t tile = ticket;
  const uint32_t first = tile * kRadixTile;
  const uint64_t lower = lane ? (~0ull >> (64u - lane)) : 0ull;
  uint64_t key[kRows];
  uint32_t digit[kRows], rank[kRows];
#pragma unroll
  for (uint32_t r = 0; r < kRows; ++r) {
    const uint32_t i = first + r * 256u + t;
    const bool valid = i < d.collider_count;
    key[r] = valid ? src[i] : 0ull;
    digit[r] = valid ? radix_digit(d, key[r], pass) : 256u;
    // Lanes holding the same digit.
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (uint32_t b = 0; b < 8; ++b) {
      const uint64_t ones = __ballot((digit[r] >> b) & 1u);
      peers &= ((digit[r] >> b) & 1u) ? ones : ~ones;
    }
    rank[r] = (uint32_t)__popcll(peers & lower);
    if (valid && rank[r] == 0u) offset[r][wave][digit[r]] = (uint32_t)__popcll(peers);
  }
  __syncthreads();
  // Thread t owns digit t: counts become offsets in tile order (row, then wave).
  uint32_t run = 0;
#pragma unroll
  for (uint32_t r = 0; r < kRows; ++r)
#pragma unroll
    for (uint32_t w = 0; w < 4; ++w) {
      const uint32_t c = offset[r][w][t];
      offset[r][w][t] = run;
      run += c;
    }
  uint32_t* status = d.rs_status + (size_t)pass * d.rs_blocks * 256u;
  uint32_t before = 0;
  if (tile == 0) {
    store_status(&status[t], kRsPrefix | run);
  } else {
    store_status(&status[(size_t)tile * 256u + t], kRsAggregate | run);
    uint32_t j = tile - 1u, spins = 0;
    for (;;) {
      const uint32_t v = load_status(&status[(size_t)j * 256u + t]);
      if ((v & ~kRsCount) == 0u) {
        if (++spins >= kRsSpinLimit) {
          atomicOr(&d.diag[tick_now(d) & 1][kDiagFlags], 1u);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
        continue;
      }
      before += v & kRsCount;
      if (v & kRsPrefix) break;
      --j;  // tile 0 always publishes a prefix, so j stays >= 0
    }
    store_status(&status[(size_t)tile * 256u + t], kRsPrefix | (before + run));
  }
  // Keys of smaller digits over all tiles, plus keys of digit t in earlier tiles.
  const uint32_t smaller = exclusive_scan256(d.rs_hist[pass * 256u + t], wave_sums);
  base[t] = smaller + before;
  __syncthreads();
#pragma unroll
  for (uint32_t r = 0; r < kRows; ++r) {
    if (digit[r] >= 256u) continue;
    const uint32_t to = base[digit[r]] + offset[r][wave][digit[r]] + rank[r];
    if (to < d.collider_count) dst[to] = key[r];
  }
}

// ---------------------------------------------------------------------------
// LBVH (Karras 2012) over the sorted keys, lbvh.wgsl:219-319.
// Internal node n covers a contiguous range of sorted leaves; its record holds
// (min, left word) and (max, escape). A child or escape is encoded as a sorted
// position with kLeafBit set for leaves; the escape of a node whose range ends
// at b is the node right after its subtree in depth-first order (the right
// child of the node split at b), or the sentinel C after the last leaf.
// A node whose leaves all share one Morton code is a bucket: its left word is
// its first leaf with kBucketBit, and traversal scans its leaves directly
// instead of descending the subtree the reference splits by collider index.
//
// k_build: one workgroup per kBlock sorted positions builds the internal nodes
// and the leaves at those positions, and the boxes of the nodes whose leaves
// all lie in its own kBlock leaves, climbing its subtrees in LDS. The sorted
// keys within kBlock of its positions are staged in LDS, which holds every
// probe of a node whose range is shorter than kBlock; the few longer nodes are
// searched by a whole wavefront, 64 probes a step. k_refit_top then gives the
// nodes whose leaves span several workgroups their boxes.
// ---------------------------------------------------------------------------
constexpr int kKeyWindow = 3 * (int)kBlock;  // sorted keys [first - kBlock, first + 2 kBlock)
constexpr uint16_t kNoParent = 0xffffu;
constexpr uint16_t kLocalLeaf = 0x8000u;  // local child word: a leaf, else an internal node

// Sorted keys, read from the LDS window when the position lies in it.
// The window as an LDS pointer: with a generic one the compiler folds the two
// loads of at() into one flat load from a selected address.
typedef const __attribute__((address_space(3))) uint64_t* LdsKeys;

struct KeyView {
  LdsKeys lds;
  const uint64_t* keys;
  int base, c;
  __device__ __forceinline__ uint64_t at(int j) const {
    const unsigned w = (unsigned)(j - base);
    return w < (unsigned)kKeyWindow ? lds[w] : keys[j];
  }
  // Common prefix length of key `ka` (at some position a != b) and key b; -1 off the ends.
  __device__ __forceinline__ int prefix(uint64_t ka, int b) const {
    if (b < 0 || b >= c) return -1;
    return __clzll((long long)(ka ^ at(b)));
  }
  // Escape of a node whose range ends at sorted position b (C = leaf count).
  __device__ __forceinline__ uint32_t escape_after(int b) const {
    if (b >= c - 1) return (uint32_t)c;
    const int n = b + 1;
    if (n <= c - 2) {
      const uint64_t kn = at(n);
      // Internal node n spans [n, ...] iff its direction is +1.
      if (prefix(kn, n + 1) > prefix(kn, n - 1)) return (uint32_t)n;
    }
    return (uint32_t)n | kLeafBit;
  }
};

// Largest x in [lo, hi) with prefix(key i, i + x dir) > threshold, given that it
// holds at lo: the common prefix of sorted keys with key i only shrinks away from
// i, so the x where it holds form a run from lo. Each step probes 64 evenly
// spaced x, one per lane, and keeps the last that holds; this is the value the
// sequential doubling and halving searches of build_topology find.
__device__ __forceinline__ int wave_search(const KeyView& kv, uint64_t ki, int i, int dir, int threshold, int lo,
                                           int hi) {
  const int lane = (int)__lane_id();
  while (hi - lo > 1) {
    const int step = (hi - lo + 63) / 64;
    const int x = lo + step * (lane + 1);
    const bool holds = x < hi && kv.prefix(ki, i + x * dir) > threshold;
    const uint64_t mask = __ballot(holds);
    const int best = mask ? lo + step * (64 - __clzll((long long)mask)) : lo;
    hi = best + step < hi ? best + step : hi;
    lo = best;
  }
  return lo;
}

// Box unions use min/max that order -0 below +0, so a union gives the same bits
// in any grouping; the tree's pairwise merges and the range unions below agree.
__device__ __forceinline__ float umin(float a, float b) {
  return a < b ? a : (b < a ? b : (__builtin_signbit(a) ? a : b));
}
__device__ __forceinline__ float umax(float a, float b) {
  return a > b ? a : (b > a ? b : (__builtin_signbit(a) ? b : a));
}

struct Box {
  float x0, y0, z0, x1, y1, z1;
};
__device__ __forceinline__ Box empty_box() {
  const float inf = __builtin_huge_valf();
  return Box{inf, inf, inf, -inf, -inf, -inf};
}
__device__ __forceinline__ void grow(Box& b, float4 mn, float4 mx) {
  b.x0 = umin(b.x0, mn.x);
  b.y0 = umin(b.y0, mn.y);
  b.z0 = umin(b.z0, mn.z);
  b.x1 = umax(b.x1, mx.x);
  b.y1 = umax(b.y1, mx.y);
  b.z1 = umax(b.z1, mx.z);
}
__device__ __forceinline__ Box unite(const Box& a, const Box& b) {
  return Box{umin(a.x0, b.x0), umin(a.y0, b.y0), umin(a.z0, b.z0),
             umax(a.x1, b.x1), umax(a.y1, b.y1), umax(a.z1, b.z1)};
}
__device__ __forceinline__ Box shfl_box(const Box& b, int lane) {
  return Box{__shfl(b.x0, lane), __shfl(b.y0, lane), __shfl(b.z0, lane),
             __shfl(b.x1, lane), __shfl(b.y1, lane), __shfl(b.z1, lane)};
}
__device__ __forceinline__ Box wave_union(Box b) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    b.x0 = umin(b.x0, __shfl_xor(b.x0, m));
    b.y0 = umin(b.y0, __shfl_xor(b.y0, m));
    b.z0 = umin(b.z0, __shfl_xor(b.z0, m));
    b.x1 = umax(b.x1, __shfl_xor(b.x1, m));
    b.y1 = umax(b.y1, __shfl_xor(b.y1, m));
    b.z1 = umax(b.z1, __shfl_xor(b.z1, m));
  }
  return b;
}

struct BuildLds {
  uint64_t keys[kKeyWindow];
  Box leaf_box[kBlock];
  Box node_box[kBlock];
  uint16_t child[kBlock][2];   // in-workgroup nodes: local child words
  uint16_t leaf_parent[kBlock];
  uint16_t node_parent[kBlock];
  uint32_t visit[kBlock];
  int2 searched[kBlock];       // long nodes: (length, split offset)
  uint32_t queue[kBlock];
  uint32_t queued;
  Box wave_total[kBlock / 64];
  uint32_t append[kBlock / 64 + 1];
};

// build_topology + prepare_leaves + build_bounds (lbvh.wgsl:233-319) for the
// nodes and leaves at sorted positions [blockIdx.x kBlock, +kBlock).
__global__ __launch_bounds__(256) void k_build(Dev d) {
  __shared__ BuildLds s;
  const uint64_t tick = tick_now(d);
  uint32_t* diag = d.diag[tick & 1];
  const int c = (int)d.collider_count;
  const uint32_t t = threadIdx.x;
  const int lane = (int)(t & 63u), wave = (int)(t >> 6);
  const int first_pos = (int)(blockIdx.x * kBlock);
  const int i = first_pos + (int)t;  // this thread's leaf and internal node
  const int block_last = min(first_pos + (int)kBlock, c) - 1;
  const KeyView kv{(LdsKeys)s.keys, d.keys, first_pos - (int)kBlock, c};
  for (int w = (int)t; w < kKeyWindow; w += (int)kBlock) {
    const int j = kv.base + w;
    if (j >= 0 && j < c) s.keys[w] = d.keys[j];
  }
  s.leaf_parent[t] = kNoParent;
  s.node_parent[t] = kNoParent;
  s.visit[t] = 0u;
  if (t == 0) s.queued = 0u;
  const bool leaf_valid = i < c;
  uint32_t collider = 0;
  float4 mn = make_float4(0.0f, 0.0f, 0.0f, 0.0f), mx = mn;
  if (leaf_valid) {
    collider = (uint32_t)(d.keys[i] & (~0ull >> (64u - d.key_index_bits)));
    mn = d.aabbs[2 * (size_t)collider];
    mx = d.aabbs[2 * (size_t)collider + 1];
  }
  __syncthreads();

  // Topology. The doubling search stays in the window up to a length of kBlock.
  const bool node_valid = i < c - 1;
  int dir = 1, len = 0, split_off = 0;
  bool searched = false;
  if (node_valid) {
    const uint64_t ki = s.keys[t + kBlock];
    dir = kv.prefix(ki, i + 1) > kv.prefix(ki, i - 1) ? 1 : -1;
    const int min_prefix = kv.prefix(ki, i - dir);
    int max_len = 2;
    while (max_len <= (int)kBlock && kv.prefix(ki, i + max_len * dir) > min_prefix) max_len *= 2;
    if (max_len > (int)kBlock) {
      searched = true;
      s.queue[atomicAdd(&s.queued, 1u)] = t;
    } else {
      for (int step = max_len / 2; step >= 1; step /= 2)
        if (kv.prefix(ki, i + (len + step) * dir) > min_prefix) len += step;
      const int node_prefix = kv.prefix(ki, i + len * dir);
      int step = (len + 1) / 2;
      while (step >= 1) {
        if (kv.prefix(ki, i + (split_off + step) * dir) > node_prefix) split_off += step;
        if (step == 1) break;
        step = (step + 1) / 2;
      }
    }
  }
  __syncthreads();
  const uint32_t queued = s.queued;
  for (uint32_t q = (uint32_t)wave; q < queued; q += kBlock / 64) {
    const int tl = (int)s.queue[q];
    const int n = first_pos + tl;
    const uint64_t kn = s.keys[tl + (int)kBlock];
    const int ndir = kv.prefix(kn, n + 1) > kv.prefix(kn, n - 1) ? 1 : -1;
    const int min_prefix = kv.prefix(kn, n - ndir);
    const int bound = ndir > 0 ? c - 1 - n : n;
    const int nlen = wave_search(kv, kn, n, ndir, min_prefix, 0, bound + 1);
    const int nsplit = wave_search(kv, kn, n, ndir, kv.prefix(kn, n + nlen * ndir), 0, nlen);
    if (lane == 0) s.searched[tl] = make_int2(nlen, nsplit);
  }
  __syncthreads();

  uint32_t left_word = 0, escape = 0;
  bool in_block = false;
  if (node_valid) {
    if (searched) {
      len = s.searched[t].x;
      split_off = s.searched[t].y;
    }
    const int other = i + len * dir;
    const int split = i + split_off * dir + (dir < 0 ? dir : 0);
    const int first = i < other ? i : other;
    const int last = i < other ? other : i;
    const uint32_t leaf_base = (uint32_t)c - 1u;
    const uint32_t left_ref = split == first ? leaf_base + (uint32_t)split : (uint32_t)split;
    const uint32_t right_ref = split + 1 == last ? leaf_base + (uint32_t)(split + 1) : (uint32_t)(split + 1);
    d.children[i] = make_uint2(left_ref, right_ref);
    d.ranges[i] = make_uint2((uint32_t)first, (uint32_t)last);
    d.parents[left_ref] = (uint32_t)i + 1u;
    d.parents[right_ref] = (uint32_t)i + 1u;
    if (i == 0) d.parents[0] = 0u;
    const bool bucket = (kv.at(first) >> d.key_index_bits) == (kv.at(last) >> d.key_index_bits);
    const uint32_t left_enc = split == first ? ((uint32_t)split | kLeafBit) : (uint32_t)split;
    left_word = bucket ? ((uint32_t)first | kBucketBit) : left_enc;
    escape = kv.escape_after(last);
    in_block = first >= first_pos && last <= block_last;
    if (in_block) {
      const int ls = split - first_pos, rs = split + 1 - first_pos;
      s.child[t][0] = (uint16_t)(split == first ? (kLocalLeaf | ls) : ls);
      s.child[t][1] = (uint16_t)(split + 1 == last ? (kLocalLeaf | rs) : rs);
      if (split == first) s.leaf_parent[ls] = (uint16_t)t;
      else s.node_parent[ls] = (uint16_t)t;
      if (split + 1 == last) s.leaf_parent[rs] = (uint16_t)t;
      else s.node_parent[rs] = (uint16_t)t;
    }
  }
  Box box = empty_box();
  if (leaf_valid) {
    d.leaves[2 * (size_t)i] = make_float4(mn.x, mn.y, mn.z, __uint_as_float(collider));
    d.leaves[2 * (size_t)i + 1] = make_float4(mx.x, mx.y, mx.z, __uint_as_float(kv.escape_after(i)));
    d.leaf_body[i] = __float_as_uint(d.colliders[4 * (size_t)collider + 3].x);
    grow(box, mn, mx);
  }
  s.leaf_box[t] = box;
  // Nodes whose leaves span several workgroups, for k_refit_top. The append's
  // barriers also publish the LDS links and leaf boxes above.
  const bool top = node_valid && !in_block;
  const uint32_t slot = block_append(&d.counters[kCountTop], top, s.append);
  if (top) d.top[slot] = (uint32_t)i;

  // Refit inside the workgroup (build_bounds): the second thread to reach a
  // node merges its children and climbs on.
  if (leaf_valid && c > 1) {
    uint32_t p = s.leaf_parent[t];
    for (int level = 0; level < (int)kBlock && p != kNoParent; ++level) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      const uint32_t arrival = __hip_atomic_fetch_add(&s.visit[p], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      if (arrival == 0u) break;
      if (arrival > 1u) {
        atomicOr(&diag[kDiagFlags], 1u);
        break;
      }
      const uint32_t a = s.child[p][0], b = s.child[p][1];
      const Box ba = (a & kLocalLeaf) ? s.leaf_box[a & ~kLocalLeaf] : s.node_box[a];
      const Box bb = (b & kLocalLeaf) ? s.leaf_box[b & ~kLocalLeaf] : s.node_box[b];
      s.node_box[p] = unite(ba, bb);
      p = s.node_parent[p];
    }
  }

  // Union of the workgroup's leaves up to and from each leaf, and of all of them.
  Box pre = box, suf = box;
#pragma unroll
  for (int m = 1; m < 64; m <<= 1) {
    const Box up = shfl_box(pre, lane - m >= 0 ? lane - m : lane);
    const Box down = shfl_box(suf, lane + m < 64 ? lane + m : lane);
    if (lane >= m) pre = unite(pre, up);
    if (lane + m < 64) suf = unite(suf, down);
  }
  if (lane == 63) s.wave_total[wave] = pre;
  __syncthreads();
  Box all = empty_box();
#pragma unroll
  for (int w = 0; w < (int)(kBlock / 64); ++w) {
    const Box tw = s.wave_total[w];
    if (w < wave) pre = unite(pre, tw);
    if (w > wave) suf = unite(suf, tw);
    all = unite(all, tw);
  }
  if (leaf_valid) {
    float4* e = d.leaf_scan + 4 * (size_t)i;
    e[0] = make_float4(pre.x0, pre.y0, pre.z0, 0.0f);
    e[1] = make_float4(pre.x1, pre.y1, pre.z1, 0.0f);
    e[2] = make_float4(suf.x0, suf.y0, suf.z0, 0.0f);
    e[3] = make_float4(suf.x1, suf.y1, suf.z1, 0.0f);
  }
  if (t == 0) {
    d.block_box[2 * (size_t)blockIdx.x] = make_float4(all.x0, all.y0, all.z0, 0.0f);
    d.block_box[2 * (size_t)blockIdx.x + 1] = make_float4(all.x1, all.y1, all.z1, 0.0f);
  }
  if (node_valid) {
    const Box nb = in_block ? s.node_box[t] : empty_box();  // k_refit_top fills the others
    d.inodes[2 * (size_t)i] = make_float4(nb.x0, nb.y0, nb.z0, __uint_as_float(left_word));
    d.inodes[2 * (size_t)i + 1] = make_float4(nb.x1, nb.y1, nb.z1, __uint_as_float(escape));
  }
}

// Nodes whose leaves span several workgroups (listed by k_build): one wavefront
// per node unites the box of its leaves in its first workgroup (a suffix box),
// in its last (a prefix box) and the boxes of the whole workgroups between.
// No node depends on another, so there is no climb and no cross-workgroup handoff.
__global__ __launch_bounds__(256) void k_refit_top(Dev d) {
  const uint32_t n = *reinterpret_cast<volatile uint32_t*>(&d.counters[kCountTop]);
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t waves = gridDim.x * (blockDim.x >> 6);
  for (uint32_t i = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); i < n; i += waves) {
    const uint32_t node = d.top[i];
    const uint2 range = d.ranges[node];
    const uint32_t b0 = range.x / kBlock, b1 = range.y / kBlock;
    Box box = empty_box();
    if (lane == 0) {
      grow(box, d.leaf_scan[4 * (size_t)range.x + 2], d.leaf_scan[4 * (size_t)range.x + 3]);
      grow(box, d.leaf_scan[4 * (size_t)range.y], d.leaf_scan[4 * (size_t)range.y + 1]);
    }
    for (uint32_t b = b0 + 1 + lane; b < b1; b += 64u) grow(box, d.block_box[2 * (size_t)b], d.block_box[2 * (size_t)b + 1]);
    box = wave_union(box);
    if (lane == 0) {
      float4* p = d.inodes + 2 * (size_t)node;
      const float lw = p[0].w, ew = p[1].w;
      p[0] = make_float4(box.x0, box.y0, box.z0, lw);
      p[1] = make_float4(box.x1, box.y1, box.z1, ew);
    }
  }
}

// ---------------------------------------------------------------------------
// Pair traversal (lbvh.wgsl:321-398). A packet of 64 consecutive sorted leaves
// (queries, one per lane) walks the tree together: it descends into a node when
// any of its queries overlaps the node and the node holds leaves right of that
// query. At a bucket it scans the bucket's leaves, 64 at a time, each lane
// testing its query against every candidate broadcast from the lane that loaded
// it. A lane only pairs with leaves right of its own,
// so each unordered overlapping pair is found once and written as
// (lower collider, higher collider): the set the reference's per-leaf stack
// traversal emits. Pairs are staged in LDS per wavefront and flushed with one
// global atomic per batch.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool suppressed(const Dev& d, uint32_t body_a, uint32_t body_b) {
  uint32_t lo = body_a < body_b ? body_a : body_b;
  uint32_t hi = body_a < body_b ? body_b : body_a;
  uint32_t left = 0, right = d.suppression_count;
  while (left < right) {
    uint32_t mid = left + (right - left) / 2;
    uint2 cand = d.suppressed[mid];
    if (cand.x < lo || (cand.x == lo && cand.y < hi)) left = mid + 1;
    else right = mid;
  }
  return left < d.suppression_count && d.suppressed[left].x == lo && d.suppressed[left].y == hi;
}

// The same membership test from one body's side: its suppressed partners are a
// sorted list (supp_off / supp_adj), whose first four a query lane holds in
// registers for the whole packet.
struct Partners {
  uint4 first;           // first four partners, kInvalid-padded
  uint32_t more, end;    // the rest: supp_adj[more, end)
};

__device__ __forceinline__ Partners load_partners(const Dev& d, uint32_t body) {
  Partners r{make_uint4(kInvalid, kInvalid, kInvalid, kInvalid), 0u, 0u};
  const uint32_t begin = d.supp_off[body], end = d.supp_off[body + 1];
  const uint32_t n = end - begin;
  if (n > 0) r.first.x = d.supp_adj[begin];
  if (n > 1) r.first.y = d.supp_adj[begin + 1];
  if (n > 2) r.first.z = d.supp_adj[begin + 2];
  if (n > 3) r.first.w = d.supp_adj[begin + 3];
  r.more = n > 4 ? begin + 4 : end;
  r.end = end;
  return r;
}

__device__ __forceinline__ bool is_partner(const Dev& d, const Partners& p, uint32_t other) {
  bool found = other == p.first.x || other == p.first.y || other == p.first.z || other == p.first.w;
  for (uint32_t i = p.more; !found && i < p.end; ++i) found = d.supp_adj[i] == other;
  return found;
}

__device__ __forceinline__ bool overlaps(float4 qmn, float4 qmx, float4 mn, float4 mx) {
  return qmn.x <= mx.x && qmn.y <= mx.y && qmn.z <= mx.z && mn.x <= qmx.x && mn.y <= qmx.y && mn.z <= qmx.z;
}

__device__ __forceinline__ float bcast(float v, uint32_t lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), (int)lane));
}
__device__ __forceinline__ uint32_t bcast(uint32_t v, uint32_t lane) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)lane);
}

__device__ __forceinline__ float uniform(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ __forceinline__ float4 uniform4(float4 v) {
  return make_float4(uniform(v.x), uniform(v.y), uniform(v.z), uniform(v.w));
}

// Lanes [0, n) of a wavefront (n wave-uniform).
__device__ __forceinline__ uint64_t lanes_below(uint32_t n) { return n >= 64u ? ~0ull : (1ull << n) - 1ull; }

constexpr uint32_t kStage = 512;  // staged pairs per wavefront

struct PairStage {
  uint2* buf;
  uint32_t count;  // wave-uniform
};

__device__ __forceinline__ void flush_pairs(const Dev& d, uint32_t* diag, PairStage& st, uint32_t lane) {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_wave_barrier();
  uint32_t base = 0;
  if (lane == 0) base = atomicAdd(&diag[kDiagPairs], st.count);
  base = bcast(base, 0u);
  for (uint32_t i = lane; i < st.count; i += 64u) {
    const uint32_t row = base + i;
    if (row < d.pair_capacity) d.pairs[row] = st.buf[i];
  }
  if (lane == 0 && base + st.count > d.pair_capacity) atomicOr(&diag[kDiagFlags], 1u);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_wave_barrier();
  st.count = 0;
}

// One packet of queries: the wavefront's 64 consecutive sorted leaves.
struct Packet {
  uint32_t first;  // sorted position of lane 0's leaf
  uint32_t p;      // this lane's leaf
  bool valid;
  float4 qmn, qmx;
  uint32_t self, self_body;
  uint32_t scene, scene_end;  // this lane's scene and the end of its sorted leaf range
  float4 umn, umx;  // union box of the packet's queries (wave-uniform)
  Partners partners;  // the query body's suppressed partners (when the scene has any)
};

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = fmn(v, __shfl_xor(v, m));
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = fmx(v, __shfl_xor(v, m));
  return v;
}

template <bool kPartners = true>
__device__ __forceinline__ Packet load_packet(const Dev& d, uint32_t first, uint32_t lane) {
  Packet k;
  k.first = first;
  k.p = first + lane;
  k.valid = k.p < d.collider_count;
  k.qmn = make_float4(0.f, 0.f, 0.f, 0.f);
  k.qmx = k.qmn;
  k.self = 0;
  k.self_body = 0;
  k.scene = 0;
  k.scene_end = d.collider_count;
  if (k.valid) {
    k.qmn = d.leaves[2 * (size_t)k.p];
    k.qmx = d.leaves[2 * (size_t)k.p + 1];
    k.self = __float_as_uint(k.qmn.w);
    k.self_body = d.leaf_body[k.p];
    if (d.scene_count > 1) {
      k.scene = d.cscene[k.self];
      k.scene_end = d.scene_first[k.scene + 1];
    }
  }
  k.partners = Partners{make_uint4(kInvalid, kInvalid, kInvalid, kInvalid), 0u, 0u};
  if (kPartners && k.valid && d.suppression_count != 0) k.partners = load_partners(d, k.self_body);
  const float inf = __builtin_huge_valf();
  k.umn = make_float4(wave_min(k.valid ? k.qmn.x : inf), wave_min(k.valid ? k.qmn.y : inf),
                      wave_min(k.valid ? k.qmn.z : inf), 0.0f);
  k.umx = make_float4(wave_max(k.valid ? k.qmx.x : -inf), wave_max(k.valid ? k.qmx.y : -inf),
                      wave_max(k.valid ? k.qmx.z : -inf), 0.0f);
  return k;
}

// Tests every lane's query against the leaves [start, last], 64 at a time: each
// lane loads one candidate, candidates that miss the packet's union box are
// dropped by one ballot, and all lanes test each remaining candidate after a
// v_readlane broadcast.
// The candidate test is branch-free: the lanes whose leaf lies left of candidate
// position x (p < x, the lanes below x - first) are a prefix of the wavefront,
// taken as a mask rather than compared per lane; invalid lanes (p >= C) never
// lie left of a candidate. The scene bound (batches) and the suppression list
// are separate instantiations, chosen once per range.
template <bool kBatch, bool kSuppress>
__device__ __forceinline__ void scan_range_as(const Dev& d, uint32_t* diag, PairStage& st, uint32_t lane,
                                              const Packet& k, uint32_t start, uint32_t last) {
  for (uint32_t base = start; base <= last; base += 64u) {
    const uint32_t q = base + lane;
    float4 cmn = make_float4(0.f, 0.f, 0.f, 0.f), cmx = cmn;
    uint32_t cbody = 0;
    if (q <= last) {
      cmn = d.leaves[2 * (size_t)q];
      cmx = d.leaves[2 * (size_t)q + 1];
      cbody = d.leaf_body[q];
    }
    uint64_t todo = __ballot(q <= last && overlaps(k.umn, k.umx, cmn, cmx));
    while (todo != 0ull) {
      const uint32_t j = (uint32_t)__ffsll((unsigned long long)todo) - 1u;
      todo &= todo - 1ull;
      const float bx0 = bcast(cmn.x, j), by0 = bcast(cmn.y, j), bz0 = bcast(cmn.z, j);
      const float bx1 = bcast(cmx.x, j), by1 = bcast(cmx.y, j), bz1 = bcast(cmx.z, j);
      const uint32_t other_body = bcast(cbody, j);
      const uint32_t x = base + j;  // > k.first: ranges start right of the packet's first leaf
      const uint64_t lanes_left = lanes_below((uint32_t)__builtin_amdgcn_readfirstlane((int)(x - k.first)));
      bool hit = (k.qmn.x <= bx1) & (k.qmn.y <= by1) & (k.qmn.z <= bz1) & (bx0 <= k.qmx.x) & (by0 <= k.qmx.y) &
                 (bz0 <= k.qmx.z) & (other_body != k.self_body);
      // A merged range may run into the next scene: candidates past the end of
      // the query's scene are no pairs.
      if constexpr (kBatch) hit &= x < k.scene_end;
      uint64_t mask = __builtin_amdgcn_ballot_w64(hit) & lanes_left;
      if constexpr (kSuppress) {
        if (mask != 0ull) {
          hit = ((mask >> lane) & 1ull) && !is_partner(d, k.partners, other_body);
          mask = __builtin_amdgcn_ballot_w64(hit);
        }
      }
      if (mask != 0ull) {
        hit = (mask >> lane) & 1ull;
        const uint32_t other = bcast(__float_as_uint(cmn.w), j);
        const uint32_t cnt = (uint32_t)__popcll(mask);
        if (st.count + cnt > kStage) flush_pairs(d, diag, st, lane);
        if (hit) {
          const uint32_t rank = (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
          st.buf[st.count + rank] = make_uint2(k.self < other ? k.self : other, k.self < other ? other : k.self);
        }
        st.count += cnt;
      }
    }
  }
}

__device__ __forceinline__ void scan_range(const Dev& d, uint32_t* diag, PairStage& st, uint32_t lane,
                                           const Packet& k, uint32_t start, uint32_t last) {
  const bool batch = d.scene_count > 1, suppress = d.suppression_count != 0;
  if (batch) {
    if (suppress) scan_range_as<true, true>(d, diag, st, lane, k, start, last);
    else scan_range_as<true, false>(d, diag, st, lane, k, start, last);
  } else {
    if (suppress) scan_range_as<false, true>(d, diag, st, lane, k, start, last);
    else scan_range_as<false, false>(d, diag, st, lane, k, start, last);
  }
}

// scan_range with the suppression list chosen at compile time (k_scan is
// instantiated for scenes with and without one).
template <bool kSuppress>
__device__ __forceinline__ void scan_range_fixed(const Dev& d, uint32_t* diag, PairStage& st, uint32_t lane,
                                                 const Packet& k, uint32_t start, uint32_t last) {
  if (d.scene_count > 1) scan_range_as<true, kSuppress>(d, diag, st, lane, k, start, last);
  else scan_range_as<false, kSuppress>(d, diag, st, lane, k, start, last);
}

constexpr uint32_t kItemStage = 64;  // work items per packet

// Block index remapped so the workgroups of one XCD (dispatched round-robin,
// blockIdx % 8) take one contiguous share of the sorted leaves: neighbouring
// packets share tree nodes, which then stay in that XCD's L2.
__device__ __forceinline__ uint32_t xcd_block() {
  const uint32_t nb = gridDim.x, b = blockIdx.x;
  const uint32_t x = b % kXcds, q = nb / kXcds, r = nb % kXcds;
  return x * q + (x < r ? x : r) + b / kXcds;
}

// Writes the staged pairs of every wavefront of the workgroup with one atomic
// reservation for the whole workgroup (all threads must call it). Appends to one
// counter serialise across the device, so they are made as few as possible.
__device__ __forceinline__ void flush_pairs_block(const Dev& d, uint32_t* diag, PairStage& st, uint32_t lane,
                                                  uint32_t* lds) {
  const uint32_t wave = threadIdx.x >> 6, waves = blockDim.x >> 6;
  if (lane == 0) lds[wave] = st.count;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t total = 0;
    for (uint32_t w = 0; w < waves; ++w) {
      const uint32_t n = lds[w];
      lds[w] = total;
      total += n;
    }
    const uint32_t base = total ? atomicAdd(&diag[kDiagPairs], total) : 0u;
    if (total && base + total > d.pair_capacity) atomicOr(&diag[kDiagFlags], 1u);
    lds[waves] = base;
  }
  __syncthreads();
  const uint32_t base = lds[waves] + lds[wave];
  for (uint32_t i = lane; i < st.count; i += 64u) {
    const uint32_t row = base + i;
    if (row < d.pair_capacity) d.pairs[row] = st.buf[i];
  }
  st.count = 0;
}

// The packet walks the tree and collects the leaf ranges (buckets and single
// leaves) it must scan, in depth-first order. Ranges that continue the previous
// one are merged, so most of a packet's leaves are scanned 64 at a time. With
// kInline the wavefront scans each merged range itself; otherwise it records
// them, cut into items of at most 64 leaves, in the packet's kItemStage slots
// for k_scan to spread over many wavefronts (ranges past the slots are scanned
// here).
//
// One workgroup per packet: its wavefronts (walkers) descend together and split
// at the first nodes where both children must be entered, each half of the group
// taking one child, until every walker walks alone. A packet whose union box
// covers a large part of the tree thus shares its walk -- the longest walks set
// the kernel's time -- and each range is recorded by one walker only (the first
// of its group), so the packet's items are the same set as one walker's.
template <bool kInline>
__global__ __launch_bounds__(256) void k_traverse(Dev d) {
  __shared__ uint2 stage_mem[4][kStage];
  __shared__ uint32_t slots;  // the packet's item slots taken so far
  const uint64_t tick = tick_now(d);
  uint32_t* diag = d.diag[tick & 1];
  const uint32_t c = d.collider_count;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t packet = xcd_block();
  const uint32_t wave_first = packet * 64u;
  if (wave_first >= c) return;  // whole workgroup past the last leaf
  const uint32_t walker = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const uint32_t walkers = blockDim.x >> 6;
  if (threadIdx.x == 0) slots = 0u;
  __syncthreads();
  const Packet k = load_packet(d, wave_first, lane);
  PairStage st{stage_mem[walker], 0u};
  uint4* items = d.work + (size_t)packet * kItemStage;

  auto emit_range = [&](uint32_t first, uint32_t last) {
    if constexpr (kInline) {
      scan_range(d, diag, st, lane, k, first, last);
    } else {
      for (uint32_t a = first; a <= last; a += 64u) {
        const uint32_t b = last - a > 63u ? a + 63u : last;
        uint32_t slot = 0u;
        if (lane == 0) slot = atomicAdd(&slots, 1u);
        slot = bcast(slot, 0u);
        if (slot < kItemStage) {
          if (lane == 0) items[slot] = make_uint4(wave_first, a, b, 0u);
        } else {
          scan_range(d, diag, st, lane, k, a, b);
        }
      }
    }
  };

  uint32_t run_first = 1u, run_last = 0u;  // pending merged range (empty while first > last)
  auto visit_range = [&](uint32_t first, uint32_t last) {
    const uint32_t start = first > wave_first ? first : wave_first + 1u;
    if (start > last) return;
    if (run_first <= run_last && run_last + 1u == start) {
      run_last = last;
    } else {
      if (run_first <= run_last) emit_range(run_first, run_last);
      run_first = start;
      run_last = last;
    }
  };
  // A node or leaf with sorted leaves [first, last] is wanted when one of the
  // packet's queries overlaps its box and lies left of its last leaf (and, in a
  // batch, the node starts before the end of that query's scene).
  // Branch-free: the lanes left of `last` (p < last) are a prefix of the
  // packet, and invalid lanes (p >= C) are never left of a node.
  const bool batch = d.scene_count > 1;
  auto wanted = [&](float4 mn, float4 mx, uint32_t first, uint32_t last) {
    bool want = (k.qmn.x <= mx.x) & (k.qmn.y <= mx.y) & (k.qmn.z <= mx.z) & (mn.x <= k.qmx.x) &
                (mn.y <= k.qmx.y) & (mn.z <= k.qmx.z);
    if (batch) want &= first < k.scene_end;
    return (__builtin_amdgcn_ballot_w64(want) & lanes_below(last > wave_first ? last - wave_first : 0u)) != 0ull;
  };

  // Depth-first walk over the internal nodes the packet enters (frames: left
  // word and sorted leaf range). Each step loads both children of the frame's
  // node at once -- in Karras order they sit at the split and right after it --
  // and tests them, so a rejected child costs no round trip of its own. Wanted
  // leaves and buckets become leaf ranges. A walker alone keeps a wanted right
  // sibling on a wave-uniform stack held one entry per lane (left word kInvalid
  // marks a leaf range), so its ranges come out in depth-first order; a group
  // splits instead, and its first walker records a range beside a frame at once.
  uint32_t group_lo = 0u, group_n = walkers;  // walkers [group_lo, group_lo + group_n) walk this path
  uint32_t stack_left = 0, stack_first = 0, stack_last = 0;
  uint32_t depth = 0;
  uint32_t fr_left, fr_first = 0u, fr_last = c - 1u;
  bool walking = false;
  {
    const float4 mn = uniform4(d.inodes[0]), mx = uniform4(d.inodes[1]);
    if (wanted(mn, mx, 0u, c - 1u)) {
      fr_left = __float_as_uint(mn.w);
      if (!(fr_left & kBucketBit)) walking = true;
      else if (walker == 0u) visit_range(0u, c - 1u);
    }
  }
  uint32_t steps_left = c;
  while (walking) {
    if (steps_left-- == 0u) {
      if (lane == 0) atomicOr(&diag[kDiagFlags], 1u);
      break;
    }
    // Wave-uniform by construction (ballots, broadcasts, scalar records); said
    // so, the compiler keeps the walk on the scalar unit.
    fr_left = (uint32_t)__builtin_amdgcn_readfirstlane((int)fr_left);
    fr_first = (uint32_t)__builtin_amdgcn_readfirstlane((int)fr_first);
    fr_last = (uint32_t)__builtin_amdgcn_readfirstlane((int)fr_last);
    const uint32_t s = fr_left & kPosMask;  // split: left child s, right child s + 1
    const bool l_leaf = s == fr_first, r_leaf = s + 1u == fr_last;
    const float4* lrec = (l_leaf ? d.leaves : d.inodes) + 2 * (size_t)s;
    const float4* rrec = (r_leaf ? d.leaves : d.inodes) + 2 * (size_t)(s + 1u);
    // The records are the same for every lane: held in scalar registers, the
    // walk's bookkeeping below runs on the scalar unit.
    const float4 lmn = uniform4(lrec[0]), lmx = uniform4(lrec[1]), rmn = uniform4(rrec[0]), rmx = uniform4(rrec[1]);
    const bool wl = wanted(lmn, lmx, fr_first, s);
    const bool wr = wanted(rmn, rmx, s + 1u, fr_last);
    const uint32_t l_word = __float_as_uint(lmn.w), r_word = __float_as_uint(rmn.w);
    const bool l_range = l_leaf || (l_word & kBucketBit);
    const bool r_range = r_leaf || (r_word & kBucketBit);
    const uint32_t r_first = s + 1u, r_last = fr_last;
    const bool lead = walker == group_lo;
    bool next = false;
    if (group_n > 1u && wl && wr && !l_range && !r_range) {
      // Two frames: the lower half of the group takes the left, the upper the right.
      const uint32_t half = group_n >> 1;
      next = true;
      if (walker < group_lo + half) {
        fr_left = l_word;
        fr_last = s;
      } else {
        fr_left = r_word;
        fr_first = r_first;
        fr_last = r_last;
        group_lo += half;
      }
      group_n = half;
    } else {
      if (wl) {
        if (l_range) {
          if (lead) visit_range(fr_first, s);
        } else {
          next = true;
          fr_left = l_word;
          fr_last = s;
        }
      }
      if (wr) {
        if (next && group_n > 1u) {
          if (lead) visit_range(r_first, r_last);  // a range: two frames would have split
        } else if (next) {
          if (depth == 64u) {  // deeper than the key bits allow
            if (lane == 0) atomicOr(&diag[kDiagFlags], 1u);
            break;
          }
          if (lane == depth) {
            stack_left = r_range ? kInvalid : r_word;
            stack_first = r_first;
            stack_last = r_last;
          }
          ++depth;
        } else if (r_range) {
          if (lead) visit_range(r_first, r_last);
        } else {
          next = true;
          fr_left = r_word;
          fr_first = r_first;
          fr_last = r_last;
        }
      }
    }
    // Only a walker alone pushes, so it alone pops.
    while (!next && depth != 0u) {
      --depth;
      const uint32_t left = bcast(stack_left, depth);
      const uint32_t first = bcast(stack_first, depth), last = bcast(stack_last, depth);
      if (left == kInvalid) {
        visit_range(first, last);
      } else {
        next = true;
        fr_left = left;
        fr_first = first;
        fr_last = last;
      }
    }
    walking = next;
  }
  if (run_first <= run_last) emit_range(run_first, run_last);
  if (st.count) flush_pairs(d, diag, st, lane);
  if constexpr (!kInline) {
    // Publish the packet's item count with its published bit as one word (the
    // Morton kernel zeroes the words each tick); the last workgroup to arrive
    // forms the exclusive prefix of all the counts, which the scan distributes
    // by. Counts and arrivals are relaxed device atomics (a release fence per
    // workgroup would write the XCD's L2 back each time): the last workgroup
    // reads each word until it is published, as every workgroup published its
    // own before arriving.
    __shared__ uint32_t last;
    __syncthreads();
    const uint32_t tag = kWalkPublished;
    if (threadIdx.x == 0) {
      const uint32_t count = slots < kItemStage ? slots : kItemStage;
      store_status(&d.work_count[packet], tag | count);
      const uint32_t arrived =
          __hip_atomic_fetch_add(&d.counters[kCountWalked], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      last = arrived == gridDim.x - 1u ? 1u : 0u;
    }
    __syncthreads();
    if (last) {
      __shared__ uint32_t wave_sums[kBlock / 64];
      const uint32_t packets = gridDim.x;
      const uint32_t per = (packets + kBlock - 1u) / kBlock;
      const uint32_t begin = threadIdx.x * per < packets ? threadIdx.x * per : packets;
      const uint32_t end = begin + per < packets ? begin + per : packets;
      uint32_t sum = 0;
      for (uint32_t i = begin; i < end; ++i) {
        uint32_t v = load_status(&d.work_count[i]);
        for (uint32_t spins = 0; (v & ~kWalkCountMask) != tag; v = load_status(&d.work_count[i])) {
          if (++spins >= kRsSpinLimit) {
            atomicOr(&diag[kDiagFlags], 1u);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
        sum += v & kWalkCountMask;
      }
      uint32_t run = exclusive_scan256(sum, wave_sums);
      for (uint32_t i = begin; i < end; ++i) {
        d.work_first[i] = run;
        run += load_status(&d.work_count[i]) & kWalkCountMask;
      }
      if (threadIdx.x == kBlock - 1u) d.work_first[packets] = run;
    }
  }
}

// Every wavefront takes a contiguous share of the work items (in packet order,
// through the prefix of the per-packet counts) and scans each item's leaves
// against its packet.
template <bool kSuppress>
__global__ __launch_bounds__(256) void k_scan(Dev d) {
  __shared__ uint2 stage_mem[4][kStage];
  __shared__ uint32_t flush_lds[5];
  const uint64_t tick = tick_now(d);
  uint32_t* diag = d.diag[tick & 1];
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t packets = (d.collider_count + 63u) / 64u;
  const uint32_t total = d.work_first[packets];
  const uint32_t waves = gridDim.x * (blockDim.x >> 6);
  const uint32_t wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const uint32_t share = (total + waves - 1u) / waves;
  const uint32_t begin = wave * share;
  const uint32_t end = begin + share < total ? begin + share : total;
  PairStage st{stage_mem[threadIdx.x >> 6], 0u};
  if (begin < end) {
    // Packet holding item `begin`: work_first[lo] <= begin < work_first[lo + 1].
    uint32_t lo = 0, hi = packets;
    while (hi - lo > 1u) {
      const uint32_t mid = (lo + hi) / 2u;
      if (d.work_first[mid] <= begin) lo = mid;
      else hi = mid;
    }
    uint32_t pk = lo, pk_first = d.work_first[pk], pk_next = d.work_first[pk + 1];
    Packet k = load_packet<kSuppress>(d, pk * 64u, lane);
    for (uint32_t i = begin; i < end; ++i) {
      if (i >= pk_next) {
        do {
          ++pk;
          pk_first = pk_next;
          pk_next = d.work_first[pk + 1];
        } while (i >= pk_next);
        k = load_packet<kSuppress>(d, pk * 64u, lane);
      }
      const uint4 it = d.work[(size_t)pk * kItemStage + (i - pk_first)];
      scan_range_fixed<kSuppress>(d, diag, st, lane, k, it.y, it.z);
    }
  }
  flush_pairs_block(d, diag, st, lane, flush_lds);
}

// One walk per leaf (the reference's per-leaf traversal, lbvh.wgsl:321-398,
// without its stack): lane p follows escapes from the root, descending into a
// node when its own query overlaps it and the node holds leaves right of p, and
// tests the leaves it reaches (a bucket's leaves one per step). Each step is one
// node or leaf per lane; hits are compacted by ballot into the wavefront's LDS
// stage as in the packet walk.
__global__ __launch_bounds__(256) void k_traverse_lane(Dev d) {
  __shared__ uint2 stage_mem[4][kStage];
  const uint64_t tick = tick_now(d);
  uint32_t* diag = d.diag[tick & 1];
  const uint32_t c = d.collider_count;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t p = xcd_block() * blockDim.x + threadIdx.x;
  if ((p & ~63u) >= c) return;  // whole wavefront past the last leaf
  PairStage st{stage_mem[threadIdx.x >> 6], 0u};
  const bool check_suppression = d.suppression_count != 0;
  bool active = p < c;
  float4 qmn = make_float4(0.f, 0.f, 0.f, 0.f), qmx = qmn;
  uint32_t self = 0, self_body = 0, scene_end = c;
  if (active) {
    qmn = d.leaves[2 * (size_t)p];
    qmx = d.leaves[2 * (size_t)p + 1];
    self = __float_as_uint(qmn.w);
    self_body = d.leaf_body[p];
    if (d.scene_count > 1) scene_end = d.scene_first[d.cscene[self] + 1];
  }
  uint32_t cur = 0u;                   // the root (c >= 2)
  uint32_t scan = 0u, scan_last = 0u;  // bucket leaves still to test, when scan <= scan_last
  uint32_t after = 0u;                 // escape of the bucket being scanned
  bool scanning = false;
  uint32_t steps_left = 4u * c + 4u;
  while (__ballot(active) != 0ull) {
    bool hit = false;
    uint32_t other = 0u;
    if (active) {
      uint32_t leaf = kInvalid;
      if (scanning) {
        leaf = scan++;
        if (scan > scan_last) {
          scanning = false;
          cur = after;
        }
      } else if (cur & kLeafBit) {
        leaf = cur & kPosMask;
        cur = __float_as_uint(d.leaves[2 * (size_t)leaf + 1].w);
      } else {
        const uint32_t pos = cur & kPosMask;
        const float4 mn = d.inodes[2 * (size_t)pos];
        const float4 mx = d.inodes[2 * (size_t)pos + 1];
        const uint32_t left_word = __float_as_uint(mn.w);
        const uint32_t esc = __float_as_uint(mx.w);
        const uint32_t range_last = (esc & kPosMask) - 1u;
        bool want = range_last > p && overlaps(qmn, qmx, mn, mx);
        if (d.scene_count > 1 && want) want = d.ranges[pos].x < scene_end;
        if (!want) {
          cur = esc;
        } else if (!(left_word & kBucketBit)) {
          cur = left_word;
        } else {
          const uint32_t first = left_word & (kBucketBit - 1u);
          scan = first > p ? first : p + 1u;
          scan_last = range_last;
          after = esc;
          scanning = scan <= scan_last;
          if (!scanning) cur = esc;
        }
      }
      if (leaf != kInvalid && leaf > p && leaf < scene_end) {
        const float4 bmn = d.leaves[2 * (size_t)leaf];
        const float4 bmx = d.leaves[2 * (size_t)leaf + 1];
        const uint32_t other_body = d.leaf_body[leaf];
        other = __float_as_uint(bmn.w);
        hit = overlaps(qmn, qmx, bmn, bmx) && other_body != self_body;
        if (hit && check_suppression) hit = !suppressed(d, self_body, other_body);
      }
      if (!scanning && cur == c) active = false;
      if (active && steps_left-- == 0u) {
        atomicOr(&diag[kDiagFlags], 1u);
        active = false;
      }
    }
    const uint64_t mask = __ballot(hit);
    if (mask != 0ull) {
      const uint32_t cnt = (uint32_t)__popcll(mask);
      if (st.count + cnt > kStage) flush_pairs(d, diag, st, lane);
      if (hit) {
        const uint32_t rank = (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
        st.buf[st.count + rank] = make_uint2(self < other ? self : other, self < other ? other : self);
      }
      st.count += cnt;
    }
  }
  if (st.count) flush_pairs(d, diag, st, lane);
}

// ---------------------------------------------------------------------------
static inline uint32_t blocks_for(uint32_t n) { return (n + kBlock - 1) / kBlock; }

void launch_morton(hipStream_t s, const Dev& d) {
  k_morton<<<blocks_for(d.sort_count), kBlock, 0, s>>>(d);
}

void launch_sort(hipStream_t s, const Dev& d) {
  if (d.t_sort_kind == 0u) {
    // k_morton wrote the keys where the last pass lands in d.keys.
    if (d.rs_blocks == 0) return;
    uint64_t* buf[2] = {d.keys, d.keys_alt};
    const uint32_t odd = d.rs_passes & 1u;
    for (uint32_t pass = 0; pass < d.rs_passes; ++pass) {
      const uint32_t from = (pass + odd) & 1u;
      k_radix_pass<<<d.rs_blocks, 256, 0, s>>>(d, buf[from], buf[from ^ 1u], pass);
    }
    return;
  }
  const uint32_t n = d.sort_count;  // power of two, >= 256
  const uint32_t tile_cap = d.t_sort_tile ? d.t_sort_tile : kTile;
  const uint32_t tile = n < tile_cap ? n : tile_cap;
  const uint32_t max_levels = d.t_sort_levels ? d.t_sort_levels : 4u;
  k_sort_tile<<<n / tile, tile / kPer, 0, s>>>(d.keys, tile);
  for (uint32_t k = tile * 2; k <= n; k <<= 1) {
    uint32_t j = k >> 1;
    while (j >= tile) {
      uint32_t span = 31u - (uint32_t)__builtin_clz(j) - (31u - (uint32_t)__builtin_clz(tile)) + 1u;
      uint32_t levels = span < max_levels ? span : max_levels;
      uint32_t threads = n >> levels;
      if (levels == 4) k_sort_global<4><<<blocks_for(threads), kBlock, 0, s>>>(d.keys, n, k, j);
      else if (levels == 3) k_sort_global<3><<<blocks_for(threads), kBlock, 0, s>>>(d.keys, n, k, j);
      else if (levels == 2) k_sort_global<2><<<blocks_for(threads), kBlock, 0, s>>>(d.keys, n, k, j);
      else k_sort_global<1><<<blocks_for(threads), kBlock, 0, s>>>(d.keys, n, k, j);
      j >>= levels;
    }
    k_sort_merge<<<n / tile, tile / kPer, 0, s>>>(d.keys, tile, k);
  }
}

void launch_lbvh(hipStream_t s, const Dev& d) {
  if (d.collider_count == 0) return;
  k_build<<<blocks_for(d.collider_count), kBlock, 0, s>>>(d);
  if (d.collider_count > 1) k_refit_top<<<1024, kBlock, 0, s>>>(d);
}

void launch_traverse(hipStream_t s, const Dev& d) {
  if (d.collider_count < 2) return;
  const uint32_t packets = (d.collider_count + 63u) / 64u;  // one workgroup of walkers each
  if (d.t_traverse_kind == 1u) {
    k_traverse_lane<<<blocks_for(d.collider_count), kBlock, 0, s>>>(d);
  } else if (d.t_traverse_kind == 2u) {
    k_traverse<true><<<packets, kBlock, 0, s>>>(d);
  } else {
    k_traverse<false><<<packets, kBlock, 0, s>>>(d);
    const uint32_t scan_blocks = d.t_scan_blocks ? d.t_scan_blocks : 4096u;
    if (d.suppression_count) k_scan<true><<<scan_blocks, kBlock, 0, s>>>(d);
    else k_scan<false><<<scan_blocks, kBlock, 0, s>>>(d);
  }
}

}  // namespace mh
