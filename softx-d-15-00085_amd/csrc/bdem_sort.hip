/*
 * bdem_sort.hip - a2: the stable sort of particle slots by cell key
 * (replaces thrust::stable_sort_by_key, Collision_Detection_BroadPhase.cuh:110-117).
 *
 * Result: keys ascending, ties in ascending slot order -- exactly what a
 * stable sort of (key, slot) produces, so sorted order is bit-exact.
 *
 * Two paths produce that same result:
 *
 * 1. Incremental (every step after the first).  The slots are in the previous
 *    step's sorted order and the step kernel flags each particle whose key
 *    changed.  The unchanged slots already form a sorted subsequence, so:
 *      k_partition   stable split into unchanged (R) and changed (S) slots,
 *                    offsets from the per-256-slot counts the step kernel wrote
 *      k_sort_small  one workgroup bitonic-sorts S as (key<<32 | slot) in LDS
 *      k_merge       every element finds its output position by rank in its
 *                    own run plus a binary search in the other run
 *    Traffic is ~29 bytes per particle instead of ~60 for three radix passes,
 *    in 3 launches instead of 9.
 * 2. Full LSD radix sort (8-bit digits, wave64 ballot ranking) for the first
 *    step after the state was set, and as the fallback when S does not fit in
 *    one workgroup's LDS.  The fallback kernels are always launched after the
 *    incremental path and return at once unless k_sort_small raised the
 *    fallback flag, so no host round trip is needed.
 */
#include "bdem_kernels.hpp"

namespace bdem {

int key_bits(uint32_t max_key) {
  int bits = 1;
  while (bits < 32 && (max_key >> bits) != 0) ++bits;
  return bits;
}

int sort_passes(uint32_t max_key) { return (key_bits(max_key) + 7) / 8; }

__device__ __forceinline__ bool gated_off(const uint32_t* gate) {
  return gate != nullptr && gate[1] == 0u;
}

/* ------------------------------------------------------------------------ */
/* Full radix sort                                                           */
/* ------------------------------------------------------------------------ */
/* Digit histogram of one 4096-key tile: hist[digit * nblocks + block]. */
__global__ __launch_bounds__(SORT_THREADS) void k_radix_hist(const uint32_t* __restrict__ keys, int n,
                                                             int shift, uint32_t* __restrict__ hist,
                                                             int nblocks, const uint32_t* gate) {
  if (gated_off(gate)) return;
  __shared__ uint32_t h[RADIX];
  for (int d = threadIdx.x; d < RADIX; d += SORT_THREADS) h[d] = 0;
  __syncthreads();
  const int base = blockIdx.x * SORT_TILE;
  for (int k = threadIdx.x; k < SORT_TILE; k += SORT_THREADS) {
    const int i = base + k;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & (RADIX - 1)], 1u);
  }
  __syncthreads();
  for (int d = threadIdx.x; d < RADIX; d += SORT_THREADS) hist[d * nblocks + blockIdx.x] = h[d];
}

/* Exclusive scan of m counters in one 1024-thread block (coalesced chunks). */
__global__ __launch_bounds__(1024) void k_scan_exclusive(uint32_t* __restrict__ data, int m,
                                                         const uint32_t* gate) {
  if (gated_off(gate)) return;
  __shared__ uint32_t sums[1024];
  __shared__ uint32_t carry;
  const int t = threadIdx.x;
  if (t == 0) carry = 0u;
  __syncthreads();
  for (int base = 0; base < m; base += 1024) {
    const int i = base + t;
    const uint32_t x = i < m ? data[i] : 0u;
    sums[t] = x;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      const uint32_t v = t >= off ? sums[t - off] : 0u;
      __syncthreads();
      sums[t] += v;
      __syncthreads();
    }
    const uint32_t excl = carry + sums[t] - x;
    if (i < m) data[i] = excl;
    __syncthreads();
    if (t == 1023) carry += sums[1023];
    __syncthreads();
  }
}

/* Scatter one tile.  Wave w owns keys [w*1024, (w+1)*1024) of the tile and
 * walks them 64 at a time in order; lanes with equal digits are matched with
 * eight ballots, ranked by lane, and the group's leader bumps the wave's LDS
 * counter.  Ranks are therefore stable across lanes, iterations, waves and
 * (through the digit-major scanned histogram) blocks. */
__global__ __launch_bounds__(SORT_THREADS) void k_radix_scatter(
    const uint32_t* __restrict__ kin, const int* __restrict__ vin, uint32_t* __restrict__ kout,
    int* __restrict__ vout, int n, int shift, const uint32_t* __restrict__ hist_scan, int nblocks,
    const uint32_t* gate) {
  if (gated_off(gate)) return;
  __shared__ uint32_t wcnt[SORT_THREADS / 64][RADIX];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  for (int d = threadIdx.x; d < RADIX; d += SORT_THREADS) {
#pragma unroll
    for (int w = 0; w < SORT_THREADS / 64; ++w) wcnt[w][d] = 0u;
  }
  __syncthreads();
  const int base = blockIdx.x * SORT_TILE + wave * (SORT_ITEMS * 64);
  const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  uint32_t key[SORT_ITEMS];
  int val[SORT_ITEMS];
  uint32_t loc[SORT_ITEMS];
#pragma unroll
  for (int it = 0; it < SORT_ITEMS; ++it) {
    const int i = base + it * 64 + lane;
    const bool valid = i < n;
    const uint32_t k = valid ? kin[i] : 0u;
    const int v = valid ? (vin ? vin[i] : i) : 0;
    const uint32_t d = (k >> shift) & (RADIX - 1);
    uint64_t eq = __ballot(valid);
#pragma unroll
    for (int bit = 0; bit < 8; ++bit) {
      const bool set = (d >> bit) & 1u;
      const uint64_t m = __ballot(set);
      eq &= set ? m : ~m;
    }
    const uint32_t rank = (uint32_t)__popcll(eq & lt);
    const uint32_t before = wcnt[wave][d];
    __builtin_amdgcn_wave_barrier();
    if (valid && (eq & lt) == 0ull) wcnt[wave][d] = before + (uint32_t)__popcll(eq);
    __builtin_amdgcn_wave_barrier();
    key[it] = k;
    val[it] = v;
    loc[it] = before + rank;
  }
  __syncthreads();
  for (int d = threadIdx.x; d < RADIX; d += SORT_THREADS) {
    uint32_t s = hist_scan[d * nblocks + blockIdx.x];
#pragma unroll
    for (int w = 0; w < SORT_THREADS / 64; ++w) {
      const uint32_t c = wcnt[w][d];
      wcnt[w][d] = s;
      s += c;
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < SORT_ITEMS; ++it) {
    const int i = base + it * 64 + lane;
    if (i < n) {
      const uint32_t d = (key[it] >> shift) & (RADIX - 1);
      const uint32_t pos = wcnt[wave][d] + loc[it];
      kout[pos] = key[it];
      vout[pos] = val[it];
    }
  }
}

static int launch_radix(const StepBuffers& b, int bits, const uint32_t* gate, hipStream_t s) {
  const int nb = sort_blocks(b.n);
  int cur = 0;
  for (int shift = 0; shift < bits; shift += 8) {
    hipLaunchKernelGGL(k_radix_hist, dim3(nb), dim3(SORT_THREADS), 0, s, b.keys[cur], b.n, shift,
                       b.radix_hist, nb, gate);
    hipLaunchKernelGGL(k_scan_exclusive, dim3(1), dim3(1024), 0, s, b.radix_hist, RADIX * nb, gate);
    hipLaunchKernelGGL(k_radix_scatter, dim3(nb), dim3(SORT_THREADS), 0, s, b.keys[cur],
                       shift == 0 ? (const int*)nullptr : b.vals[cur], b.keys[cur ^ 1],
                       b.vals[cur ^ 1], b.n, shift, b.radix_hist, nb, gate);
    cur ^= 1;
  }
  return cur;
}

/* ------------------------------------------------------------------------ */
/* Incremental path                                                          */
/* ------------------------------------------------------------------------ */
constexpr int PART_TILE = SORT_TILE;                 /* 4096 slots, 16 step blocks */
constexpr int STEP_BLOCK = 256;                      /* k_step block size */
constexpr int SMALL_MAX = 16384;                     /* S capacity of one workgroup (128 KiB LDS) */

/* Stable split of slots into unchanged (R, written to rkeys/rvals in order)
 * and changed (S, written to sbuf as key<<32|slot).  The S offset of a tile
 * is the sum of the step kernel's per-block changed counts before it. */
__global__ __launch_bounds__(SORT_THREADS) void k_partition(
    int n, const uint32_t* __restrict__ keys, const uint8_t* __restrict__ changed,
    const uint32_t* __restrict__ blk_changed, uint32_t* __restrict__ rkeys, int* __restrict__ rvals,
    unsigned long long* __restrict__ sbuf, uint32_t* __restrict__ meta) {
  __shared__ uint32_t red[SORT_THREADS / 64];
  __shared__ uint32_t wave_s[SORT_THREADS / 64];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int tile_base = blockIdx.x * PART_TILE;
  /* S slots before this tile */
  const int nprev = tile_base / STEP_BLOCK;
  uint32_t acc = 0u;
  for (int b = threadIdx.x; b < nprev; b += SORT_THREADS) acc += blk_changed[b];
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if (lane == 0) red[wave] = acc;
  /* per-wave S counts of this tile */
  const int wbase = tile_base + wave * (SORT_ITEMS * 64);
  uint32_t cnt = 0u;
#pragma unroll
  for (int it = 0; it < SORT_ITEMS; ++it) {
    const int i = wbase + it * 64 + lane;
    cnt += (uint32_t)__popcll(__ballot(i < n && changed[i] != 0));
  }
  if (lane == 0) wave_s[wave] = cnt;
  __syncthreads();
  uint32_t s_off = 0u;
  for (int w = 0; w < SORT_THREADS / 64; ++w) s_off += red[w];
  for (int w = 0; w < wave; ++w) s_off += wave_s[w];
  const uint32_t tile_s_before = s_off;  /* includes earlier waves of this tile */
  uint32_t r_off = (uint32_t)wbase - tile_s_before;
  const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
#pragma unroll
  for (int it = 0; it < SORT_ITEMS; ++it) {
    const int i = wbase + it * 64 + lane;
    const bool valid = i < n;
    const bool chg = valid && changed[i] != 0;
    const uint64_t sm = __ballot(chg);
    const uint64_t rm = __ballot(valid && !chg);
    if (valid) {
      const uint32_t k = keys[i];
      if (chg) {
        sbuf[s_off + (uint32_t)__popcll(sm & lt)] = ((unsigned long long)k << 32) | (uint32_t)i;
      } else {
        const uint32_t r = r_off + (uint32_t)__popcll(rm & lt);
        rkeys[r] = k;
        rvals[r] = i;
      }
    }
    s_off += (uint32_t)__popcll(sm);
    r_off += (uint32_t)__popcll(rm);
  }
  /* the tile holding the last slot publishes |S| */
  if (tile_base <= n - 1 && n - 1 < tile_base + PART_TILE && wave == SORT_THREADS / 64 - 1 && lane == 0) {
    uint32_t total = 0u;
    for (int w = 0; w < SORT_THREADS / 64; ++w) total += red[w] + wave_s[w];
    meta[0] = total;
  }
}

/* Sort S (|S| <= SMALL_MAX) in one workgroup: bitonic sort of 64-bit
 * (key<<32 | slot) composites, padded with all-ones.  Raises the fallback
 * flag (meta[1]) when S is too large. */
__global__ __launch_bounds__(1024) void k_sort_small(unsigned long long* __restrict__ sbuf,
                                                     uint32_t* __restrict__ meta) {
  __shared__ unsigned long long v[SMALL_MAX];
  const uint32_t ns = meta[0];
  if (ns > (uint32_t)SMALL_MAX) {
    if (threadIdx.x == 0) meta[1] = 1u;
    return;
  }
  if (threadIdx.x == 0) meta[1] = 0u;
  if (ns <= 1u) return;
  uint32_t m = 1u;
  while (m < ns) m <<= 1;
  for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) v[i] = i < ns ? sbuf[i] : ~0ull;
  __syncthreads();
  for (uint32_t k = 2; k <= m; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) {
        const uint32_t l = i ^ j;
        if (l > i) {
          const unsigned long long a = v[i], b = v[l];
          const bool up = (i & k) == 0;
          if ((a > b) == up) { v[i] = b; v[l] = a; }
        }
      }
      __syncthreads();
    }
  }
  for (uint32_t i = threadIdx.x; i < ns; i += blockDim.x) sbuf[i] = v[i];
}

/* Merge R (in rkeys/rvals) and sorted S (in sbuf) into the output arrays.
 * Thread t < |R| places R[t] at t + #{S < R[t]}; thread t >= |R| places
 * S[q = t-|R|] at q + #{R < S[q]}.  R and S are disjoint in slot, so the
 * (key, slot) order is strict.  R threads of one block are consecutive and
 * sorted, so their search runs inside the S window of the block's first and
 * last element. */
__global__ __launch_bounds__(256) void k_merge(int n, const uint32_t* __restrict__ rkeys,
                                               const int* __restrict__ rvals,
                                               const unsigned long long* __restrict__ sbuf,
                                               const uint32_t* __restrict__ meta,
                                               uint32_t* __restrict__ kout, int* __restrict__ vout) {
  if (meta[1] != 0u) return;
  const uint32_t ns = meta[0];
  const int nr = n - (int)ns;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  __shared__ uint32_t win[2];
  const int b0 = blockIdx.x * blockDim.x;
  if (threadIdx.x == 0) {
    /* S window of this block's R range */
    uint32_t lo = 0u, hi = ns;
    if (b0 < nr) {
      const int last = min(b0 + (int)blockDim.x, nr) - 1;
      const unsigned long long first_key = ((unsigned long long)rkeys[b0] << 32) | (uint32_t)rvals[b0];
      const unsigned long long last_key = ((unsigned long long)rkeys[last] << 32) | (uint32_t)rvals[last];
      uint32_t a = 0u, z = ns;
      while (a < z) { const uint32_t mid = (a + z) >> 1; if (sbuf[mid] < first_key) a = mid + 1; else z = mid; }
      lo = a;
      z = ns;
      while (a < z) { const uint32_t mid = (a + z) >> 1; if (sbuf[mid] < last_key) a = mid + 1; else z = mid; }
      hi = a;
    }
    win[0] = lo;
    win[1] = hi;
  }
  __syncthreads();
  if (t >= n) return;
  if (t < nr) {
    const uint32_t k = rkeys[t];
    const int s = rvals[t];
    const unsigned long long c = ((unsigned long long)k << 32) | (uint32_t)s;
    uint32_t a = win[0], z = win[1];
    while (a < z) { const uint32_t mid = (a + z) >> 1; if (sbuf[mid] < c) a = mid + 1; else z = mid; }
    const int pos = t + (int)a;
    kout[pos] = k;
    vout[pos] = s;
  } else {
    const uint32_t q = (uint32_t)(t - nr);
    const unsigned long long c = sbuf[q];
    int a = 0, z = nr;
    while (a < z) {
      const int mid = (a + z) >> 1;
      const unsigned long long rm = ((unsigned long long)rkeys[mid] << 32) | (uint32_t)rvals[mid];
      if (rm < c) a = mid + 1; else z = mid;
    }
    const int pos = (int)q + a;
    kout[pos] = (uint32_t)(c >> 32);
    vout[pos] = (int)(uint32_t)(c & 0xFFFFFFFFull);
  }
}

int launch_sort(const StepBuffers& b, uint32_t max_key, bool incremental, hipStream_t s) {
  const int bits = key_bits(max_key);
  const int final_buf = sort_passes(max_key) & 1;
  if (!incremental) {
    const int cur = launch_radix(b, bits, nullptr, s);
    return cur;
  }
  const int np = (b.n + PART_TILE - 1) / PART_TILE;
  hipLaunchKernelGGL(k_partition, dim3(np), dim3(SORT_THREADS), 0, s, b.n, b.keys[0], b.changed,
                     b.blk_changed, b.rkeys, b.rvals, b.sbuf, b.sort_meta);
  hipLaunchKernelGGL(k_sort_small, dim3(1), dim3(1024), 0, s, b.sbuf, b.sort_meta);
  hipLaunchKernelGGL(k_merge, dim3((b.n + 255) / 256), dim3(256), 0, s, b.n, b.rkeys, b.rvals,
                     b.sbuf, b.sort_meta, b.keys[final_buf], b.vals[final_buf]);
  const int cur = launch_radix(b, bits, b.sort_meta, s);
  (void)cur;
  return final_buf;
}

}  // namespace bdem
