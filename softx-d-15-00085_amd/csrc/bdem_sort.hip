/*
 * bdem_sort.hip - a2-a4: stable sort of particle slots by cell key, the
 * sorted gather of the state and the cell (start, end) table.
 * Replaces thrust::stable_sort_by_key + Kernel_SortArrays +
 * Kernel_SpatialDecomp_ReorderPDA + the 0xFF fill + Kernel_SpatialDecomp_BinData
 * (Collision_Detection_BroadPhase.cuh:110-325, Device/DeviceInterface.cu:2097-2105).
 *
 * Result: keys ascending, ties in ascending slot order -- exactly what a
 * stable sort of (key, slot) produces -- so the sorted order is bit-exact.
 *
 * Incremental path (every step after the first).  The slots are in the
 * previous step's sorted order; the step kernel appended every particle
 * whose key changed to a small list S.  Unchanged particles already form a
 * sorted subsequence R, so with S sorted both by (key, slot) and by slot:
 *   unchanged slot i with key k:  pos = i - #{S: slot < i} + #{S: (key,slot) < (k,i)}
 *   changed  slot i with key k:   pos = rank of (k,i) in S + #{R: (key,slot) < (k,i)},
 *     where #{R < (k,i)} = #{all: (oldkey,slot) < (k,i)} - #{S: (oldkey,slot) < (k,i)}
 *     (old keys are the previous sorted keys; unchanged elements keep theirs).
 *   k_sort_small   one workgroup bitonic-sorts S by (key<<32|slot) and by slot in
 *                  LDS, and clears the cells the changed particles left
 *   k_place        every slot computes its position (binary searches within
 *                  the block's window of S), writes the sorted key and slot,
 *                  and gathers the three state streams
 *   k_bin          writes (start, end) of every occupied cell
 * Three launches, ~110 bytes per particle, no full-table clear.
 *
 * Full path: one cooperative kernel does the LSD radix sort (8-bit digits,
 * wave64 ballot ranking), the gather, the table clear and the binning.  It
 * runs on the first step after a host upload and, gated on a device flag, as
 * the fallback when S does not fit one workgroup; otherwise it returns at
 * once, so the fallback costs one empty launch and no host round trip.
 */
#include <hip/hip_cooperative_groups.h>
#include "bdem_kernels.hpp"

namespace cg = cooperative_groups;

namespace bdem {

int key_bits(uint32_t max_key) {
  int bits = 1;
  while (bits < 32 && (max_key >> bits) != 0) ++bits;
  return bits;
}

int sort_passes(uint32_t max_key) { return (key_bits(max_key) + 7) / 8; }

/* ------------------------------------------------------------------------ */
/* Incremental path                                                          */
/* ------------------------------------------------------------------------ */
__global__ __launch_bounds__(1024) void k_sort_small(SortArgs a) {
  __shared__ unsigned long long v[SMALL_MAX];
  const uint32_t ns = a.meta[META_APPEND];
  __syncthreads();
  if (threadIdx.x == 0) {
    a.meta[META_APPEND] = 0u;          /* the next step appends afresh */
    a.meta[META_COUNT] = ns;
    a.meta[META_FALLBACK] = ns > (uint32_t)SMALL_MAX ? 1u : 0u;
  }
  if (ns > (uint32_t)SMALL_MAX || ns == 0u) return;
  uint32_t m = 1u;
  while (m < ns) m <<= 1;
  /* pass 1: by (key, slot) */
  for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) v[i] = i < ns ? a.sbuf[i] : ~0ull;
  __syncthreads();
  for (int pass = 0; pass < 2; ++pass) {
    for (uint32_t k = 2; k <= m; k <<= 1) {
      for (uint32_t j = k >> 1; j > 0; j >>= 1) {
        for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) {
          const uint32_t l = i ^ j;
          if (l > i) {
            const unsigned long long x = v[i], y = v[l];
            if ((x > y) == ((i & k) == 0)) { v[i] = y; v[l] = x; }
          }
        }
        __syncthreads();
      }
    }
    if (pass == 0) {
      for (uint32_t i = threadIdx.x; i < ns; i += blockDim.x) a.skey[i] = v[i];
      __syncthreads();
      /* pass 2: by slot (as 64-bit values slot<<32) */
      for (uint32_t i = threadIdx.x; i < m; i += blockDim.x)
        v[i] = i < ns ? ((a.sbuf[i] & 0xFFFFFFFFull) << 32) : ~0ull;
      __syncthreads();
    } else {
      for (uint32_t i = threadIdx.x; i < ns; i += blockDim.x) a.sslot[i] = (uint32_t)(v[i] >> 32);
    }
  }
  /* clear the cells the changed particles left; k_bin rewrites every occupied cell */
  for (uint32_t i = threadIdx.x; i < ns; i += blockDim.x) {
    const uint32_t slot = (uint32_t)(a.sbuf[i] & 0xFFFFFFFFull);
    const uint32_t ok = a.okeys[slot];
    if (ok < a.C) a.cells[ok] = make_uint2(0xFFFFFFFFu, 0xFFFFFFFFu);
  }
}

__device__ __forceinline__ uint32_t lb_u64(const unsigned long long* __restrict__ arr, uint32_t lo,
                                           uint32_t hi, unsigned long long x) {
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (arr[mid] < x) lo = mid + 1; else hi = mid;
  }
  return lo;
}
__device__ __forceinline__ uint32_t lb_u32(const uint32_t* __restrict__ arr, uint32_t lo, uint32_t hi,
                                           uint32_t x) {
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (arr[mid] < x) lo = mid + 1; else hi = mid;
  }
  return lo;
}
__device__ __forceinline__ unsigned long long comp(uint32_t k, uint32_t s) {
  return ((unsigned long long)k << 32) | s;
}

/* Position of every slot in the new order, then the gather.  The S windows
 * of a block are found once: slots [b0, b0+256) bound the slot searches,
 * and the block's unchanged elements (sorted by (key, slot)) bound the key
 * searches between its first and last unchanged element. */
__global__ __launch_bounds__(256) void k_place(SortArgs a) {
  if (a.meta[META_FALLBACK] != 0u) return;
  const uint32_t ns = a.meta[META_COUNT];
  const int n = a.n;
  const int b0 = blockIdx.x * blockDim.x;
  const int i = b0 + threadIdx.x;
  __shared__ uint32_t win[4];
  __shared__ unsigned long long ends[2];
  const bool in = i < n;
  const uint32_t k = in ? a.keys_new[i] : 0u;
  const bool chg = in && a.changed[i] != 0;
  if (threadIdx.x == 0) { ends[0] = ~0ull; ends[1] = 0ull; }
  __syncthreads();
  if (in && !chg) {
    atomicMin(&ends[0], comp(k, (uint32_t)i));
    atomicMax(&ends[1], comp(k, (uint32_t)i));
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int b1 = min(b0 + (int)blockDim.x, n);
    win[0] = lb_u32(a.sslot, 0u, ns, (uint32_t)b0);
    win[1] = lb_u32(a.sslot, win[0], ns, (uint32_t)b1);
    win[2] = lb_u64(a.skey, 0u, ns, ends[0]);
    win[3] = ends[1] >= ends[0] ? lb_u64(a.skey, win[2], ns, ends[1]) : win[2];
  }
  __syncthreads();
  if (!in) return;
  uint32_t pos;
  const unsigned long long c = comp(k, (uint32_t)i);
  if (!chg) {
    pos = (uint32_t)i - lb_u32(a.sslot, win[0], win[1], (uint32_t)i) + lb_u64(a.skey, win[2], win[3], c);
  } else {
    const uint32_t q = lb_u64(a.skey, 0u, ns, c);
    /* all slots with (oldkey, slot) < (k, i): old keys are sorted in slot order */
    uint32_t lo = 0u, hi = (uint32_t)n;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (comp(a.okeys[mid], mid) < c) lo = mid + 1; else hi = mid;
    }
    const uint32_t all_before = lo;
    /* changed slots with (oldkey, slot) < (k, i): sslot is in slot order, so
     * (oldkey, slot) is increasing along it too */
    lo = 0u; hi = ns;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      const uint32_t s = a.sslot[mid];
      if (comp(a.okeys[s], s) < c) lo = mid + 1; else hi = mid;
    }
    pos = q + (all_before - lo);
  }
  if (pos >= (uint32_t)n) {   /* cannot happen for a consistent changed set; never write out of range */
    atomicOr(a.status, ST_SORT_ERROR);
    return;
  }
  a.keys_sorted[pos] = k;
  a.idx_sorted[pos] = i;
  a.pos_srt[pos] = a.pos_cur[i];
  a.vel_srt[pos] = a.vel_cur[i];
  a.omg_srt[pos] = a.omg_cur[i];
  if (a.quat_cur) a.quat_srt[pos] = a.quat_cur[i];
}

/* Kernel_SpatialDecomp_BinData (Collision_Detection_BroadPhase.cuh:279-325):
 * (start, end) of every occupied cell; sentinel keys own no cell. */
__device__ __forceinline__ void bin_one(int n, int i, const uint32_t* __restrict__ keys,
                                        uint2* __restrict__ cells, uint32_t C) {
  const uint32_t k = keys[i];
  if (k >= C) return;
  if (i == 0 || keys[i - 1] != k) cells[k].x = (uint32_t)i;
  if (i == n - 1 || keys[i + 1] != k) cells[k].y = (uint32_t)(i + 1);
}

__global__ __launch_bounds__(256) void k_bin_gated(SortArgs a) {
  if (a.meta[META_FALLBACK] != 0u) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < a.n) bin_one(a.n, i, a.keys_sorted, a.cells, a.C);
}

/* ------------------------------------------------------------------------ */
/* Full path: one cooperative kernel                                          */
/* ------------------------------------------------------------------------ */
__global__ __launch_bounds__(SORT_THREADS) void k_full_sort(SortArgs a, int gated) {
  if (gated && a.meta[META_FALLBACK] == 0u) return;   /* uniform across the grid */
  cg::grid_group grid = cg::this_grid();
  __shared__ uint32_t wcnt[SORT_THREADS / 64][RADIX];
  __shared__ uint32_t red[SORT_THREADS];
  __shared__ uint32_t carry;
  const int n = a.n;
  const int ntiles = sort_blocks(n);
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  const int m = RADIX * ntiles;

  const uint32_t* kin = a.keys_new;
  const int* vin = nullptr;
  for (int p = 0; p < a.passes; ++p) {
    const int shift = 8 * p;
    const bool last = p == a.passes - 1;
    uint32_t* kout = last ? a.keys_sorted : a.rk[p & 1];
    int* vout = last ? a.idx_sorted : a.rv[p & 1];
    /* 1. digit histogram per tile: hist[digit * ntiles + tile] */
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
      uint32_t* h = &wcnt[0][0];
      for (int d = threadIdx.x; d < RADIX; d += SORT_THREADS) h[d] = 0u;
      __syncthreads();
      for (int q = threadIdx.x; q < SORT_TILE; q += SORT_THREADS) {
        const int i = t * SORT_TILE + q;
        if (i < n) atomicAdd(&h[(kin[i] >> shift) & (RADIX - 1)], 1u);
      }
      __syncthreads();
      for (int d = threadIdx.x; d < RADIX; d += SORT_THREADS) a.hist[d * ntiles + t] = h[d];
      __syncthreads();
    }
    grid.sync();
    /* 2. exclusive scan of the m counters: chunk sums, one block scans them, chunk rescan */
    const int chunk = (m + gridDim.x - 1) / gridDim.x;
    const int c0 = blockIdx.x * chunk, c1 = min(c0 + chunk, m);
    {
      uint32_t s = 0u;
      for (int i = c0 + threadIdx.x; i < c1; i += SORT_THREADS) s += a.hist[i];
      red[threadIdx.x] = s;
      __syncthreads();
      for (int o = SORT_THREADS / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
      }
      if (threadIdx.x == 0) a.part[blockIdx.x] = red[0];
    }
    grid.sync();
    if (blockIdx.x == 0) {
      if (threadIdx.x == 0) carry = 0u;
      __syncthreads();
      for (int base = 0; base < (int)gridDim.x; base += SORT_THREADS) {
        const int i = base + threadIdx.x;
        const uint32_t x = i < (int)gridDim.x ? a.part[i] : 0u;
        red[threadIdx.x] = x;
        __syncthreads();
        for (int o = 1; o < SORT_THREADS; o <<= 1) {
          const uint32_t y = threadIdx.x >= (unsigned)o ? red[threadIdx.x - o] : 0u;
          __syncthreads();
          red[threadIdx.x] += y;
          __syncthreads();
        }
        if (i < (int)gridDim.x) a.part[i] = carry + red[threadIdx.x] - x;
        __syncthreads();
        if (threadIdx.x == SORT_THREADS - 1) carry += red[SORT_THREADS - 1];
        __syncthreads();
      }
    }
    grid.sync();
    {
      if (threadIdx.x == 0) carry = a.part[blockIdx.x];
      __syncthreads();
      for (int base = c0; base < c1; base += SORT_THREADS) {
        const int i = base + threadIdx.x;
        const uint32_t x = i < c1 ? a.hist[i] : 0u;
        red[threadIdx.x] = x;
        __syncthreads();
        for (int o = 1; o < SORT_THREADS; o <<= 1) {
          const uint32_t y = threadIdx.x >= (unsigned)o ? red[threadIdx.x - o] : 0u;
          __syncthreads();
          red[threadIdx.x] += y;
          __syncthreads();
        }
        if (i < c1) a.hist[i] = carry + red[threadIdx.x] - x;
        __syncthreads();
        if (threadIdx.x == SORT_THREADS - 1) carry += red[SORT_THREADS - 1];
        __syncthreads();
      }
    }
    grid.sync();
    /* 3. scatter each tile (stable wave-ballot ranking, as a one-pass kernel would) */
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
      for (int d = threadIdx.x; d < RADIX; d += SORT_THREADS)
        for (int w = 0; w < SORT_THREADS / 64; ++w) wcnt[w][d] = 0u;
      __syncthreads();
      const int base = t * SORT_TILE + wave * (SORT_ITEMS * 64);
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
          const uint64_t mm = __ballot(set);
          eq &= set ? mm : ~mm;
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
        uint32_t s = a.hist[d * ntiles + t];
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
      __syncthreads();
    }
    grid.sync();
    kin = kout;
    vin = vout;
  }
  /* gather + table clear, then bin */
  const int gsize = gridDim.x * SORT_THREADS;
  for (int i = blockIdx.x * SORT_THREADS + threadIdx.x; i < n; i += gsize) {
    const int j = a.idx_sorted[i];
    a.pos_srt[i] = a.pos_cur[j];
    a.vel_srt[i] = a.vel_cur[j];
    a.omg_srt[i] = a.omg_cur[j];
    if (a.quat_cur) a.quat_srt[i] = a.quat_cur[j];
  }
  for (uint32_t c = blockIdx.x * SORT_THREADS + threadIdx.x; c < a.C; c += (uint32_t)gsize)
    a.cells[c] = make_uint2(0xFFFFFFFFu, 0xFFFFFFFFu);
  grid.sync();
  for (int i = blockIdx.x * SORT_THREADS + threadIdx.x; i < n; i += gsize)
    bin_one(n, i, a.keys_sorted, a.cells, a.C);
}

static int coop_grid = 0;

static int full_sort_grid() {
  if (coop_grid > 0) return coop_grid;
  int dev = 0, cus = 0, per = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, reinterpret_cast<const void*>(k_full_sort),
                                                     SORT_THREADS, 0);
  int g = cus * (per > 0 ? per : 1);
  if (g > 1024) g = 1024;     /* part[] capacity */
  if (g < 1) g = 1;
  coop_grid = g;
  return g;
}

hipError_t launch_full_sort(const SortArgs& a, bool gated, hipStream_t s) {
  SortArgs args = a;
  int gflag = gated ? 1 : 0;
  void* params[] = {&args, &gflag};
  return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_full_sort), dim3(full_sort_grid()),
                                    dim3(SORT_THREADS), params, 0, s);
}

__global__ void k_reset_append(uint32_t* meta) { meta[META_APPEND] = 0u; }

hipError_t launch_sort(const SortArgs& a, bool incremental, hipStream_t s, hipEvent_t* ev) {
  if (!incremental) {
    /* the step kernel will append afresh */
    hipLaunchKernelGGL(k_reset_append, dim3(1), dim3(1), 0, s, a.meta);
    if (ev) { (void)hipEventRecord(ev[0], s); (void)hipEventRecord(ev[1], s); }
    hipError_t e = launch_full_sort(a, false, s);
    if (ev) (void)hipEventRecord(ev[2], s);
    return e;
  }
  hipLaunchKernelGGL(k_sort_small, dim3(1), dim3(1024), 0, s, a);
  if (ev) (void)hipEventRecord(ev[0], s);
  const dim3 g((a.n + 255) / 256);
  hipLaunchKernelGGL(k_place, g, dim3(256), 0, s, a);
  if (ev) (void)hipEventRecord(ev[1], s);
  hipLaunchKernelGGL(k_bin_gated, g, dim3(256), 0, s, a);
  hipError_t e = launch_full_sort(a, true, s);
  if (ev) (void)hipEventRecord(ev[2], s);
  return e;
}

}  // namespace bdem
