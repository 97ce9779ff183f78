/*
 * bdem_sort.hip - a2-a4: stable sort of particle slots by cell key, the
 * sorted gather of the state and the cell (start, end) table.
 * Replaces thrust::stable_sort_by_key + Kernel_SortArrays +
 * Kernel_SpatialDecomp_ReorderPDA + the 0xFF fill + Kernel_SpatialDecomp_BinData
 * (Collision_Detection_BroadPhase.cuh:110-325, Device/DeviceInterface.cu:2097-2105).
 *
 * Result: keys ascending, ties in ascending slot order -- exactly what a
 * stable sort of (key, slot) produces -- so the sorted order is bit-exact.
 * Elements are compared as composites c = key << 32 | slot, which are unique.
 *
 * Incremental path (every step after the first): one kernel, k_resort.
 * The slots are in the previous step's sorted order, so the old composites
 * increase with the slot.  The step kernel flagged every particle whose key
 * changed and appended its new composite to a list S (any order).  Unchanged
 * particles keep their composite and their relative order.
 *
 * Block b owns slots [s0, s1).  Let W_lo be the composite of the first
 * unchanged slot >= s0 (0 for block 0) and W_hi that of the first unchanged
 * slot >= s1 (infinity past the end).  The windows [W_lo, W_hi) partition
 * the composite space, so every element of the new order belongs to exactly
 * one block: its unchanged slots U_b and the changed composites L_b that
 * fall in its window.  With below = #{S < W_lo} and chg_before = #{S: slot < s0}
 * the block's output starts at base = (s0 - chg_before) + below, and
 *   unchanged u:  pos = base + #{U_b: slot < u} + #{L_b < c_u}
 *   changed  l:   pos = base + #{L_b < l}      + #{U_b < l}
 * Each block reads all of S once (it is small), sorts L_b in LDS, writes
 * its keys, slots and gathered state, then the cell bounds of its output:
 * the neighbours across its output edges are the largest element below its
 * window and the smallest above it, both known from the same pass over S.
 * A cell is cleared only when it becomes empty: every particle of its old
 * run changed key and no changed particle arrives in it.  The block owning
 * the first slot of that run decides it, so no cell has two writers.
 * When L_b exceeds LDS the block sorts it in global scratch instead (slow,
 * correct): the kernel handles any S.
 *
 * Full path: an LSD radix sort (8-bit digits, wave64 ballot ranking) by
 * ordinary launches, then the gather, the table clear and the binning.  The
 * host runs it on the first step after an upload and after a discharge
 * check, when many particles change key at once.
 */
#include "bdem_kernels.hpp"


namespace bdem {

int key_bits(uint32_t max_key) {
  int bits = 1;
  while (bits < 32 && (max_key >> bits) != 0) ++bits;
  return bits;
}

int sort_passes(uint32_t max_key) { return (key_bits(max_key) + 7) / 8; }

__device__ __forceinline__ unsigned long long comp(uint32_t k, uint32_t s) {
  return ((unsigned long long)k << 32) | s;
}
constexpr unsigned long long COMP_INF = ~0ull;

/* ------------------------------------------------------------------------ */
/* Incremental path                                                          */
/* ------------------------------------------------------------------------ */
constexpr int RS_ITEMS = RS_TILE / RS_THREADS;
constexpr int RS_ROWS = RS_ITEMS * (RS_THREADS / 64);   /* (item row, wave) counts per block */
constexpr int RS_UNROLL = 4;
static_assert(RS_ROWS <= 64 && RS_THREADS >= 256, "one wave (the fourth) scans the per-(item row, wave) counts");

/* Ascending sort of v[0, cnt) by a bitonic network whose first stage of
 * every merge compares i with i ^ (k - 1): every comparator puts the
 * smaller value at the lower index, so indices >= cnt act as +infinity and
 * are skipped.  V is an LDS or global array; the caller's threads all call. */
template <class V>
__device__ __forceinline__ void block_sort_u64(V v, uint32_t cnt) {
  uint32_t m = 1u;
  while (m < cnt) m <<= 1;
  for (uint32_t k = 2; k <= m; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) {
        const uint32_t l = (j == (k >> 1)) ? (i ^ (k - 1u)) : (i ^ j);
        if (l > i && l < cnt) {
          const unsigned long long x = v[i], y = v[l];
          if (x > y) { v[i] = y; v[l] = x; }
        }
      }
      __syncthreads();
    }
  }
}

template <class V>
__device__ __forceinline__ uint32_t lower_bound_u64(V v, uint32_t cnt, unsigned long long x) {
  uint32_t lo = 0u, hi = cnt;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (v[mid] < x) lo = mid + 1; else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long w = __shfl_xor(v, o, 64);
    v = w < v ? w : v;
  }
  return v;
}
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long w = __shfl_xor(v, o, 64);
    v = w > v ? w : v;
  }
  return v;
}
__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += (uint32_t)__shfl_xor((int)v, o, 64);
  return v;
}

/* First slot >= from (step +1) or last slot <= from (step -1) that is
 * unchanged, scanning 64 flags per wave iteration; -1 if none.  One wave. */
__device__ __forceinline__ int wave_find_unchanged(const uint8_t* __restrict__ changed, int n, int from,
                                                   int dir) {
  const int lane = threadIdx.x & 63;
  for (int base = from; base >= 0 && base < n; base += 64 * dir) {
    const int s = base + dir * lane;
    const bool u = s >= 0 && s < n && changed[s] == 0;
    const uint64_t m = __ballot(u);
    if (m) return base + dir * (__ffsll((unsigned long long)m) - 1);
  }
  return -1;
}

__global__ __launch_bounds__(RS_THREADS) void k_resort(SortArgs a, int par) {
  __shared__ uint32_t okey_l[RS_TILE];           /* old keys of the block's slots */
  __shared__ uint16_t ucnt[RS_TILE + 1];         /* exclusive prefix of unchanged slots */
  __shared__ uint8_t cand[RS_TILE];              /* 1: candidate cell to clear, 2: refilled */
  __shared__ unsigned long long Ls[RS_LCAP];     /* L_b */
  __shared__ uint32_t wcnt[RS_ITEMS * (RS_THREADS / 64)];   /* unchanged per (item row, wave) */
  __shared__ uint32_t wpre[RS_ITEMS * (RS_THREADS / 64)];
  __shared__ unsigned long long s_wlo, s_whi, s_upred, s_maxb, s_mina;
  __shared__ uint32_t s_below, s_chgb, s_nl, s_n2, s_ncand;

  const int n = a.n;
  const uint32_t C = a.C;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int s0 = blockIdx.x * RS_TILE;
  const int cnt = min(RS_TILE, n - s0);
  uint32_t ns = a.meta[META_APPEND + par];
  if (ns > (uint32_t)n) {   /* each slot appends at most once: never read past the list */
    if (tid == 0) atomicOr(a.status, ST_SORT_ERROR);
    ns = (uint32_t)n;
  }
  if (blockIdx.x == 0 && tid == 0) {
    a.meta[META_APPEND + (par ^ 1)] = 0u;     /* the next step kernel appends there */
    a.meta[META_COUNT] = ns;
  }

  /* 1. the block's slots: old keys, changed flags, unchanged prefix.  Item
   * it of thread tid is slot t = it * RS_THREADS + tid, so every load and
   * store of a wave covers 64 consecutive slots. */
  uint32_t ok[RS_ITEMS];
  bool ch[RS_ITEMS];
  uint64_t bal[RS_ITEMS];
  const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
#pragma unroll
  for (int it = 0; it < RS_ITEMS; ++it) {
    const int t = it * RS_THREADS + tid;
    ok[it] = t < cnt ? a.okeys[s0 + t] : 0xFFFFFFFFu;
    ch[it] = t < cnt ? a.changed[s0 + t] != 0 : true;
    if (t < cnt) okey_l[t] = ok[it];
    cand[t] = 0;
    bal[it] = __ballot(!ch[it]);
    if (lane == 0) wcnt[it * (RS_THREADS / 64) + wave] = (uint32_t)__popcll(bal[it]);
  }
  if (tid == 0) {
    s_below = 0u; s_chgb = 0u; s_nl = 0u; s_n2 = 0u; s_ncand = 0u;
    s_maxb = 0ull; s_mina = COMP_INF;
  }
  __syncthreads();

  if (wave == 3) {
    /* exclusive scan of the RS_ROWS (item row, wave) counts, in slot order */
    const uint32_t v = lane < RS_ROWS ? wcnt[lane] : 0u;
    uint32_t incl = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t u = (uint32_t)__shfl_up((int)incl, o, 64);
      if (lane >= o) incl += u;
    }
    if (lane < RS_ROWS) wpre[lane] = incl - v;
    if (lane == RS_ROWS - 1) ucnt[RS_TILE] = (uint16_t)incl;
  } else if (wave == 0) {
    /* window ends and the last unchanged element before the block */
    unsigned long long wlo = COMP_INF;
    if (blockIdx.x == 0) {
      wlo = 0ull;
    } else {
      const int u = wave_find_unchanged(a.changed, n, s0, 1);
      if (u >= 0) wlo = comp(a.okeys[u], (uint32_t)u);
    }
    if (lane == 0) s_wlo = wlo;
  } else if (wave == 1) {
    const int u = wave_find_unchanged(a.changed, n, s0 + cnt, 1);
    if (lane == 0) s_whi = u >= 0 ? comp(a.okeys[u], (uint32_t)u) : COMP_INF;
  } else if (wave == 2) {
    const int u = s0 > 0 ? wave_find_unchanged(a.changed, n, s0 - 1, -1) : -1;
    if (lane == 0) s_upred = u >= 0 ? comp(a.okeys[u], (uint32_t)u) + 1ull : 0ull;   /* +1; 0: none */
  }
  /* cells the changed particles may leave empty: first slot of an old run,
   * every slot of the run changed (the run may continue past the block) */
#pragma unroll
  for (int it = 0; it < RS_ITEMS; ++it) {
    const int t = it * RS_THREADS + tid, s = s0 + t;
    if (t >= cnt || !ch[it] || ok[it] >= C) continue;
    if (s > 0 && (t > 0 ? okey_l[t - 1] : a.okeys[s - 1]) == ok[it]) continue;
    bool all = true;
    for (int j = s + 1; j < n && a.okeys[j] == ok[it]; ++j)
      if (a.changed[j] == 0) { all = false; break; }
    if (all) { cand[t] = 1; atomicAdd(&s_ncand, 1u); }
  }
  __syncthreads();
  uint32_t myc[RS_ITEMS];   /* unchanged slots of the block before my items */
#pragma unroll
  for (int it = 0; it < RS_ITEMS; ++it) {
    myc[it] = wpre[it * (RS_THREADS / 64) + wave] + (uint32_t)__popcll(bal[it] & lt);
    ucnt[it * RS_THREADS + tid] = (uint16_t)myc[it];
  }

  /* 2. one pass over S */
  const unsigned long long wlo = s_wlo, whi = s_whi;
  const uint32_t kfirst = okey_l[0], klast = okey_l[cnt - 1];
  uint32_t below = 0u, chgb = 0u;
  unsigned long long maxb = 0ull, mina = COMP_INF;   /* maxb holds c + 1; 0: none */
  const bool any_cand = s_ncand != 0u;
  /* RS_UNROLL independent loads in flight per thread */
  for (uint32_t q0 = 0; q0 < ns; q0 += RS_THREADS * RS_UNROLL) {
    unsigned long long cc[RS_UNROLL];
#pragma unroll
    for (int u = 0; u < RS_UNROLL; ++u) {
      const uint32_t q = q0 + (uint32_t)(u * RS_THREADS + tid);
      cc[u] = q < ns ? a.sbuf[q] : 0ull;
    }
#pragma unroll
    for (int u = 0; u < RS_UNROLL; ++u) {
      if (q0 + (uint32_t)(u * RS_THREADS + tid) >= ns) continue;
      const unsigned long long c = cc[u];
      const uint32_t k = (uint32_t)(c >> 32), s = (uint32_t)c;
      chgb += s < (uint32_t)s0 ? 1u : 0u;
      if (c < wlo) {
        ++below;
        maxb = c + 1ull > maxb ? c + 1ull : maxb;
      } else if (c < whi) {
        const uint32_t idx = atomicAdd(&s_nl, 1u);
        if (idx < (uint32_t)RS_LCAP) Ls[idx] = c;
      } else {
        mina = c < mina ? c : mina;
      }
      if (any_cand && k >= kfirst && k <= klast && k < C) {
        uint32_t lo = 0u, hi = (uint32_t)cnt;
        while (lo < hi) {
          const uint32_t mid = (lo + hi) >> 1;
          if (okey_l[mid] < k) lo = mid + 1; else hi = mid;
        }
        if (lo < (uint32_t)cnt && okey_l[lo] == k && cand[lo] != 0) cand[lo] = 2;
      }
    }
  }
  below = wave_sum_u32(below);
  chgb = wave_sum_u32(chgb);
  maxb = wave_max_u64(maxb);
  mina = wave_min_u64(mina);
  if (lane == 0) {
    if (below) atomicAdd(&s_below, below);
    if (chgb) atomicAdd(&s_chgb, chgb);
    if (maxb) atomicMax(&s_maxb, maxb);
    if (mina != COMP_INF) atomicMin(&s_mina, mina);
  }
  __syncthreads();

  /* 3. sort L_b: in LDS, or in global scratch past the LDS capacity */
  const uint32_t nl = s_nl;
  below = s_below;
  if (s_chgb > (uint32_t)s0 || below + nl > ns) {   /* inconsistent list (uniform: return together) */
    if (tid == 0) atomicOr(a.status, ST_SORT_ERROR);
    return;
  }
  const bool big = nl > (uint32_t)RS_LCAP;
  unsigned long long* gl = a.skey + below;   /* the block's own range of the scratch */
  if (!big) {
    block_sort_u64(Ls, nl);
  } else {
    for (uint32_t q = tid; q < ns; q += RS_THREADS) {
      const unsigned long long c = a.sbuf[q];
      if (c >= wlo && c < whi) gl[atomicAdd(&s_n2, 1u)] = c;
    }
    __syncthreads();
    block_sort_u64(gl, nl);
  }
  __syncthreads();

  /* 4. positions, keys, slots and the gather */
  const uint32_t base = (uint32_t)s0 - s_chgb + below;
  const uint32_t nu_b = ucnt[RS_TILE];
  const uint32_t total = nu_b + nl;
#pragma unroll
  for (int it = 0; it < RS_ITEMS; ++it) {
    const int t = it * RS_THREADS + tid, s = s0 + t;
    if (t >= cnt || ch[it]) continue;
    const unsigned long long c = comp(ok[it], (uint32_t)s);
    const uint32_t r = nl == 0u ? 0u : (big ? lower_bound_u64(gl, nl, c) : lower_bound_u64(Ls, nl, c));
    const uint32_t pos = base + myc[it] + r;
    if (pos >= (uint32_t)n) { atomicOr(a.status, ST_SORT_ERROR); continue; }
    a.keys_sorted[pos] = ok[it];
    a.idx_sorted[pos] = s;
    if (a.inv_map) a.inv_map[s] = pos;
    a.pos_srt[pos] = a.pos_cur[s];
    a.vel_srt[pos] = a.vel_cur[s];
    a.omg_srt[pos] = a.omg_cur[s];
    if (a.quat_cur) a.quat_srt[pos] = a.quat_cur[s];
    if (a.hwc_cur) a.hwc_srt[pos] = a.hwc_cur[s];
    if (a.mask_cur) a.mask_srt[pos] = a.mask_cur[s];
  }
  for (uint32_t r = tid; r < nl; r += RS_THREADS) {
    const unsigned long long c = big ? gl[r] : Ls[r];
    const uint32_t k = (uint32_t)(c >> 32), s = (uint32_t)c;
    /* #{U_b < c}: old composites increase along the block's slots */
    uint32_t lo = 0u, hi = (uint32_t)cnt;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (comp(okey_l[mid], (uint32_t)s0 + mid) < c) lo = mid + 1; else hi = mid;
    }
    const uint32_t pos = base + r + ucnt[lo];
    if (pos >= (uint32_t)n || s >= (uint32_t)n) { atomicOr(a.status, ST_SORT_ERROR); continue; }
    a.keys_sorted[pos] = k;
    a.idx_sorted[pos] = (int)s;
    if (a.inv_map) a.inv_map[s] = pos;
    a.pos_srt[pos] = a.pos_cur[s];
    a.vel_srt[pos] = a.vel_cur[s];
    a.omg_srt[pos] = a.omg_cur[s];
    if (a.quat_cur) a.quat_srt[pos] = a.quat_cur[s];
    if (a.hwc_cur) a.hwc_srt[pos] = a.hwc_cur[s];
    if (a.mask_cur) a.mask_srt[pos] = a.mask_cur[s];
  }
  __syncthreads();   /* the block's keys_sorted writes are visible to the block */

  /* 5. cell bounds of the block's output range */
  if (base + total <= (uint32_t)n) {
    const unsigned long long pred = s_maxb > s_upred ? s_maxb : s_upred;   /* composite + 1; 0: none */
    const unsigned long long succ = s_mina < whi ? s_mina : whi;            /* INF: none */
    for (uint32_t q = tid; q < total; q += RS_THREADS) {
      const uint32_t p = base + q;
      const uint32_t k = a.keys_sorted[p];
      if (k >= C) continue;
      bool first, last;
      if (q > 0) first = a.keys_sorted[p - 1] != k;
      else first = p == 0 || pred == 0ull || (uint32_t)((pred - 1ull) >> 32) != k;
      if (q + 1 < total) last = a.keys_sorted[p + 1] != k;
      else last = succ == COMP_INF || (uint32_t)(succ >> 32) != k;
      if (first) a.cells[k].x = p;
      if (last) a.cells[k].y = p + 1;
    }
  }
  /* 6. cells left empty */
#pragma unroll
  for (int it = 0; it < RS_ITEMS; ++it) {
    const int t = it * RS_THREADS + tid;
    if (t < cnt && cand[t] == 1) a.cells[ok[it]] = make_uint2(0xFFFFFFFFu, 0xFFFFFFFFu);
  }
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

/* ------------------------------------------------------------------------ */
/* Full path: LSD radix sort by ordinary launches (per 8-bit digit: tile     */
/* histograms, one scan, stable scatter), then gather, table clear and bin.  */
/* No cooperative launch: those hold the device's global wave-sync resource,  */
/* which one process at a time can own.                                       */
/* ------------------------------------------------------------------------ */
/* Digit histogram of tile blockIdx.x: hist[digit * ntiles + tile].  The
 * first pass also resets the next step kernel's append counter. */
__global__ __launch_bounds__(SORT_THREADS) void k_fs_hist(SortArgs a, int par, const uint32_t* __restrict__ kin,
                                                          int shift, int first) {
  __shared__ uint32_t h[RADIX];
  const int n = a.n;
  const int ntiles = sort_blocks(n);
  const int t = blockIdx.x;
  if (first && t == 0 && threadIdx.x == 0) {
    a.meta[META_APPEND + (par ^ 1)] = 0u;   /* the next step kernel appends there */
    a.meta[META_COUNT] = (uint32_t)n;
  }
  for (int d = threadIdx.x; d < RADIX; d += SORT_THREADS) h[d] = 0u;
  __syncthreads();
  for (int q = threadIdx.x; q < SORT_TILE; q += SORT_THREADS) {
    const int i = t * SORT_TILE + q;
    if (i < n) atomicAdd(&h[(kin[i] >> shift) & (RADIX - 1)], 1u);
  }
  __syncthreads();
  for (int d = threadIdx.x; d < RADIX; d += SORT_THREADS) a.hist[d * ntiles + t] = h[d];
}

/* Exclusive scan of the m digit-major counters in one block, a tile of
 * 4 x 1024 counters at a time: each thread loads four consecutive counters
 * (one 16-byte load), the 1024 thread sums are scanned by wave shuffles and
 * one LDS pass over the 16 wave totals, and the running total carries to
 * the next tile. */
constexpr int SCAN_THREADS = 1024;
__global__ __launch_bounds__(SCAN_THREADS) void k_fs_scan(uint32_t* __restrict__ hist, int m) {
  __shared__ uint32_t wsum[SCAN_THREADS / 64];
  __shared__ uint32_t carry_s;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (threadIdx.x == 0) carry_s = 0u;
  __syncthreads();
  for (int base = 0; base < m; base += 4 * SCAN_THREADS) {
    const int i0 = base + 4 * (int)threadIdx.x;
    uint32_t x[4];
    if (i0 + 3 < m) {
      const uint4 v = *reinterpret_cast<const uint4*>(hist + i0);   /* m is a multiple of 256 */
      x[0] = v.x; x[1] = v.y; x[2] = v.z; x[3] = v.w;
    } else {
      for (int k = 0; k < 4; ++k) x[k] = i0 + k < m ? hist[i0 + k] : 0u;
    }
    const uint32_t t = x[0] + x[1] + x[2] + x[3];
    uint32_t incl = t;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(incl, o, 64);
      if (lane >= o) incl += y;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    uint32_t before = carry_s;
    for (int w = 0; w < wave; ++w) before += wsum[w];
    uint32_t run = before + incl - t;
    if (i0 + 3 < m) {
      uint4 o;
      o.x = run; run += x[0];
      o.y = run; run += x[1];
      o.z = run; run += x[2];
      o.w = run;
      *reinterpret_cast<uint4*>(hist + i0) = o;
    } else {
      for (int k = 0; k < 4; ++k)
        if (i0 + k < m) { hist[i0 + k] = run; run += x[k]; }
    }
    __syncthreads();
    if (threadIdx.x == SCAN_THREADS - 1) {
      uint32_t tot = 0u;
      for (int w = 0; w < SCAN_THREADS / 64; ++w) tot += wsum[w];
      carry_s += tot;
    }
    __syncthreads();
  }
}

/* Stable scatter of tile blockIdx.x: each wave ranks its 16 x 64 keys by
 * ballots over the digit bits (equal digits in lane order), the waves'
 * counts are offset in wave order on top of the tile's scanned base. */
__global__ __launch_bounds__(SORT_THREADS) void k_fs_scatter(SortArgs a, const uint32_t* __restrict__ kin,
                                                             const int* __restrict__ vin, uint32_t* __restrict__ kout,
                                                             int* __restrict__ vout, int shift) {
  __shared__ uint32_t wcnt[SORT_THREADS / 64][RADIX];
  const int n = a.n;
  const int ntiles = sort_blocks(n);
  const int t = blockIdx.x;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
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
}

/* The sorted order's state gather, and every cell marked empty. */
__global__ __launch_bounds__(SORT_THREADS) void k_fs_gather(SortArgs a) {
  const int gsize = gridDim.x * SORT_THREADS;
  for (int i = blockIdx.x * SORT_THREADS + threadIdx.x; i < a.n; i += gsize) {
    const int j = a.idx_sorted[i];
    a.pos_srt[i] = a.pos_cur[j];
    a.vel_srt[i] = a.vel_cur[j];
    a.omg_srt[i] = a.omg_cur[j];
    if (a.quat_cur) a.quat_srt[i] = a.quat_cur[j];
    if (a.hwc_cur) a.hwc_srt[i] = a.hwc_cur[j];
  }
  for (uint32_t c = blockIdx.x * SORT_THREADS + threadIdx.x; c < a.C; c += (uint32_t)gsize)
    a.cells[c] = make_uint2(0xFFFFFFFFu, 0xFFFFFFFFu);
}

__global__ __launch_bounds__(SORT_THREADS) void k_fs_bin(SortArgs a) {
  const int i = blockIdx.x * SORT_THREADS + threadIdx.x;
  if (i < a.n) bin_one(a.n, i, a.keys_sorted, a.cells, a.C);
}

static hipError_t launch_full_sort(const SortArgs& a, int par, hipStream_t s) {
  const int ntiles = sort_blocks(a.n);
  const uint32_t* kin = a.keys_new;
  const int* vin = nullptr;
  for (int p = 0; p < a.passes; ++p) {
    const int shift = 8 * p;
    const bool last = p == a.passes - 1;
    uint32_t* kout = last ? a.keys_sorted : a.rk[p & 1];
    int* vout = last ? a.idx_sorted : a.rv[p & 1];
    hipLaunchKernelGGL(k_fs_hist, dim3(ntiles), dim3(SORT_THREADS), 0, s, a, par, kin, shift, p == 0 ? 1 : 0);
    hipLaunchKernelGGL(k_fs_scan, dim3(1), dim3(SCAN_THREADS), 0, s, a.hist, RADIX * ntiles);
    hipLaunchKernelGGL(k_fs_scatter, dim3(ntiles), dim3(SORT_THREADS), 0, s, a, kin, vin, kout, vout, shift);
    kin = kout;
    vin = vout;
  }
  int g = (int)((std::max<uint64_t>((uint64_t)a.n, (uint64_t)a.C) + SORT_THREADS - 1) / SORT_THREADS);
  if (g > 8192) g = 8192;
  if (g < 1) g = 1;
  hipLaunchKernelGGL(k_fs_gather, dim3(g), dim3(SORT_THREADS), 0, s, a);
  hipLaunchKernelGGL(k_fs_bin, dim3((a.n + SORT_THREADS - 1) / SORT_THREADS), dim3(SORT_THREADS), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_sort(const SortArgs& a, bool incremental, int par, hipStream_t s) {
  if (incremental) {
    hipLaunchKernelGGL(k_resort, dim3((a.n + RS_TILE - 1) / RS_TILE), dim3(RS_THREADS), 0, s, a, par);
    return hipGetLastError();
  }
  return launch_full_sort(a, par, s);
}

}  // namespace bdem
