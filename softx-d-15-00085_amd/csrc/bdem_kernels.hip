/*
 * bdem_kernels.hip - gfx950 kernels of one DEM step.
 *
 *   k_hash           a1  cell key per particle (only when the state was set
 *                        from the host; otherwise the step kernel emits keys)
 *   k_radix_*        a2  stable LSD radix sort, 8-bit digits, wave64 ballot
 *                        ranking (replaces thrust::stable_sort_by_key)
 *   k_gather         a3  one gather of the three state streams into sorted
 *                        order (replaces Kernel_SortArrays + ReorderPDA)
 *   k_fill/k_bin     a4  cell (start,end) table in the reference format
 *   k_step           a5+a6+a8+a10+a11 fused: contact search, pair force,
 *                        wall force, symplectic Euler, discharge check, and
 *                        the next step's key
 *   k_contact_list,
 *   k_forces             the same code paths writing intermediate results
 *                        (contact lists, per-particle forces) for parity.
 */
#include "bdem_kernels.hpp"

namespace bdem {

/* ------------------------------------------------------------------------ */
__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* __restrict__ p, size_t count, uint32_t v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < count; i += stride) p[i] = v;
}

void launch_fill_u32(uint32_t* p, size_t count, uint32_t v, hipStream_t s) {
  if (count == 0) return;
  size_t blocks = (count + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, s, p, count, v);
}

/* ------------------------------------------------------------------------ */
/* a1: Kernel_SpatialDecomp_CalcPHash (Collision_Detection_BroadPhase.cuh:75-100) */
__global__ __launch_bounds__(256) void k_hash(const DevParams* __restrict__ Pp, int n,
                                              const float4* __restrict__ pos_r,
                                              const float4* __restrict__ omg_t,
                                              uint32_t* __restrict__ keys, uint32_t* status) {
  const DevParams& P = *Pp;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 p = pos_r[i];
  keys[i] = key_of(P, xyz(p), packed_pid(omg_t[i].w), status);
}

void launch_hash(const DevParams* dP, const StepBuffers& b, hipStream_t s) {
  hipLaunchKernelGGL(k_hash, dim3((b.n + 255) / 256), dim3(256), 0, s, dP, b.n, b.pos_cur,
                     b.omg_cur, b.keys[0], b.status);
}

/* ------------------------------------------------------------------------ */
/* a2: stable LSD radix sort                                                 */
/* ------------------------------------------------------------------------ */
int key_bits(uint32_t max_key) {
  int bits = 1;
  while (bits < 32 && (max_key >> bits) != 0) ++bits;
  return bits;
}

/* Digit histogram of one 4096-key tile: hist[digit * nblocks + block]. */
__global__ __launch_bounds__(SORT_THREADS) void k_radix_hist(const uint32_t* __restrict__ keys, int n,
                                                             int shift, uint32_t* __restrict__ hist,
                                                             int nblocks) {
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

/* Exclusive scan of m counters in one 1024-thread block. */
__global__ __launch_bounds__(1024) void k_scan_exclusive(uint32_t* __restrict__ data, int m) {
  __shared__ uint32_t sums[1024];
  const int t = threadIdx.x;
  const int per = (m + 1023) / 1024;
  const int b = t * per;
  const int e = min(b + per, m);
  uint32_t s = 0;
  for (int i = b; i < e; ++i) s += data[i];
  sums[t] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const uint32_t v = t >= off ? sums[t - off] : 0u;
    __syncthreads();
    sums[t] += v;
    __syncthreads();
  }
  uint32_t run = t ? sums[t - 1] : 0u;
  for (int i = b; i < e; ++i) {
    const uint32_t x = data[i];
    data[i] = run;
    run += x;
  }
}

/* Scatter one tile.  Wave w owns keys [w*1024, (w+1)*1024) of the tile and
 * walks them 64 at a time in order; lanes with equal digits are matched with
 * eight ballots, ranked by lane, and the group's leader bumps the wave's
 * LDS counter.  Ranks are therefore stable across lanes, iterations, waves
 * and (through the digit-major scanned histogram) blocks. */
__global__ __launch_bounds__(SORT_THREADS) void k_radix_scatter(
    const uint32_t* __restrict__ kin, const int* __restrict__ vin, uint32_t* __restrict__ kout,
    int* __restrict__ vout, int n, int shift, const uint32_t* __restrict__ hist_scan, int nblocks) {
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

int launch_sort(const StepBuffers& b, int bits, hipStream_t s) {
  const int nb = sort_blocks(b.n);
  int cur = 0;
  for (int shift = 0; shift < bits; shift += 8) {
    hipLaunchKernelGGL(k_radix_hist, dim3(nb), dim3(SORT_THREADS), 0, s, b.keys[cur], b.n, shift,
                       b.radix_hist, nb);
    hipLaunchKernelGGL(k_scan_exclusive, dim3(1), dim3(1024), 0, s, b.radix_hist, RADIX * nb);
    hipLaunchKernelGGL(k_radix_scatter, dim3(nb), dim3(SORT_THREADS), 0, s, b.keys[cur],
                       shift == 0 ? (const int*)nullptr : b.vals[cur], b.keys[cur ^ 1],
                       b.vals[cur ^ 1], b.n, shift, b.radix_hist, nb);
    cur ^= 1;
  }
  return cur;
}

/* ------------------------------------------------------------------------ */
/* a3: one gather into sorted order.                                          */
/* ------------------------------------------------------------------------ */
__global__ __launch_bounds__(256) void k_gather(int n, const int* __restrict__ idx,
                                                const float4* __restrict__ p_in,
                                                const float4* __restrict__ v_in,
                                                const float4* __restrict__ o_in,
                                                const float4* __restrict__ q_in,
                                                float4* __restrict__ p_out, float4* __restrict__ v_out,
                                                float4* __restrict__ o_out, float4* __restrict__ q_out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int j = idx[i];
  p_out[i] = p_in[j];
  v_out[i] = v_in[j];
  o_out[i] = o_in[j];
  if (q_in) q_out[i] = q_in[j];
}

void launch_gather(const StepBuffers& b, const int* idx, bool orient, hipStream_t s) {
  hipLaunchKernelGGL(k_gather, dim3((b.n + 255) / 256), dim3(256), 0, s, b.n, idx, b.pos_cur,
                     b.vel_cur, b.omg_cur, orient ? b.quat_cur : nullptr, b.pos_srt, b.vel_srt,
                     b.omg_srt, orient ? b.quat_srt : nullptr);
}

/* ------------------------------------------------------------------------ */
/* a4: Kernel_SpatialDecomp_BinData (Collision_Detection_BroadPhase.cuh:279-325)
 * after the 0xFFFFFFFF fill (Device/DeviceInterface.cu:2097-2098).  Keys at or
 * beyond C (outside the grid, dead) own no cell. */
__global__ __launch_bounds__(256) void k_bin(int n, const uint32_t* __restrict__ keys,
                                             uint2* __restrict__ cells, uint32_t C) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t k = keys[i];
  if (k >= C) return;
  if (i == 0 || keys[i - 1] != k) cells[k].x = (uint32_t)i;
  if (i == n - 1 || keys[i + 1] != k) cells[k].y = (uint32_t)(i + 1);
}

void launch_cells(const DevParams* dP, const StepBuffers& b, const uint32_t* sorted_keys,
                  uint32_t ncells, hipStream_t s) {
  (void)dP;
  launch_fill_u32(reinterpret_cast<uint32_t*>(b.cells), (size_t)ncells * 2, 0xFFFFFFFFu, s);
  hipLaunchKernelGGL(k_bin, dim3((b.n + 255) / 256), dim3(256), 0, s, b.n, sorted_keys, b.cells,
                     ncells);
}

/* ------------------------------------------------------------------------ */
/* Pair forces of one particle (all laws).                                    */
/* ------------------------------------------------------------------------ */
struct PairHistCtx {
  const HistRec* in; int in_cnt; int cursor;
  HistRec* out; int out_cnt;
};

template <int LAW>
__device__ __forceinline__ int particle_pair_forces(const DevParams& P, const uint2* __restrict__ cells,
                                                    const float4* __restrict__ pos_r,
                                                    const float4* __restrict__ vel_m,
                                                    const float4* __restrict__ omg_t, int i, f3 pa,
                                                    float ra, f3 va, float ma, f3 wa, int ta, f3& F,
                                                    f3& T, PairHistCtx* hc) {
  return for_each_contact(P, cells, pos_r, i, pa, ra, [&](int j, float4 pb, int k) {
    const float4 vb = vel_m[j];
    const float4 ob = omg_t[j];
    if (LAW == BDEM_LAW_REFERENCE) {
      pair_ref(P, pa, ra, va, ma, wa, ta, xyz(pb), pb.w, xyz(vb), vb.w, xyz(ob),
               packed_type(ob.w), F, T);
    } else {
      const f3 rel = pa - xyz(pb);
      const float d = length(rel);
      const f3 n = rel / d;
      const float delta = (ra + pb.w) - d;
      const f3 la = (-1.0f * n) * ra;
      const f3 lb = n * pb.w;
      const f3 vc = (va + cross(wa, la)) - (xyz(vb) + cross(xyz(ob), lb));
      const float rstar = (ra * pb.w) / (ra + pb.w);
      const float mstar = (ma * vb.w) / (ma + vb.w);
      const int idb = packed_pid(ob.w);
      f3 xi = find_spring(hc->in, hc->in_cnt, idb, hc->cursor);
      const f3 f = history_law(P, n, delta, vc, rstar, mstar, P.law_pp[ta][packed_type(ob.w)], xi);
      F = F + f;
      if (P.rotation) T = T + cross(la, f);
      if (hc->out_cnt < HMAX) {
        hc->out[hc->out_cnt] = make_float4(xi.x, xi.y, xi.z, __int_as_float(idb));
        ++hc->out_cnt;
      }
    }
    (void)k;
  });
}

/* ------------------------------------------------------------------------ */
/* Fused step: contacts -> pair forces (+gravity) -> walls -> integrate ->    */
/* discharge check -> next key.  Reads the sorted scratch, writes the state.  */
/* ------------------------------------------------------------------------ */
template <int LAW, bool ORIENT>
__global__ __launch_bounds__(256) void k_step(
    const DevParams* __restrict__ Pp, int n, const uint2* __restrict__ cells,
    const float4* __restrict__ pos_in, const float4* __restrict__ vel_in,
    const float4* __restrict__ omg_in, const float4* __restrict__ quat_in,
    float4* __restrict__ pos_out, float4* __restrict__ vel_out, float4* __restrict__ omg_out,
    float4* __restrict__ quat_out, uint32_t* __restrict__ keys_out,
    const int* __restrict__ hc_in, const HistRec* __restrict__ hr_in, int* __restrict__ hc_out,
    HistRec* __restrict__ hr_out, const int* __restrict__ wc_in, const HistRec* __restrict__ wr_in,
    int* __restrict__ wc_out, HistRec* __restrict__ wr_out, int hist_cap,
    float* __restrict__ wall_force, int do_kill, uint32_t* status, int* kill_count) {
  const DevParams& P = *Pp;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 pr = pos_in[i];
  const float4 vm = vel_in[i];
  const float4 ot = omg_in[i];
  int pid = packed_pid(ot.w);
  const int ta = packed_type(ot.w);
  if (pid < 0) {
    pos_out[i] = pr;
    vel_out[i] = vm;
    omg_out[i] = ot;
    if (ORIENT) quat_out[i] = quat_in[i];
    keys_out[i] = P.C + 1u;
    if (wall_force) { wall_force[3 * i] = 0.f; wall_force[3 * i + 1] = 0.f; wall_force[3 * i + 2] = 0.f; }
    return;
  }
  f3 p = xyz(pr), v = xyz(vm), w = xyz(ot);
  const float r = pr.w, m = vm.w;

  /* gravity is part of the pair force, DEM_Kernels_Compute.cuh:654 */
  f3 Fpp = cvt(P.gravity) * m;
  f3 Tpp = mk3(0.f, 0.f, 0.f);
  PairHistCtx hc{nullptr, 0, 0, nullptr, 0};
  WallHist wh{nullptr, 0, nullptr, 0};
  const bool hist_slot = LAW != BDEM_LAW_REFERENCE && pid < hist_cap;
  if (LAW != BDEM_LAW_REFERENCE) {
    if (hist_slot) {
      hc.in = hr_in + (size_t)pid * HMAX;
      hc.in_cnt = hc_in[pid];
      hc.out = hr_out + (size_t)pid * HMAX;
      wh.in = wr_in + (size_t)pid * WMAX;
      wh.in_cnt = wc_in[pid];
      wh.out = wr_out + (size_t)pid * WMAX;
    }
  }
  HistRec scratch_pair[1];
  HistRec scratch_wall[1];
  if (LAW != BDEM_LAW_REFERENCE && !hist_slot) {
    hc.out = scratch_pair; hc.out_cnt = HMAX;   /* no slot: springs are not kept */
    wh.out = scratch_wall; wh.out_cnt = WMAX;
  }
  const int nc = particle_pair_forces<LAW>(P, cells, pos_in, vel_in, omg_in, i, p, r, v, m, w, ta,
                                           Fpp, Tpp, &hc);
  if (nc > LIST_STRIDE) atomicOr(status, ST_LIST_OVERFLOW);
  if (LAW != BDEM_LAW_REFERENCE && nc > HMAX) atomicOr(status, ST_HIST_OVERFLOW);
  if (!P.rotation) Tpp = mk3(0.f, 0.f, 0.f);

  f3 Fw, Tw;
  wall_forces(P, ta, p, r, v, w, m, Fw, Tw, &wh);
  if (wall_force) {
    wall_force[3 * i] = Fw.x; wall_force[3 * i + 1] = Fw.y; wall_force[3 * i + 2] = Fw.z;
  }
  if (hist_slot) {
    hc_out[pid] = hc.out_cnt;
    wc_out[pid] = wh.out_cnt;
  }

  const float inv_inertia = P.per_particle ? 1.0f / (0.4f * r * r * m) : P.t_inertia[ta];
  QuatF q{1.f, 0.f, 0.f, 0.f};
  if (ORIENT) { const float4 q4 = quat_in[i]; q = QuatF{q4.x, q4.y, q4.z, q4.w}; }
  integrate(P, inv_inertia, m, Fpp, Tpp, Fw, Tw, p, v, w, ORIENT ? &q : nullptr);

  /* a11: Remove_Silo_Discharged_Particles, GPU_Memory_Transfer.cuh:650-664 */
  if (do_kill && p.y < P.kill_height) {
    p = cvt(P.max_size);
    pid = -1;
    atomicAdd(kill_count, 1);
  }
  pos_out[i] = make_float4(p.x, p.y, p.z, r);
  vel_out[i] = make_float4(v.x, v.y, v.z, m);
  omg_out[i] = make_float4(w.x, w.y, w.z, __int_as_float(pack_id_type(pid, ta)));
  if (ORIENT) quat_out[i] = make_float4(q.w, q.x, q.y, q.z);
  keys_out[i] = key_of(P, p, pid, status);
}

template <int LAW, bool ORIENT>
static void launch_step_t(const DevParams* dP, const StepBuffers& b, int hp, bool do_kill,
                          bool want_wf, hipStream_t s) {
  hipLaunchKernelGGL((k_step<LAW, ORIENT>), dim3((b.n + 255) / 256), dim3(256), 0, s, dP, b.n,
                     b.cells, b.pos_srt, b.vel_srt, b.omg_srt, b.quat_srt, b.pos_cur, b.vel_cur,
                     b.omg_cur, b.quat_cur, b.keys[0], b.hcount[hp], b.hrec[hp], b.hcount[hp ^ 1],
                     b.hrec[hp ^ 1], b.wcount[hp], b.wrec[hp], b.wcount[hp ^ 1], b.wrec[hp ^ 1],
                     b.hist_cap, want_wf ? b.wall_force : nullptr, do_kill ? 1 : 0, b.status,
                     b.kill_count);
}

void launch_step(const DevParams* dP, const StepBuffers& b, int law, bool orient, int hp,
                 bool do_kill, bool want_wf, hipStream_t s) {
  if (law == BDEM_LAW_REFERENCE) {
    if (orient) launch_step_t<BDEM_LAW_REFERENCE, true>(dP, b, hp, do_kill, want_wf, s);
    else launch_step_t<BDEM_LAW_REFERENCE, false>(dP, b, hp, do_kill, want_wf, s);
  } else if (law == BDEM_LAW_LSD_COULOMB) {
    if (orient) launch_step_t<BDEM_LAW_LSD_COULOMB, true>(dP, b, hp, do_kill, want_wf, s);
    else launch_step_t<BDEM_LAW_LSD_COULOMB, false>(dP, b, hp, do_kill, want_wf, s);
  } else {
    if (orient) launch_step_t<BDEM_LAW_HERTZ_MINDLIN, true>(dP, b, hp, do_kill, want_wf, s);
    else launch_step_t<BDEM_LAW_HERTZ_MINDLIN, false>(dP, b, hp, do_kill, want_wf, s);
  }
}

/* ------------------------------------------------------------------------ */
/* Debug stages                                                              */
/* ------------------------------------------------------------------------ */
/* a5: Kernel_BroadCollisionDetection_NonSymmetry output format: the first 32
 * contacts per particle, NN_NUM = full count. */
__global__ __launch_bounds__(256) void k_contact_list(const DevParams* __restrict__ Pp, int n,
                                                      const uint2* __restrict__ cells,
                                                      const float4* __restrict__ pos_r,
                                                      const float4* __restrict__ omg_t,
                                                      uint32_t* __restrict__ list32,
                                                      uint32_t* __restrict__ nn, uint32_t* status) {
  const DevParams& P = *Pp;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (packed_pid(omg_t[i].w) < 0) { nn[i] = 0u; return; }
  const float4 pr = pos_r[i];
  const int c = for_each_contact(P, cells, pos_r, i, xyz(pr), pr.w, [&](int j, float4, int k) {
    if (k < LIST_STRIDE) list32[(size_t)i * LIST_STRIDE + k] = (uint32_t)j;
  });
  nn[i] = (uint32_t)c;
  if (c > LIST_STRIDE) atomicOr(status, ST_LIST_OVERFLOW);
}

void launch_contact_list(const DevParams* dP, const StepBuffers& b, uint32_t* list32, uint32_t* nn,
                         hipStream_t s) {
  hipLaunchKernelGGL(k_contact_list, dim3((b.n + 255) / 256), dim3(256), 0, s, dP, b.n, b.cells,
                     b.pos_srt, b.omg_srt, list32, nn, b.status);
}

template <int LAW>
__global__ __launch_bounds__(256) void k_forces(
    const DevParams* __restrict__ Pp, int n, const uint2* __restrict__ cells,
    const float4* __restrict__ pos_in, const float4* __restrict__ vel_in,
    const float4* __restrict__ omg_in, const int* __restrict__ hc_in,
    const HistRec* __restrict__ hr_in, HistRec* __restrict__ hr_out,
    const int* __restrict__ wc_in, const HistRec* __restrict__ wr_in, HistRec* __restrict__ wr_out,
    int hist_cap, float* __restrict__ fpp, float* __restrict__ tpp, float* __restrict__ fw,
    float* __restrict__ tw) {
  const DevParams& P = *Pp;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 pr = pos_in[i];
  const float4 vm = vel_in[i];
  const float4 ot = omg_in[i];
  const int pid = packed_pid(ot.w);
  const int ta = packed_type(ot.w);
  f3 F = mk3(0.f, 0.f, 0.f), T = F, Fw = F, Tw = F;
  if (pid >= 0) {
    const f3 p = xyz(pr), v = xyz(vm), w = xyz(ot);
    const float r = pr.w, m = vm.w;
    F = cvt(P.gravity) * m;
    PairHistCtx hc{nullptr, 0, 0, nullptr, 0};
    WallHist wh{nullptr, 0, nullptr, 0};
    HistRec scratch_pair[1];
    HistRec scratch_wall[1];
    if (LAW != BDEM_LAW_REFERENCE) {
      if (pid < hist_cap) {
        hc.in = hr_in + (size_t)pid * HMAX; hc.in_cnt = hc_in[pid]; hc.out = hr_out + (size_t)pid * HMAX;
        wh.in = wr_in + (size_t)pid * WMAX; wh.in_cnt = wc_in[pid]; wh.out = wr_out + (size_t)pid * WMAX;
      } else {
        hc.out = scratch_pair; hc.out_cnt = HMAX;
        wh.out = scratch_wall; wh.out_cnt = WMAX;
      }
    }
    particle_pair_forces<LAW>(P, cells, pos_in, vel_in, omg_in, i, p, r, v, m, w, ta, F, T, &hc);
    if (!P.rotation) T = mk3(0.f, 0.f, 0.f);
    wall_forces(P, ta, p, r, v, w, m, Fw, Tw, &wh);
  }
  fpp[3 * i] = F.x; fpp[3 * i + 1] = F.y; fpp[3 * i + 2] = F.z;
  tpp[3 * i] = T.x; tpp[3 * i + 1] = T.y; tpp[3 * i + 2] = T.z;
  fw[3 * i] = Fw.x; fw[3 * i + 1] = Fw.y; fw[3 * i + 2] = Fw.z;
  tw[3 * i] = Tw.x; tw[3 * i + 1] = Tw.y; tw[3 * i + 2] = Tw.z;
}

void launch_forces(const DevParams* dP, const StepBuffers& b, int law, int hp, float* fpp, float* tpp,
                   float* fw, float* tw, hipStream_t s) {
  const dim3 g((b.n + 255) / 256), blk(256);
  /* springs are written to the idle half so the state's history is untouched */
  if (law == BDEM_LAW_REFERENCE)
    hipLaunchKernelGGL(k_forces<BDEM_LAW_REFERENCE>, g, blk, 0, s, dP, b.n, b.cells, b.pos_srt,
                       b.vel_srt, b.omg_srt, b.hcount[hp], b.hrec[hp], b.hrec[hp ^ 1], b.wcount[hp],
                       b.wrec[hp], b.wrec[hp ^ 1], b.hist_cap, fpp, tpp, fw, tw);
  else if (law == BDEM_LAW_LSD_COULOMB)
    hipLaunchKernelGGL(k_forces<BDEM_LAW_LSD_COULOMB>, g, blk, 0, s, dP, b.n, b.cells, b.pos_srt,
                       b.vel_srt, b.omg_srt, b.hcount[hp], b.hrec[hp], b.hrec[hp ^ 1], b.wcount[hp],
                       b.wrec[hp], b.wrec[hp ^ 1], b.hist_cap, fpp, tpp, fw, tw);
  else
    hipLaunchKernelGGL(k_forces<BDEM_LAW_HERTZ_MINDLIN>, g, blk, 0, s, dP, b.n, b.cells, b.pos_srt,
                       b.vel_srt, b.omg_srt, b.hcount[hp], b.hrec[hp], b.hrec[hp ^ 1], b.wcount[hp],
                       b.wrec[hp], b.wrec[hp ^ 1], b.hist_cap, fpp, tpp, fw, tw);
}

}  // namespace bdem
