/*
 * bdem_kernels.hip - gfx950 kernels of one DEM step.
 *
 *   k_hash           a1  cell key per particle (only when the state was set
 *                        from the host; otherwise the step kernel emits keys)
 *   (a2, the stable sort, lives in bdem_sort.hip)
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

/* Plain 16-byte copy: the calibration kernel for hardware byte counters. */
__global__ __launch_bounds__(256) void k_copy_u4(uint4* __restrict__ dst, const uint4* __restrict__ src,
                                                 size_t count) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < count; i += stride) dst[i] = src[i];
}

void launch_copy_u4(uint4* dst, const uint4* src, size_t count, hipStream_t s) {
  if (count == 0) return;
  size_t blocks = (count + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_copy_u4, dim3((unsigned)blocks), dim3(256), 0, s, dst, src, count);
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
                     b.omg_cur, b.keys_new, b.status);
}

/* Keys of many simulations sharing one sort: slot i belongs to the
 * simulation of its 64-slot tile, whose key_base offsets its keys. */
__global__ __launch_bounds__(256) void k_hash_batch(const DevParams* __restrict__ Pv,
                                                    const uint32_t* __restrict__ tile_sim, int n,
                                                    const float4* __restrict__ pos_r,
                                                    const float4* __restrict__ omg_t,
                                                    uint32_t* __restrict__ keys, uint32_t* status) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const DevParams& P = Pv[tile_sim[i >> 6]];
  keys[i] = P.key_base + key_of(P, xyz(pos_r[i]), packed_pid(omg_t[i].w), status);
}

void launch_hash_batch(const DevParams* Pv, const uint32_t* tile_sim, const StepBuffers& b, hipStream_t s) {
  hipLaunchKernelGGL(k_hash_batch, dim3((b.n + 255) / 256), dim3(256), 0, s, Pv, tile_sim, b.n, b.pos_cur,
                     b.omg_cur, b.keys_new, b.status);
}

/* ------------------------------------------------------------------------ */
/* Pair forces of one particle (all laws).                                    */
/* ------------------------------------------------------------------------ */
struct PairHistCtx {
  const HistRec* in; int in_cnt; int cursor;
  HistRec* out; int out_cnt;
};

/* Force and torque on particle A from contact partner j (all laws), and the
 * spring record of that contact.  The caller accumulates Fc and Tc. */
template <int LAW>
__device__ __forceinline__ void contact_force(const DevParams& P, int j, const float4* __restrict__ pos_r,
                                              const float4* __restrict__ vel_m,
                                              const float4* __restrict__ omg_t, f3 pa, float ra, f3 va,
                                              float ma, f3 wa, int ta, const HistRec* hin, int hin_cnt,
                                              int cursor, HistRec* hrec, f3& Fc, f3& Tc,
                                              float* tally = nullptr) {
  const float4 pb = pos_r[j];
  const float4 vb = vel_m[j];
  const float4 ob = omg_t[j];
  HistRec first = make_float4(0.f, 0.f, 0.f, 0.f);
  if (LAW != BDEM_LAW_REFERENCE && hin_cnt > 0) first = hin[spring_start(hin_cnt, cursor)];
  if (LAW == BDEM_LAW_REFERENCE) {
    pair_ref(P, pa, ra, va, ma, wa, ta, xyz(pb), pb.w, xyz(vb), vb.w, xyz(ob), packed_type(ob.w), Fc, Tc,
             tally);
  } else {
    const f3 rel = pa - xyz(pb);
    const float d = length(rel);
    const f3 n = rel * (1.0f / d);
    const float delta = (ra + pb.w) - d;
    const f3 la = (-1.0f * n) * ra;
    const f3 lb = n * pb.w;
    const f3 vc = (va + cross(wa, la)) - (xyz(vb) + cross(xyz(ob), lb));
    /* R* and m* from one division */
    const float sr = ra + pb.w, sm = ma + vb.w;
    const float q = 1.0f / (sr * sm);
    const float rstar = ((ra * pb.w) * sm) * q;
    const float mstar = ((ma * vb.w) * sr) * q;
    const int idb = packed_pid(ob.w);
    f3 xi = find_spring_first(hin, hin_cnt, idb, cursor, first);
    /* one type: a wave-uniform address, read by scalar loads that need not
     * wait for the partner's type word */
    float e = 0.0f;
    Fc = history_law(P, n, delta, vc, rstar, mstar, P.one_type ? P.law_pp[0][0] : P.law_pp[ta][packed_type(ob.w)],
                     xi, tally ? &e : nullptr);
    /* half the pair's dissipation from each side (the pair is met from both) */
    if (tally) atomicAdd(tally + TALLY_PP, (0.5f * e) * P.dt);
    Tc = cross(la, Fc);
    if (hrec) *hrec = make_float4(xi.x, xi.y, xi.z, __int_as_float(idb));
  }
}

/* Per-particle (uncompacted) loop over the contacts found by the scan. */
template <int LAW>
__device__ __forceinline__ void particle_contact(const DevParams& P, int j, const float4* __restrict__ pos_r,
                                                 const float4* __restrict__ vel_m,
                                                 const float4* __restrict__ omg_t, f3 pa, float ra, f3 va,
                                                 float ma, f3 wa, int ta, f3& F, f3& T, PairHistCtx* hc,
                                                 float* tally = nullptr) {
  f3 Fc, Tc;
  HistRec* rec = (LAW != BDEM_LAW_REFERENCE && hc->out_cnt < HMAX) ? hc->out + hc->out_cnt : nullptr;
  contact_force<LAW>(P, j, pos_r, vel_m, omg_t, pa, ra, va, ma, wa, ta, hc->in, hc->in_cnt,
                     hc->out_cnt, rec, Fc, Tc, tally);
  if (LAW != BDEM_LAW_REFERENCE && hc->out_cnt < HMAX) ++hc->out_cnt;
  F = F + Fc;
  if (P.rotation) T = T + Tc;
}

/* Pair forces of one particle in two phases: the candidate scan only
 * records contact partners (in the reference order) into this thread's
 * column of an LDS list, then the force loop walks the list.  A wave's force
 * body then runs max(contacts) times instead of max(candidates) times.
 * Particles with more than CLIST contacts (beyond the reference's 32-entry
 * list) finish the tail with a second scan, in the same order. */
constexpr int CLIST = 12;   /* LDS list entries per thread: 12 KiB + wave staging fits 6 blocks per CU */
static_assert(CLIST <= HMAX, "springs of the listed contacts are stored first");
constexpr int STEP_THREADS = 256;

template <int LAW>
__device__ __forceinline__ int particle_pair_forces(const DevParams& P, const uint2* __restrict__ cells,
                                                    const float4* __restrict__ pos_r,
                                                    const float4* __restrict__ vel_m,
                                                    const float4* __restrict__ omg_t, int i, f3 pa,
                                                    float ra, f3 va, float ma, f3 wa, int ta, f3& F,
                                                    f3& T, PairHistCtx* hc, uint32_t* clist) {
  const int tid = threadIdx.x;
  const int count = for_each_contact(P, cells, pos_r, i, pa, ra, [&](int j, float4, int k) {
    if (k < CLIST) clist[k * STEP_THREADS + tid] = (uint32_t)j;
  });
  const int nl = count < CLIST ? count : CLIST;
  for (int k = 0; k < nl; ++k)
    particle_contact<LAW>(P, (int)clist[k * STEP_THREADS + tid], pos_r, vel_m, omg_t, pa, ra, va, ma,
                          wa, ta, F, T, hc);
  if (count > CLIST) {
    for_each_contact(P, cells, pos_r, i, pa, ra, [&](int j, float4, int k) {
      if (k >= CLIST)
        particle_contact<LAW>(P, j, pos_r, vel_m, omg_t, pa, ra, va, ma, wa, ta, F, T, hc);
    });
  }
  return count;
}

/* Wave-compacted force loop.  The 64 lanes' listed contacts are numbered
 * 0..T-1 in (lane, contact) order; each pass computes 64 of them, one per
 * lane, into LDS, and every owner then adds its own results in contact
 * order -- the same additions in the same order as the per-particle loop,
 * with the wave's force body running ceil(T/64) times instead of
 * max(contacts) times.  A map from contact number to (owner lane, index)
 * replaces a search per contact.  The owner's own state (position, velocity,
 * spin, spring count) comes from the owner lane's registers by wave shuffle,
 * not by a second gather.  All lanes of the wave must call it (count 0 for
 * idle lanes). */
struct WaveScratch {
  float f[6][64];
  uint16_t own[64 * CLIST];   /* contact q -> owner lane | index << 6 */
};
constexpr int WAVE_LDS_BYTES =
    sizeof(WaveStage) > sizeof(WaveScratch) ? (int)sizeof(WaveStage) : (int)sizeof(WaveScratch);

template <int LAW, int BT = STEP_THREADS>
__device__ __forceinline__ void wave_pair_forces(const DevParams& P, const float4* __restrict__ pos_r,
                                                 const float4* __restrict__ vel_m,
                                                 const float4* __restrict__ omg_t,
                                                 const uint32_t* clist, int nl, int base_i,
                                                 float4 own_p, float4 own_v, float4 own_o, uint32_t own_hwc,
                                                 const HistRec* __restrict__ hr_in, HistRec* hr_out,
                                                 int hist_cap, WaveScratch& ws, f3& F, f3& T,
                                                 float* tally = nullptr) {
  const int lane = threadIdx.x & 63;
  const int wbase_tid = threadIdx.x & ~63;
  int incl = nl;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int v = __shfl_up(incl, o, 64);
    if (lane >= o) incl += v;
  }
  const int excl = incl - nl;
  const int total = __shfl(incl, 63, 64);
  for (int k = 0; k < nl; ++k) ws.own[excl + k] = (uint16_t)(lane | (k << 6));
  __builtin_amdgcn_wave_barrier();
  for (int base = 0; base < total; base += 64) {
    const int p = base + lane;
    f3 Fc = mk3(0.f, 0.f, 0.f), Tc = Fc;
    /* every lane shuffles (the source lane must be active); lanes past the
     * last contact read their own registers and discard them */
    const uint32_t e = p < total ? (uint32_t)ws.own[p] : (uint32_t)lane;
    const int src = (int)(e & 63u);
    const float4 po = make_float4(__shfl(own_p.x, src, 64), __shfl(own_p.y, src, 64), __shfl(own_p.z, src, 64),
                                  __shfl(own_p.w, src, 64));
    const float4 vo = make_float4(__shfl(own_v.x, src, 64), __shfl(own_v.y, src, 64), __shfl(own_v.z, src, 64),
                                  __shfl(own_v.w, src, 64));
    const float4 oo = make_float4(__shfl(own_o.x, src, 64), __shfl(own_o.y, src, 64), __shfl(own_o.z, src, 64),
                                  __shfl(own_o.w, src, 64));
    const uint32_t ohwc = (uint32_t)__shfl((int)own_hwc, src, 64);
    if (p < total) {
      const int k = (int)(e >> 6);
      const int otid = wbase_tid + src;
      const int j = (int)clist[k * BT + otid];
      const int opid = packed_pid(oo.w);
      const HistRec* hin = nullptr;
      int hin_cnt = 0;
      HistRec* rec = nullptr;
      if (LAW != BDEM_LAW_REFERENCE && opid >= 0 && opid < hist_cap) {
        hin = hr_in + (size_t)opid * HMAX;
        hin_cnt = min((int)(ohwc & 0xFFFFu), HMAX);
        if (k < HMAX) rec = hr_out + (size_t)opid * HMAX + k;
      }
      contact_force<LAW>(P, j, pos_r, vel_m, omg_t, xyz(po), po.w, xyz(vo), vo.w, xyz(oo),
                         packed_type(oo.w), hin, hin_cnt, k, rec, Fc, Tc, tally);
    }
    ws.f[0][lane] = Fc.x; ws.f[1][lane] = Fc.y; ws.f[2][lane] = Fc.z;
    ws.f[3][lane] = Tc.x; ws.f[4][lane] = Tc.y; ws.f[5][lane] = Tc.z;
    __builtin_amdgcn_wave_barrier();
    const int a = max(excl, base), b = min(excl + nl, base + 64);
    for (int q = a; q < b; ++q) {
      const int s = q - base;
      F = F + mk3(ws.f[0][s], ws.f[1][s], ws.f[2][s]);
      if (P.rotation) T = T + mk3(ws.f[3][s], ws.f[4][s], ws.f[5][s]);
    }
    __builtin_amdgcn_wave_barrier();
  }
}

/* ------------------------------------------------------------------------ */
/* Fused step: contacts -> pair forces (+gravity) -> walls -> integrate ->    */
/* discharge check -> next key.  Reads the sorted scratch, writes the state.  */
/* ------------------------------------------------------------------------ */

#define BDEM_STEP_PARAMS                                                                            \
  int n, const float4* __restrict__ pos_in, const float4* __restrict__ vel_in,                      \
      const float4* __restrict__ omg_in, const float4* __restrict__ quat_in,                        \
      float4* __restrict__ pos_out, float4* __restrict__ vel_out, float4* __restrict__ omg_out,     \
      float4* __restrict__ quat_out, uint32_t* __restrict__ keys_out,                               \
      const uint32_t* __restrict__ sorted_keys, uint8_t* __restrict__ changed,                      \
      unsigned long long* __restrict__ sbuf, uint32_t* __restrict__ append,                         \
      const uint32_t* __restrict__ hwc_in, uint32_t* __restrict__ hwc_out,                          \
      const HistRec* __restrict__ hr_in, HistRec* __restrict__ hr_out,                              \
      const HistRec* __restrict__ wr_in, HistRec* __restrict__ wr_out, int hist_cap,                \
      float* __restrict__ wall_force, int do_kill, uint32_t *status, int *kill_count
#define BDEM_STEP_ARGS                                                                               \
  n, pos_in, vel_in, omg_in, quat_in, pos_out, vel_out, omg_out, quat_out, keys_out, sorted_keys, changed, \
      sbuf, append, hwc_in, hwc_out, hr_in, hr_out, wr_in, wr_out, hist_cap, wall_force, do_kill, status,  \
      kill_count

/* One wave's 64 consecutive sorted slots (lane i & 63 of the wave).  P and
 * cells are the slot range's simulation: its parameter block and its cell
 * table (indexed by the simulation's own hash); key_base is added to every
 * key written, so that simulations sharing one sort keep disjoint key ranges
 * (0 for a lone simulation).  clist is the block's contact list, wl this
 * wave's staging area.  Every lane of the wave must call it.  The streams
 * are restrict parameters (not struct members) so that the inlined body
 * keeps their no-alias guarantee. */
template <int LAW, bool ORIENT, bool SKIN, bool ENERGY, int BT = STEP_THREADS>
__device__ __forceinline__ void step_wave(const DevParams& P, const uint2* __restrict__ cells, uint32_t key_base,
                                          int i, uint32_t* clist, unsigned char* wl, const SkinArgs& sk,
                                          uint32_t* __restrict__ dbg_list, uint32_t* __restrict__ dbg_nn,
                                          BDEM_STEP_PARAMS) {
  float* const tally = ENERGY ? P.tally : nullptr;   /* dissipation tallies (EnergyCalc) */
  const int base_i = i - (int)(threadIdx.x & 63);
  const bool in = i < n;
  float4 pr = make_float4(0.f, 0.f, 0.f, 0.f), vm = pr, ot = pr;
  uint32_t hwc = 0u;
  if (in) {
    pr = pos_in[i]; vm = vel_in[i]; ot = omg_in[i];
    if (hwc_in) hwc = hwc_in[i];
  }
  int pid = in ? packed_pid(ot.w) : -1;
  const int ta = packed_type(ot.w);
  const bool alive = in && pid >= 0;
  f3 p = xyz(pr), v = xyz(vm), w = xyz(ot);
  const float r = pr.w, m = vm.w;

  /* phase 1: contact partners in the reference order, into this thread's list */
  auto record = [&](int j, float4, int k) {
    if (k < CLIST) clist[k * BT + threadIdx.x] = (uint32_t)j;
    if (dbg_list && k < LIST_STRIDE) dbg_list[(size_t)i * LIST_STRIDE + k] = (uint32_t)j;
  };
  int nc = 0;
  uint64_t new_mask = 0ull;
  if (SKIN) {
    /* skin masks: rows flagged (or all, at an epoch start) are tested in
     * full and re-masked, the others only at their masked candidates */
    const bool full = sk.vmeta[VM_FULL] != 0u;
    uint32_t flags = 0x1FFu;
    if (in && !full) {
      flags = sk.rowflag[i];
      if (flags) sk.rowflag[i] = 0u;
    }
    uint64_t m0 = in ? sk.mask_in[i] : 0ull;
#pragma unroll
    for (int rr = 0; rr < 9; ++rr)
      if ((flags >> rr) & 1u) m0 &= ~(0x7Full << (7 * rr));
    MaskedRows rows{m0, flags, sk.skin};
    nc = for_each_contact_wave_rows(P, cells, pos_in, (uint32_t)n, i, p, r, alive,
                                    *reinterpret_cast<WaveStage*>(wl), record, rows);
    new_mask = rows.mask;
  } else {
    nc = for_each_contact_wave(P, cells, pos_in, (uint32_t)n, i, p, r, alive,
                               *reinterpret_cast<WaveStage*>(wl), record);
  }
  if (dbg_nn && in) dbg_nn[i] = (uint32_t)nc;
  __builtin_amdgcn_wave_barrier();
  /* phase 2: gravity first (DEM_Kernels_Compute.cuh:654), then the contacts */
  f3 Fpp = alive ? cvt(P.gravity) * m : mk3(0.f, 0.f, 0.f);
  f3 Tpp = mk3(0.f, 0.f, 0.f);
  const int nl = nc < CLIST ? nc : CLIST;
  wave_pair_forces<LAW, BT>(P, pos_in, vel_in, omg_in, clist, nl, base_i, pr, vm, ot, hwc, hr_in, hr_out, hist_cap,
                        *reinterpret_cast<WaveScratch*>(wl), Fpp, Tpp, tally);

  bool chg = false;
  float dmove = 0.0f;
  if (alive) {
    const bool hist_slot = LAW != BDEM_LAW_REFERENCE && pid < hist_cap;
    if (nc > CLIST) {
      /* beyond the list: rescan, same order (rare).  Springs of contacts
       * CLIST .. HMAX-1 are stored after the list's, as in the restatement */
      PairHistCtx hc{nullptr, 0, 0, nullptr, HMAX};   /* no slot: looked up, not stored */
      if (hist_slot) {
        hc.in = hr_in + (size_t)pid * HMAX;
        hc.in_cnt = min((int)(hwc & 0xFFFFu), HMAX);
        hc.out = hr_out + (size_t)pid * HMAX;
        hc.out_cnt = CLIST;
      }
      for_each_contact(P, cells, pos_in, i, p, r, [&](int j, float4, int k) {
        if (k >= CLIST)
          particle_contact<LAW>(P, j, pos_in, vel_in, omg_in, p, r, v, m, w, ta, Fpp, Tpp, &hc, tally);
        if (dbg_list && k >= CLIST && k < LIST_STRIDE) dbg_list[(size_t)i * LIST_STRIDE + k] = (uint32_t)j;
      });
      /* the reference's 32-entry list overflows (Collision_Detection_BroadPhase.cuh:481) */
      if (nc > LIST_STRIDE) atomicOr(status, ST_LIST_OVERFLOW);
    }
    if (LAW != BDEM_LAW_REFERENCE && nc > HMAX) atomicOr(status, ST_HIST_OVERFLOW);
    if (!P.rotation) Tpp = mk3(0.f, 0.f, 0.f);

    WallHist wh{nullptr, 0, nullptr, 0};
    if (LAW != BDEM_LAW_REFERENCE) {
      if (hist_slot) {
        wh.in = wr_in + (size_t)pid * WMAX;
        wh.in_cnt = min((int)((hwc >> 16) & 0xFFu), WMAX);
        wh.out = wr_out + (size_t)pid * WMAX;
      } else {
        wh.out = nullptr; wh.out_cnt = WMAX;   /* no slot: springs are not kept (full count: never written) */
      }
    }
    f3 Fw, Tw;
    wall_forces(P, ta, p, r, v, w, m, Fw, Tw, &wh, tally);
    if (wall_force) {
      wall_force[3 * i] = Fw.x; wall_force[3 * i + 1] = Fw.y; wall_force[3 * i + 2] = Fw.z;
    }
    /* a9: lifters; the contact type of the last check is kept for the next step */
    f3 Fl = mk3(0.f, 0.f, 0.f), Tl = Fl;
    uint32_t ltype = 0u;
    if (P.num_dobj > 0)
      ltype = (uint32_t)lifter_forces(P, ta, p, r, v, w, m, (int)((hwc >> 24) & 3u), Fl, Tl, tally);
    if (hwc_out) {
      uint32_t h = ltype << 24;
      if (LAW != BDEM_LAW_REFERENCE && hist_slot)
        h |= (uint32_t)(nc < HMAX ? nc : HMAX) | ((uint32_t)wh.out_cnt << 16);
      hwc_out[i] = h;
    }

    const float inv_inertia = P.per_particle ? 1.0f / (0.4f * r * r * m) : P.t_inertia[ta];
    QuatF q{1.f, 0.f, 0.f, 0.f};
    if (ORIENT) { const float4 q4 = quat_in[i]; q = QuatF{q4.x, q4.y, q4.z, q4.w}; }
    integrate(P, inv_inertia, m, Fpp, Tpp, Fw, Tw, Fl, Tl, p, v, w, ORIENT ? &q : nullptr);
    if (SKIN) {
      /* |p_new - p_old| <= sum |v dt| (1 + 2^-20) + sum |p| 2^-22 (rounding of p += v dt) */
      dmove = ((fabsf(v.x) + fabsf(v.y) + fabsf(v.z)) * P.dt) * 1.00000095367431640625f +
              (fabsf(p.x) + fabsf(p.y) + fabsf(p.z)) * 2.384185791015625e-07f;
    }

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
    if (SKIN) sk.mask_out[i] = new_mask;
  } else if (in) {
    /* dead: inert, copied through */
    pos_out[i] = pr;
    vel_out[i] = vm;
    omg_out[i] = ot;
    if (ORIENT) quat_out[i] = quat_in[i];
    if (SKIN) sk.mask_out[i] = 0ull;
    if (hwc_out) hwc_out[i] = hwc;
    if (wall_force) { wall_force[3 * i] = 0.f; wall_force[3 * i + 1] = 0.f; wall_force[3 * i + 2] = 0.f; }
  }
  uint32_t new_key = 0u;
  if (in) {
    new_key = key_base + key_of(P, p, pid, status);
    keys_out[i] = new_key;
    chg = new_key != sorted_keys[i];
    changed[i] = chg ? 1 : 0;
  }
  if (SKIN) {
    /* the step's largest displacement bound, for the next step's epoch test */
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) dmove = fmaxf(dmove, __shfl_xor(dmove, o, 64));
    if ((threadIdx.x & 63) == 0 && dmove > 0.0f) atomicMax(&sk.vmeta[VM_DMAX + sk.par], __float_as_uint(dmove));
  }
  /* append changed slots for the next incremental sort (one atomic per wave) */
  const uint64_t mask = __ballot(chg);
  if (mask) {
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((unsigned long long)mask) - 1;
    uint32_t base = 0u;
    if (lane == leader) base = atomicAdd(append, (uint32_t)__popcll(mask));
    base = __shfl(base, leader, 64);
    const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    const uint32_t at = base + (uint32_t)__popcll(mask & lt);
    if (chg) {
      if (at < (uint32_t)n) sbuf[at] = ((unsigned long long)new_key << 32) | (uint32_t)i;
      else atomicOr(status, ST_SORT_ERROR);   /* a counter left uncleared: never write past the list */
    }
  }
}

/* One simulation: block b (renumbered XCD-aware) owns sorted slots
 * [256 b, 256 b + 256). */
template <int LAW, bool ORIENT, bool SKIN, bool ENERGY>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void k_step(
    SkinArgs sk, uint32_t* __restrict__ dbg_list, uint32_t* __restrict__ dbg_nn,
    const DevParams* __restrict__ Pp, const uint2* __restrict__ cells, BDEM_STEP_PARAMS) {
  __shared__ uint32_t clist[CLIST * STEP_THREADS];
  /* per wave: candidate staging in phase 1, pair results in phase 2 */
  __shared__ __align__(16) unsigned char wlds[STEP_THREADS / 64][WAVE_LDS_BYTES];
  const int i = xcd_block(blockIdx.x, gridDim.x) * STEP_THREADS + threadIdx.x;
  step_wave<LAW, ORIENT, SKIN, ENERGY>(*Pp, cells, 0u, i, clist, wlds[threadIdx.x >> 6], sk, dbg_list, dbg_nn,
                                       BDEM_STEP_ARGS);
}

/* Many independent simulations in one launch.  Simulation s owns the sorted
 * slots [64 t, 64 t + 64) of every tile t with tile_sim[t] == s (its slot
 * range is padded to whole tiles with dead slots), its own parameter block
 * Pv[s] and the cell table cells + Pv[s].key_base.
 *
 * PERSIST: a persistent grid deals the tiles from NUM_XCD queues, queue x
 * holding the x-th contiguous eighth of the tiles.  A wave first drains the
 * queue of the XCD its workgroup was dispatched to (workgroup w runs on XCD
 * w % 8), so neighbouring tiles, which read each other's cells and
 * particles, share an L2; then it helps the other queues in turn.  Each grab
 * takes `grab` consecutive tiles.  Counters queue[8 par .. 8 par + 7] serve
 * this step; the other set is cleared here for the next step (both start at
 * zero).  Otherwise each wave of the grid takes one tile, the workgroups
 * renumbered XCD-aware as in k_step. */
template <int LAW, int BT, bool PERSIST>
__device__ __forceinline__ void batch_body(const DevParams* __restrict__ Pv, const uint32_t* __restrict__ tile_sim,
                                           uint32_t ntiles, uint32_t grab, uint32_t* __restrict__ queue, int par,
                                           const uint2* __restrict__ cells, uint32_t* clist,
                                           unsigned char (*wlds)[WAVE_LDS_BYTES], BDEM_STEP_PARAMS) {
  const int lane = threadIdx.x & 63;
  const SkinArgs none{};
  if (!PERSIST) {
    const uint32_t tile = (uint32_t)xcd_block(blockIdx.x, gridDim.x) * (BT / 64) + (threadIdx.x >> 6);
    if (tile >= ntiles) return;   /* whole waves */
    const DevParams& P = Pv[__builtin_amdgcn_readfirstlane(tile_sim[tile])];
    step_wave<LAW, false, false, false, BT>(P, cells + P.key_base, P.key_base, (int)(tile * 64u) + lane, clist,
                                            wlds[threadIdx.x >> 6], none, nullptr, nullptr, BDEM_STEP_ARGS);
    return;
  }
  if (blockIdx.x == 0 && threadIdx.x < NUM_XCD) queue[NUM_XCD * (par ^ 1) + threadIdx.x] = 0u;
  uint32_t* const q = queue + NUM_XCD * par;
  const int home = (int)(blockIdx.x % NUM_XCD);
  for (int hop = 0; hop < NUM_XCD;) {
    const int x = (home + hop) % NUM_XCD;
    const uint32_t lo = (uint32_t)(((uint64_t)ntiles * (uint32_t)x) / NUM_XCD);
    const uint32_t hi = (uint32_t)(((uint64_t)ntiles * (uint32_t)(x + 1)) / NUM_XCD);
    uint32_t t = 0u;
    if (lane == 0) t = atomicAdd(q + x, grab);
    t = __builtin_amdgcn_readfirstlane(t);
    if (t >= hi - lo) { ++hop; continue; }
    const uint32_t tend = min(lo + t + grab, hi);
    for (uint32_t tile = lo + t; tile < tend; ++tile) {
      const DevParams& P = Pv[__builtin_amdgcn_readfirstlane(tile_sim[tile])];
      step_wave<LAW, false, false, false, BT>(P, cells + P.key_base, P.key_base, (int)(tile * 64u) + lane, clist,
                                              wlds[threadIdx.x >> 6], none, nullptr, nullptr, BDEM_STEP_ARGS);
    }
  }
}

/* The persistent loop keeps more values live across tiles: it is held to
 * four waves per SIMD (128 VGPRs, no spills) where the one-tile-per-wave
 * grid runs at six. */
template <int LAW, int BT, bool PERSIST>
__global__ __launch_bounds__(BT) __attribute__((amdgpu_waves_per_eu(PERSIST ? 4 : 6))) void k_step_batch(
    const DevParams* __restrict__ Pv, const uint32_t* __restrict__ tile_sim, uint32_t ntiles, uint32_t grab,
    uint32_t* __restrict__ queue, int par, const uint2* __restrict__ cells, BDEM_STEP_PARAMS) {
  __shared__ uint32_t clist[CLIST * BT];
  __shared__ __align__(16) unsigned char wlds[BT / 64][WAVE_LDS_BYTES];
  batch_body<LAW, BT, PERSIST>(Pv, tile_sim, ntiles, grab, queue, par, cells, clist, wlds, BDEM_STEP_ARGS);
}

template <int LAW, int BT, bool PERSIST>
static void launch_batch_t(const BatchLaunch& L, const StepBuffers& b, const uint32_t* sk, int hp, int ap,
                           hipStream_t s) {
  const unsigned blocks = PERSIST ? (unsigned)L.blocks : (L.ntiles + (BT / 64) - 1) / (BT / 64);
  hipLaunchKernelGGL((k_step_batch<LAW, BT, PERSIST>), dim3(blocks), dim3(BT), 0, s, L.params, L.tile_sim, L.ntiles,
                     L.grab, L.queue, L.par, L.cells, b.n, b.pos_srt, b.vel_srt, b.omg_srt, nullptr, b.pos_cur,
                     b.vel_cur, b.omg_cur, nullptr, b.keys_new, sk, b.changed, b.sbuf, b.meta + META_APPEND + ap,
                     b.hwc_srt, b.hwc_cur, b.hrec[hp], b.hrec[hp ^ 1], b.wrec[hp], b.wrec[hp ^ 1], b.hist_cap,
                     nullptr, 0, b.status, b.kill_count);
}

template <int LAW>
static void launch_batch_l(const BatchLaunch& L, const StepBuffers& b, const uint32_t* sk, int hp, int ap,
                           hipStream_t s) {
  if (L.blocks > 0) {
    if (L.threads == 64) launch_batch_t<LAW, 64, true>(L, b, sk, hp, ap, s);
    else if (L.threads == 128) launch_batch_t<LAW, 128, true>(L, b, sk, hp, ap, s);
    else launch_batch_t<LAW, 256, true>(L, b, sk, hp, ap, s);
  } else {
    if (L.threads == 64) launch_batch_t<LAW, 64, false>(L, b, sk, hp, ap, s);
    else if (L.threads == 128) launch_batch_t<LAW, 128, false>(L, b, sk, hp, ap, s);
    else launch_batch_t<LAW, 256, false>(L, b, sk, hp, ap, s);
  }
}

void launch_step_batch(const BatchLaunch& L, const StepBuffers& b, const uint32_t* sorted_keys, int hp, int ap,
                       hipStream_t s) {
  if (L.law == BDEM_LAW_REFERENCE) launch_batch_l<BDEM_LAW_REFERENCE>(L, b, sorted_keys, hp, ap, s);
  else if (L.law == BDEM_LAW_LSD_COULOMB) launch_batch_l<BDEM_LAW_LSD_COULOMB>(L, b, sorted_keys, hp, ap, s);
  else launch_batch_l<BDEM_LAW_HERTZ_MINDLIN>(L, b, sorted_keys, hp, ap, s);
}

template <int LAW, bool ORIENT, bool SKIN, bool ENERGY>
static void launch_step_t(const DevParams* dP, const StepBuffers& b, const uint32_t* sk, int hp, int ap,
                          bool do_kill, bool want_wf, const SkinArgs& sa, uint32_t* dbg_list, uint32_t* dbg_nn,
                          hipStream_t s) {
  hipLaunchKernelGGL((k_step<LAW, ORIENT, SKIN, ENERGY>), dim3((b.n + 255) / 256), dim3(256), 0, s, sa, dbg_list, dbg_nn,
                     dP, b.cells, b.n, b.pos_srt, b.vel_srt, b.omg_srt, b.quat_srt, b.pos_cur, b.vel_cur,
                     b.omg_cur, b.quat_cur, b.keys_new, sk, b.changed, b.sbuf, b.meta + META_APPEND + ap,
                     b.hwc_srt, b.hwc_cur, b.hrec[hp], b.hrec[hp ^ 1], b.wrec[hp], b.wrec[hp ^ 1],
                     b.hist_cap, want_wf ? b.wall_force : nullptr, do_kill ? 1 : 0, b.status, b.kill_count);
}

template <int LAW, bool ORIENT>
static void launch_step_s(const DevParams* dP, const StepBuffers& b, const uint32_t* sk, int hp, int ap,
                          bool do_kill, bool want_wf, bool energy, const SkinArgs& sa, uint32_t* dl, uint32_t* dn,
                          hipStream_t s) {
  if (energy) {   /* dissipation tallies (EnergyCalc) */
    if (sa.on) launch_step_t<LAW, ORIENT, true, true>(dP, b, sk, hp, ap, do_kill, want_wf, sa, dl, dn, s);
    else launch_step_t<LAW, ORIENT, false, true>(dP, b, sk, hp, ap, do_kill, want_wf, sa, dl, dn, s);
    return;
  }
  if (sa.on) launch_step_t<LAW, ORIENT, true, false>(dP, b, sk, hp, ap, do_kill, want_wf, sa, dl, dn, s);
  else launch_step_t<LAW, ORIENT, false, false>(dP, b, sk, hp, ap, do_kill, want_wf, sa, dl, dn, s);
}

void launch_step(const DevParams* dP, const StepBuffers& b, const uint32_t* sk, int law, bool orient,
                 int hp, int ap, bool do_kill, bool want_wf, bool energy, const SkinArgs& sa, uint32_t* dl,
                 uint32_t* dn, hipStream_t s) {
  if (law == BDEM_LAW_REFERENCE) {
    if (orient) launch_step_s<BDEM_LAW_REFERENCE, true>(dP, b, sk, hp, ap, do_kill, want_wf, energy, sa, dl, dn, s);
    else launch_step_s<BDEM_LAW_REFERENCE, false>(dP, b, sk, hp, ap, do_kill, want_wf, energy, sa, dl, dn, s);
  } else if (law == BDEM_LAW_LSD_COULOMB) {
    if (orient) launch_step_s<BDEM_LAW_LSD_COULOMB, true>(dP, b, sk, hp, ap, do_kill, want_wf, energy, sa, dl, dn, s);
    else launch_step_s<BDEM_LAW_LSD_COULOMB, false>(dP, b, sk, hp, ap, do_kill, want_wf, energy, sa, dl, dn, s);
  } else {
    if (orient) launch_step_s<BDEM_LAW_HERTZ_MINDLIN, true>(dP, b, sk, hp, ap, do_kill, want_wf, energy, sa, dl, dn, s);
    else launch_step_s<BDEM_LAW_HERTZ_MINDLIN, false>(dP, b, sk, hp, ap, do_kill, want_wf, energy, sa, dl, dn, s);
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
    const float4* __restrict__ omg_in, const uint32_t* __restrict__ hwc_in,
    const HistRec* __restrict__ hr_in, HistRec* __restrict__ hr_out,
    const HistRec* __restrict__ wr_in, HistRec* __restrict__ wr_out,
    int hist_cap, float* __restrict__ fpp, float* __restrict__ tpp, float* __restrict__ fw,
    float* __restrict__ tw) {
  const DevParams& P = *Pp;
  __shared__ uint32_t clist[CLIST * STEP_THREADS];
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
    if (LAW != BDEM_LAW_REFERENCE) {
      if (pid < hist_cap) {
        const uint32_t hwc = hwc_in[i];
        hc.in = hr_in + (size_t)pid * HMAX; hc.in_cnt = min((int)(hwc & 0xFFFFu), HMAX);
        hc.out = hr_out + (size_t)pid * HMAX;
        wh.in = wr_in + (size_t)pid * WMAX; wh.in_cnt = min((int)(hwc >> 16), WMAX);
        wh.out = wr_out + (size_t)pid * WMAX;
      } else {
        hc.out = nullptr; hc.out_cnt = HMAX;   /* no slot: full counts, never written */
        wh.out = nullptr; wh.out_cnt = WMAX;
      }
    }
    particle_pair_forces<LAW>(P, cells, pos_in, vel_in, omg_in, i, p, r, v, m, w, ta, F, T, &hc, clist);
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
                       b.vel_srt, b.omg_srt, b.hwc_srt, b.hrec[hp], b.hrec[hp ^ 1], b.wrec[hp],
                       b.wrec[hp ^ 1], b.hist_cap, fpp, tpp, fw, tw);
  else if (law == BDEM_LAW_LSD_COULOMB)
    hipLaunchKernelGGL(k_forces<BDEM_LAW_LSD_COULOMB>, g, blk, 0, s, dP, b.n, b.cells, b.pos_srt,
                       b.vel_srt, b.omg_srt, b.hwc_srt, b.hrec[hp], b.hrec[hp ^ 1], b.wrec[hp],
                       b.wrec[hp ^ 1], b.hist_cap, fpp, tpp, fw, tw);
  else
    hipLaunchKernelGGL(k_forces<BDEM_LAW_HERTZ_MINDLIN>, g, blk, 0, s, dP, b.n, b.cells, b.pos_srt,
                       b.vel_srt, b.omg_srt, b.hwc_srt, b.hrec[hp], b.hrec[hp ^ 1], b.wrec[hp],
                       b.wrec[hp ^ 1], b.hist_cap, fpp, tpp, fw, tw);
}

}  // namespace bdem
