/*
 * bdem_device.hpp - device-side data layout, parameters and contact physics
 * for the gfx950 DEM step.
 *
 * Every floating-point expression here is written so that, compiled with
 * -ffp-contract=off and IEEE division/sqrt, it performs exactly the
 * operations of the reference arithmetic in the same order (the reference
 * lines are cited per function).  No FMA contraction, no reciprocal tricks:
 * the sphere path is memory-bound, so exactness costs nothing measurable.
 *
 * Particle state is SoA, three float4 streams in sorted (cell) order:
 *   pos_r = (x, y, z, radius)     vel_m = (vx, vy, vz, mass)
 *   omg_t = (wx, wy, wz, bits(P_ID * 8 + type))
 * plus an optional float4 orientation (w, x, y, z).
 */
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "bdem_types.h"

namespace bdem {

constexpr int LIST_STRIDE = 32;   /* Broad_List stride, Collision_Detection_BroadPhase.cuh:481 */
constexpr int HMAX = 16;          /* pair tangential-history records per particle */
constexpr int WMAX = 4;           /* wall tangential-history records per particle */
constexpr int MAX_TYPES = 8;      /* __constant__ POBJ ParticleObject[8], GPU_Memory_Transfer.cuh:81 */
constexpr int MAX_WOBJ = 8;       /* __constant__ WOBJ WorldObject[8], GPU_Memory_Transfer.cuh:80 */
constexpr int MAX_DOBJ = 32;      /* __constant__ DOBJ DynamicObject[32], GPU_Memory_Transfer.cuh:82 */

/* Status bits raised by kernels into the context's status word. */
enum : uint32_t {
  ST_OUT_OF_GRID   = 1u,   /* a live particle left the hash grid */
  ST_LIST_OVERFLOW = 2u,   /* more than 32 contacts (reference list capacity) */
  ST_HIST_OVERFLOW = 4u,   /* more contacts than history slots */
  ST_SORT_ERROR    = 8u,   /* incremental sort produced an out-of-range position */
};

/* MI355X dispatches workgroup w to XCD w % 8, and each XCD has its own L2.
 * Renumbering so that XCD x runs one contiguous eighth of the logical blocks
 * keeps the neighbour rows a block reads (one y-layer of cells either side,
 * about 19 blocks of 256 away in sorted order) in the same L2 that loaded
 * them for the blocks next to it.  A bijection on [0, nb) for any nb. */
constexpr int NUM_XCD = 8;
__device__ __forceinline__ int xcd_block(int w, int nb) {
  const int x = w % NUM_XCD, q = w / NUM_XCD, per = nb / NUM_XCD, rem = nb % NUM_XCD;
  return x * per + (x < rem ? x : rem) + q;
}

struct f3 { float x, y, z; };
__device__ __forceinline__ f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ f3 operator*(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ f3 operator*(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ f3 operator/(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
/* Utilities/Device/Functions.cuh:198-201 */
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
/* Utilities/Device/Functions.cuh:370-373 */
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
  return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
/* Utilities/Device/Functions.cuh:241-244 */
__device__ __forceinline__ float length(f3 a) { return __builtin_sqrtf(dot(a, a)); }
__device__ __forceinline__ f3 cvt(bdem_float3 v) { return f3{v.x, v.y, v.z}; }
__device__ __forceinline__ f3 xyz(float4 v) { return f3{v.x, v.y, v.z}; }

__device__ __forceinline__ int packed_pid(float w) { return __float_as_int(w) >> 3; }
__device__ __forceinline__ int packed_type(float w) { return __float_as_int(w) & 7; }
__host__ __device__ __forceinline__ int pack_id_type(int pid, int type) { return pid * 8 + (type & 7); }

/* Per type-pair constants of the history laws (derived on the host). */
struct LawPair { float kn, kt, beta, mu, estar, gstar, tg, pad1; };   /* tg = sqrt(4 G* / E*) */

/* Tangential spring record: (xi.x, xi.y, xi.z, bits(partner id)). */
typedef float4 HistRec;

/* Everything a kernel reads besides particle streams; lives in device memory
 * (one copy per context), read through uniform scalar loads. */
struct DevParams {
  /* grid (SimulationInfo worldOrigin/cellSize/num_NNCells) */
  bdem_float3 origin;
  bdem_float3 cell;
  int nx, ny, nz;
  uint32_t C;                 /* nx*ny*nz */
  uint32_t key_base;          /* added to every key written: simulations sharing one sort (0 alone) */
  /* time stepping */
  float dt;
  bdem_float3 gravity;
  float roll_res;
  int rotation, mode2d, sphere_orient;
  int poly;                   /* particle_type 1: polyhedra */
  int law;                    /* BDEM_LAW_* */
  int per_particle;           /* 1: inverse inertia from per-particle radius/mass */
  /* walls */
  int num_wobj;               /* Num_WorldObjects (Hatch decrements it) */
  int cyl_rotation;           /* drum_rotation() flag */
  int rotating_wobj;          /* SimulationInfo::Rotating_WObject */
  float mill_rot_vel;
  int num_cones;
  /* discharge */
  float kill_height;
  bdem_float3 max_size;
  int kill_enabled;
  /* type tables (POBJ fields the sphere path reads) */
  float t_radius[MAX_TYPES], t_mass[MAX_TYPES], t_inertia[MAX_TYPES];
  float pp_kn[MAX_TYPES][MAX_TYPES], pp_cn[MAX_TYPES][MAX_TYPES], pp_fk[MAX_TYPES][MAX_TYPES];
  float s_kn[MAX_TYPES], s_cn[MAX_TYPES], s_fs[MAX_TYPES], s_fk[MAX_TYPES];
  LawPair law_pp[MAX_TYPES][MAX_TYPES];
  int one_type;               /* one particle type: every pair uses law_pp[0][0] */
  float vbound_scale;         /* polyhedra: largest |q|^2 (>= 1) of the last orientation upload */
  int shape_by_pid;           /* polyhedra: shape-table rows are P_IDs (else types) */
  LawPair law_w[MAX_TYPES];
  bdem_cone cone[BDEM_MAX_CONES];
  bdem_WOBJ wobj[MAX_WOBJ];
  /* lifters (VolumeObject_InteractionParticle) */
  int num_dobj;
  int mill_sim;               /* SimulationInfo::isMillSim */
  bdem_float3 mill_cent;      /* SimulationInfo::cent */
  float dobj_bound_r;         /* SimulationInfo::DObject_CylinderBoundR */
  float* tally;               /* {particle, lifter, surface} dissipation tallies (device) */
  float d_kn[MAX_TYPES], d_cn[MAX_TYPES], d_fk[MAX_TYPES];   /* POBJ Dsurface_* */
  bdem_DOBJ dobj[MAX_DOBJ];
};

/* ------------------------------------------------------------------------ */
/* a1: grid position + hash                                                 */
/* SpatialDecomp_CalcGridPos, Collision_Detection_BroadPhase.cuh:20-29       */
/* SpatialDecomp_CalcGridHash, Collision_Detection_BroadPhase.cuh:42-50      */
/* ------------------------------------------------------------------------ */
__device__ __forceinline__ void grid_pos(const DevParams& P, f3 p, int g[3]) {
  g[0] = (int)floorf((p.x - P.origin.x) / P.cell.x);
  g[1] = (int)floorf((p.y - P.origin.y) / P.cell.y);
  g[2] = (int)floorf((p.z - P.origin.z) / P.cell.z);
}
__device__ __forceinline__ uint32_t hash_of(const DevParams& P, int gx, int gy, int gz) {
  return ((uint32_t)gx + (uint32_t)gz * (uint32_t)P.nz) + ((uint32_t)gy * (uint32_t)P.nx) * (uint32_t)P.nz;
}
/* Key of a particle: reference hash when in the grid, C for a live particle
 * outside it, C+1 for a dead one.  Both sentinels sort behind every cell. */
__device__ __forceinline__ uint32_t key_of(const DevParams& P, f3 p, int pid, uint32_t* status) {
  if (pid < 0) return P.C + 1u;
  int g[3];
  grid_pos(P, p, g);
  if (g[0] < 0 || g[1] < 0 || g[2] < 0 || g[0] >= P.nx || g[1] >= P.ny || g[2] >= P.nz) {
    atomicOr(status, ST_OUT_OF_GRID);
    return P.C;
  }
  uint32_t h = hash_of(P, g[0], g[1], g[2]);
  return h < P.C ? h : P.C;
}

/* ------------------------------------------------------------------------ */
/* a5 traversal: 27-cell stencil in the reference order (y outer, z, x      */
/* inner, ascending sorted index inside a cell), clamped to the grid.        */
/* Collision_Detection_BroadPhase.cuh:428-489, 369-383.                      */
/* The three x-neighbours of a (y,z) row have consecutive hashes, so their   */
/* particles are one contiguous run of the sorted arrays: a row is scanned   */
/* as a single index range.                                                  */
/* ------------------------------------------------------------------------ */
/* The reference's overlap test, length(pA-pB) - (rA+rB) < 0
 * (Collision_Detection_BroadPhase.cuh:376), decided without a square root
 * whenever the answer is certain.  With d2 = fl(dot(d,d)) (the same value
 * the reference takes the root of) and s = fl(rA+rB):
 *   d2 < fl(fl(s*s)*(1-2^-20))  =>  sqrt(d2) < s(1-2^-22), at least two ulps
 *                                   below s, so fl(sqrt(d2)) < s: overlap;
 *   d2 > fl(fl(s*s)*(1+2^-20))  =>  sqrt(d2) > s(1+2^-22), so
 *                                   fl(sqrt(d2)) >= s: no overlap.
 * Only the band in between evaluates the reference expression itself, so
 * the decision is bit-identical to it. */
__device__ __forceinline__ bool overlaps(f3 pa, float ra, float4 pb) {
  const f3 d = pa - xyz(pb);
  const float d2 = dot(d, d);
  const float s = ra + pb.w;
  const float s2 = s * s;
  bool hit = d2 < s2 * 0.99999904632568359375f;                          /* 1 - 2^-20 */
  const bool band = !hit && !(d2 > s2 * 1.00000095367431640625f);        /* 1 + 2^-20 */
  /* a branch on the rare band only, so the root is skipped by whole waves */
  if (__builtin_expect(band, 0)) hit = __builtin_sqrtf(d2) - s < 0.0f;
  return hit;
}

/* The contact predicate of the reference (exact, above) and the membership
 * predicate of the skin lists: |pa - pb| < ra + rb + skin, decided on the
 * safe side (every pair that close is included, a few slightly farther may
 * be: their contact is still tested exactly each step).  With
 * d2 = fl(dot(d,d)) <= D^2 (1 + 2^-22) and s = fl(fl(ra+rb) + skin) >=
 * S (1 - 2^-23), comparing against fl(s*s) (1 + 2^-19) includes every D < S. */
struct OverlapTest {
  __device__ __forceinline__ bool operator()(f3 pa, float ra, float4 pb) const { return overlaps(pa, ra, pb); }
};
struct SkinTest {
  float skin;
  __device__ __forceinline__ bool operator()(f3 pa, float ra, float4 pb) const {
    const f3 d = pa - xyz(pb);
    const float d2 = dot(d, d);
    const float s = (ra + pb.w) + skin;
    return d2 < (s * s) * 1.0000019073486328125f;   /* 1 + 2^-19 */
  }
};

template <class Visit, class Pred = OverlapTest>
__device__ __forceinline__ int for_each_contact(const DevParams& P, const uint2* __restrict__ cells,
                                                const float4* __restrict__ pos_r, int i, f3 pa,
                                                float ra, Visit&& visit, Pred pred = Pred()) {
  int g[3];
  grid_pos(P, pa, g);
  const int xs = max(g[0] - 1, 0), xe = min(g[0] + 1, P.nx - 1);
  const int ys = max(g[1] - 1, 0), ye = min(g[1] + 1, P.ny - 1);
  const int zs = max(g[2] - 1, 0), ze = min(g[2] + 1, P.nz - 1);
  int count = 0;
  for (int y = ys; y <= ye; ++y) {
    for (int z = zs; z <= ze; ++z) {
      if (xs > xe) continue;
      const uint32_t h0 = hash_of(P, xs, y, z);
      const uint32_t h1 = hash_of(P, xe, y, z);
      if (h1 >= P.C) continue;
      uint32_t lo = 0xFFFFFFFFu, hi = 0u;
      for (uint32_t h = h0; h <= h1; ++h) {
        const uint2 c = cells[h];
        if (c.x != 0xFFFFFFFFu) {
          lo = min(lo, c.x);
          hi = c.y;
        }
      }
      if (lo == 0xFFFFFFFFu) continue;
      for (uint32_t j = lo; j < hi; ++j) {
        if ((int)j == i) continue;
        const float4 pb = pos_r[j];
        if (pred(pa, ra, pb)) {
          visit((int)j, pb, count);
          ++count;
        }
      }
    }
  }
  return count;
}

/* Wave-cooperative form of the same traversal.  The 64 lanes are 64
 * consecutive sorted particles, so their cells lie in a short key range
 * [Hmin, Hmax].  For a stencil row offset `off` (dy, dz) every lane's three
 * cells are inside [Hmin + off - 1, Hmax + off + 1]: the wave loads that run
 * of cell entries once (coalesced, no per-row reduction), takes the first
 * and last occupied entry by ballot, and loads the contiguous run of
 * candidates between them into LDS.  Each lane then scans its own cells and
 * candidates from LDS in the reference order.  A wave whose key range is too
 * wide, or a row with too many candidates, reads global memory instead; a
 * live particle outside the grid (key C) takes the per-lane traversal.  Every
 * lane of the wave must call it; `active` false for lanes without a live
 * particle.  Same visits, same order as for_each_contact. */
constexpr int STAGE_POS = 128;
constexpr int STAGE_CELL = 128;
struct WaveStage {
  float4 pos[STAGE_POS];
  uint2 cell[STAGE_CELL];
};

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, o, 64));
  return v;
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o, 64));
  return v;
}

/* Wave window of cell entries for stencil row r = 3 (dy + 1) + (dz + 1):
 * [hmin + off - 1, hmax + off + 1] clipped to the table (uniform). */
struct RowWin { uint32_t w0, wn; bool ok; };
__device__ __forceinline__ RowWin row_window(const DevParams& P, uint32_t hmin, uint32_t hmax, int r) {
  const int dy = r / 3 - 1, dz = r % 3 - 1;
  const long long off = (long long)dz * P.nz + (long long)dy * P.nx * P.nz;
  const long long wl = max((long long)hmin + off - 1, 0ll);
  const long long wh = min((long long)hmax + off + 1, (long long)P.C - 1);
  RowWin w;
  w.ok = r < 9 && wl <= wh;
  w.w0 = w.ok ? (uint32_t)wl : 0u;
  w.wn = w.ok ? (uint32_t)(wh - wl) + 1u : 0u;
  return w;
}

/* Candidate run of a row window from its two cell entries per lane: the
 * start of the first occupied entry and the end of the last (by ballot). */
struct RowStage { uint32_t pmin, pmax; bool any, staged; };
__device__ __forceinline__ RowStage row_stage(uint2 e0, uint2 e1, uint32_t n) {
  RowStage s{0u, 0u, false, false};
  const uint64_t m0 = __ballot(e0.x != 0xFFFFFFFFu);
  const uint64_t m1 = __ballot(e1.x != 0xFFFFFFFFu);
  if ((m0 | m1) == 0ull) return s;
  const int f = m0 ? __ffsll((unsigned long long)m0) - 1 : 64 + __ffsll((unsigned long long)m1) - 1;
  const int l = m1 ? 127 - __clzll((long long)m1) : 63 - __clzll((long long)m0);
  s.pmin = f < 64 ? (uint32_t)__builtin_amdgcn_readlane((int)e0.x, f)
                  : (uint32_t)__builtin_amdgcn_readlane((int)e1.x, f - 64);
  s.pmax = l < 64 ? (uint32_t)__builtin_amdgcn_readlane((int)e0.y, l)
                  : (uint32_t)__builtin_amdgcn_readlane((int)e1.y, l - 64);
  s.pmax = min(s.pmax, n);   /* table entries never exceed n; kept in range regardless */
  s.any = s.pmin < s.pmax;
  s.staged = s.any && s.pmax - s.pmin <= (uint32_t)STAGE_POS;
  return s;
}

/* Candidates [lo, hi) of one lane, read from `src` (LDS stage or global). */
template <class Visit, class Pred>
__device__ __forceinline__ void scan_candidates(const float4* src, uint32_t lo, uint32_t hi, uint32_t shift,
                                                int i, f3 pa, float ra, int& count, Visit& visit,
                                                const Pred& pred) {
  for (uint32_t j = lo; j < hi; ++j) {
    const float4 pb = src[j - shift];
    if ((int)j != i && pred(pa, ra, pb)) {
      visit((int)j, pb, count);
      ++count;
    }
  }
}

/* How one lane scans the candidate run [lo, hi) of stencil row r
 * (r = 3 (dy + 1) + (dz + 1)): every candidate with a predicate. */
template <class Pred>
struct PlainRows {
  Pred pred;
  template <class Visit>
  __device__ __forceinline__ void row(const float4* src, uint32_t lo, uint32_t hi, uint32_t shift, int, int i,
                                      f3 pa, float ra, int& count, Visit& visit) {
    scan_candidates(src, lo, hi, shift, i, pa, ra, count, visit, pred);
  }
  __device__ __forceinline__ Pred lane_pred() { return pred; }
};

/* Skin masks.  Bits 7r .. 7r+6 of `mask` mark which of the first seven
 * candidates lo .. lo+6 of row r were within ra + rb + skin when the row was
 * last tested in full; candidates beyond the seventh are always tested.
 * While no particle enters or leaves the row's three cells the row's run
 * holds the same particles in the same order (the re-sort is stable), so
 * the bits keep their meaning however the run shifts in the sorted arrays;
 * and while the epoch's displacement bound holds (twice it below the skin)
 * no unmarked candidate can have come into contact.  Rows in `full_rows`
 * are tested in full and their bits rebuilt.  Contacts are visited in the
 * run's order either way. */
struct MaskedRows {
  uint64_t mask;
  uint32_t full_rows;
  float skin;
  template <class Visit>
  __device__ __forceinline__ void row(const float4* src, uint32_t lo, uint32_t hi, uint32_t shift, int r, int i,
                                      f3 pa, float ra, int& count, Visit& visit) {
    const uint32_t sh = 7u * (uint32_t)r;
    if ((full_rows >> r) & 1u) {
      const SkinTest near{skin};
      uint64_t nb = 0ull;
      for (uint32_t j = lo; j < hi; ++j) {
        const float4 pb = src[j - shift];
        if ((int)j == i) continue;
        if (j - lo < 7u && near(pa, ra, pb)) nb |= 1ull << (j - lo);
        if (overlaps(pa, ra, pb)) {
          visit((int)j, pb, count);
          ++count;
        }
      }
      mask |= nb << sh;   /* the row's field was cleared when it was marked full */
    } else {
      uint32_t bits = (uint32_t)(mask >> sh) & 0x7Fu;
      while (bits) {
        const uint32_t j = lo + (uint32_t)__builtin_ctz(bits);
        bits &= bits - 1u;
        if (j >= hi || (int)j == i) continue;
        const float4 pb = src[j - shift];
        if (overlaps(pa, ra, pb)) {
          visit((int)j, pb, count);
          ++count;
        }
      }
      for (uint32_t j = lo + 7u; j < hi; ++j) {
        const float4 pb = src[j - shift];
        if ((int)j != i && overlaps(pa, ra, pb)) {
          visit((int)j, pb, count);
          ++count;
        }
      }
    }
  }
  /* a lane scanned on its own (outside the grid): no mask */
  __device__ __forceinline__ OverlapTest lane_pred() {
    mask = 0ull;
    return OverlapTest();
  }
};

template <class Visit, class Rows>
__device__ __forceinline__ int for_each_contact_wave_rows(const DevParams& P, const uint2* __restrict__ cells,
                                                          const float4* __restrict__ pos_r, uint32_t n, int i,
                                                          f3 pa, float ra, bool active, WaveStage& st,
                                                          Visit& visit, Rows& rows) {
  const int lane = threadIdx.x & 63;
  int g[3];
  grid_pos(P, pa, g);
  const bool ingrid = active && g[0] >= 0 && g[1] >= 0 && g[2] >= 0 && g[0] < P.nx && g[1] < P.ny &&
                      g[2] < P.nz && hash_of(P, g[0], g[1], g[2]) < P.C;
  const uint32_t H = ingrid ? hash_of(P, g[0], g[1], g[2]) : 0u;
  const uint32_t hmin = wave_min_u32(ingrid ? H : 0xFFFFFFFFu);
  const uint32_t hmax = wave_max_u32(ingrid ? H : 0u);
  const int xs = max(g[0] - 1, 0), xe = min(g[0] + 1, P.nx - 1);
  int count = 0;
  if (hmin <= hmax) {   /* some lane is in the grid (wave-uniform) */
    if (hmax - hmin + 3u <= (uint32_t)STAGE_CELL) {
      /* Software pipeline over the nine rows r = 3 (dy + 1) + (dz + 1): while
       * row r is scanned from LDS, the candidates of row r + 1 and the cell
       * entries of row r + 2 are already in flight. */
      const uint2 empty_cell = make_uint2(0xFFFFFFFFu, 0u);
      RowWin wc = row_window(P, hmin, hmax, 0), w1 = row_window(P, hmin, hmax, 1);
      /* (plain ifs: a select between a load and a constant struct would go
       * through a private copy and a flat load) */
      uint2 ce0 = empty_cell, ce1 = empty_cell, ne0 = empty_cell, ne1 = empty_cell;
      if (wc.ok && (uint32_t)lane < wc.wn) ce0 = cells[wc.w0 + lane];
      if (wc.ok && (uint32_t)lane + 64u < wc.wn) ce1 = cells[wc.w0 + 64u + lane];
      if (w1.ok && (uint32_t)lane < w1.wn) ne0 = cells[w1.w0 + lane];
      if (w1.ok && (uint32_t)lane + 64u < w1.wn) ne1 = cells[w1.w0 + 64u + lane];
      RowStage sc = row_stage(ce0, ce1, n);
      float4 q0 = make_float4(0.f, 0.f, 0.f, 0.f), q1 = q0;
      if (sc.staged) {
        if ((uint32_t)lane < sc.pmax - sc.pmin) q0 = pos_r[sc.pmin + lane];
        if ((uint32_t)lane + 64u < sc.pmax - sc.pmin) q1 = pos_r[sc.pmin + 64u + lane];
      }
      for (int r = 0; r < 9; ++r) {
        /* 1. row r into LDS */
        st.cell[lane] = ce0;
        st.cell[lane + 64] = ce1;
        if (sc.staged) {
          st.pos[lane] = q0;
          st.pos[lane + 64] = q1;
        }
        __builtin_amdgcn_wave_barrier();
        /* 2. issue row r + 1's candidates and row r + 2's cells */
        const RowStage sn = row_stage(ne0, ne1, n);
        float4 nq0 = make_float4(0.f, 0.f, 0.f, 0.f), nq1 = nq0;
        if (r + 1 < 9 && sn.staged) {
          if ((uint32_t)lane < sn.pmax - sn.pmin) nq0 = pos_r[sn.pmin + lane];
          if ((uint32_t)lane + 64u < sn.pmax - sn.pmin) nq1 = pos_r[sn.pmin + 64u + lane];
        }
        const RowWin w2 = row_window(P, hmin, hmax, r + 2);
        uint2 nn0 = empty_cell, nn1 = empty_cell;
        if (r + 2 < 9 && w2.ok) {
          if ((uint32_t)lane < w2.wn) nn0 = cells[w2.w0 + lane];
          if ((uint32_t)lane + 64u < w2.wn) nn1 = cells[w2.w0 + 64u + lane];
        }
        /* 3. scan row r: each lane's cells, then its candidates */
        if (wc.ok && sc.any) {
          const int y = g[1] + (r / 3 - 1), z = g[2] + (r % 3 - 1);
          const bool rv0 = ingrid && y >= 0 && y < P.ny && z >= 0 && z < P.nz && xs <= xe;
          const uint32_t h0 = rv0 ? hash_of(P, xs, y, z) : 0u;
          const uint32_t h1 = rv0 ? hash_of(P, xe, y, z) : 0u;
          uint32_t lo = sc.pmin, hi = sc.pmin;
          if (rv0 && h1 < P.C) {
            uint32_t l2 = 0xFFFFFFFFu, h2 = 0u;
            for (uint32_t h = h0; h <= h1; ++h) {
              const uint2 c = st.cell[h - wc.w0];
              if (c.x != 0xFFFFFFFFu) {
                l2 = min(l2, c.x);
                h2 = c.y;
              }
            }
            if (l2 != 0xFFFFFFFFu) { lo = l2; hi = min(h2, n); }
          }
          /* a consistent cell table puts every lane's run inside the stage;
           * checked, so LDS is never indexed outside it whatever the table holds */
          if (sc.staged && lo >= sc.pmin && hi <= sc.pmax)
            rows.row(st.pos, lo, hi, sc.pmin, r, i, pa, ra, count, visit);
          else
            rows.row(pos_r, lo, hi, 0u, r, i, pa, ra, count, visit);
        }
        __builtin_amdgcn_wave_barrier();
        ce0 = ne0; ce1 = ne1; ne0 = nn0; ne1 = nn1;
        sc = sn; q0 = nq0; q1 = nq1;
        wc = w1; w1 = w2;
      }
    } else {
      /* keys too spread for one stage: per-lane cells, shared candidates when they fit */
      for (int dy = -1; dy <= 1; ++dy) {
        for (int dz = -1; dz <= 1; ++dz) {
          const int y = g[1] + dy, z = g[2] + dz;
          const bool rv0 = ingrid && y >= 0 && y < P.ny && z >= 0 && z < P.nz && xs <= xe;
          const uint32_t h0 = rv0 ? hash_of(P, xs, y, z) : 0u;
          const uint32_t h1 = rv0 ? hash_of(P, xe, y, z) : 0u;
          const bool rv = rv0 && h1 < P.C;
          uint32_t lo = 0xFFFFFFFFu, hi = 0u;
          if (rv) {
            for (uint32_t h = h0; h <= h1; ++h) {
              const uint2 c = cells[h];
              if (c.x != 0xFFFFFFFFu) {
                lo = min(lo, c.x);
                hi = c.y;
              }
            }
          }
          const bool ne = lo != 0xFFFFFFFFu;
          const uint32_t pmin = wave_min_u32(ne ? lo : 0xFFFFFFFFu);
          const uint32_t pmax = min(wave_max_u32(ne ? hi : 0u), n);
          if (pmin >= pmax) continue;
          if (!ne) { lo = pmin; hi = pmin; }
          hi = min(hi, n);
          if (pmax - pmin <= (uint32_t)STAGE_POS) {
            for (uint32_t c = lane; c < pmax - pmin; c += 64) st.pos[c] = pos_r[pmin + c];
            __builtin_amdgcn_wave_barrier();
            if (lo >= pmin && hi <= pmax) rows.row(st.pos, lo, hi, pmin, (dy + 1) * 3 + (dz + 1), i, pa, ra, count, visit);
            else rows.row(pos_r, lo, hi, 0u, (dy + 1) * 3 + (dz + 1), i, pa, ra, count, visit);
          } else {
            rows.row(pos_r, lo, hi, 0u, (dy + 1) * 3 + (dz + 1), i, pa, ra, count, visit);
          }
          __builtin_amdgcn_wave_barrier();
        }
      }
    }
  }
  /* a live particle outside the grid (key C): the per-lane traversal */
  if (active && !ingrid) count = for_each_contact(P, cells, pos_r, i, pa, ra, visit, rows.lane_pred());
  return count;
}

template <class Visit, class Pred = OverlapTest>
__device__ __forceinline__ int for_each_contact_wave(const DevParams& P, const uint2* __restrict__ cells,
                                                     const float4* __restrict__ pos_r, uint32_t n, int i,
                                                     f3 pa, float ra, bool active, WaveStage& st,
                                                     Visit&& visit, Pred pred = Pred()) {
  PlainRows<Pred> rows{pred};
  return for_each_contact_wave_rows(P, cells, pos_r, n, i, pa, ra, active, st, visit, rows);
}


/* ------------------------------------------------------------------------ */
/* a6: reference pair law on A from B.                                       */
/* Kernel_ParticleInteraction_Spheres_NonSymmetry DEM_Kernels_Compute.cuh:656-712 */
/* Collision_Response_Particle_Particle Collision_Response.cuh:374-494       */
/* ------------------------------------------------------------------------ */
/* Collision_Response_Particle_Particle (Collision_Response.cuh:374-494) for
 * normal N (towards A), depth dis and lever arms RA, RB from the centres. */
/* Energy-dissipation tallies {particle, lifter, surface}
 * (Tally_EnergyDis_PP / DisLift / DisSurf, GPU_Memory_Transfer.cuh:97-99):
 * each contact adds its increment on its own with a float atomic, as the
 * reference does.  tally == nullptr: not counted. */
enum { TALLY_PP = 0, TALLY_LIFT = 1, TALLY_SURF = 2 };

/* The wall and lifter increment (Collision_Response.cuh:25-89, 127-214,
 * 266-344): |Fdiss| |Vn| dt unless the force was clipped as attractive,
 * plus |Ft| |Vt| dt when there is tangential slip. */
__device__ __forceinline__ void wall_tally(const DevParams& P, float* tally, int k, bool attractive, f3 Fdiss,
                                           f3 Vn, f3 Tang, float mag) {
  if (!tally) return;
  float e = (!attractive && dot(Vn, Vn) > 0.0f) ? (length(Fdiss) * length(Vn)) * P.dt : 0.0f;
  if (mag > 0.0f) e = e + ((length(Tang) * mag) * P.dt);
  atomicAdd(tally + k, e);
}

__device__ __forceinline__ void pair_resp(const DevParams& P, f3 va, float ma, f3 wa, int ta, f3 vb,
                                          float mb, f3 wb, int tb, f3 N, float dis, f3 RA, f3 RB,
                                          f3& Fc, f3& Tc, float* tally = nullptr) {
  const f3 RVel = va - vb;
  const float Kn = P.pp_kn[ta][tb], Cn = P.pp_cn[ta][tb], Fk = P.pp_fk[ta][tb];
  const f3 Vnormal = dot(RVel, N) * N;
  const float Fnormal = Kn * dis;
  const f3 Fdiss = Cn * Vnormal;
  f3 Ft = Fnormal * N - Fdiss;
  const f3 V_CPA = va + cross(wa, RA);
  const f3 V_CPB = vb + cross(wb, RB);
  const f3 V_CPNA = dot(V_CPA, N) * N;
  const f3 V_CPNB = dot(V_CPB, N) * N;
  const f3 Vt_A = V_CPA - V_CPNA;
  const f3 Vt_B = V_CPB - V_CPNB;
  const f3 Vt = Vt_A - Vt_B;
  float mag = dot(Vt, Vt);
  f3 Tang = mk3(0.f, 0.f, 0.f);
  if (mag > 0.0f) {
    mag = __builtin_sqrtf(mag);
    float Velf = dot(RVel, RVel);
    Velf = Velf > 0.0f ? __builtin_sqrtf(Velf) : 1.0f;
    const float tan_f = (((0.50f * (ma + mb)) / P.dt) * 0.050f) * Fk * (mag / Velf);
    Tang = (-1.0f * tan_f) * (Vt / mag);
  }
  if (tally) {
    /* half the normal work (each pair is visited from both sides), the whole
     * tangential work (:401-475) */
    float e = dot(Vnormal, Vnormal) > 0.0f ? ((0.50f * length(Fdiss)) * length(Vnormal)) * P.dt : 0.0f;
    if (mag > 0.0f) e = e + ((length(Tang) * mag) * P.dt);
    atomicAdd(tally + TALLY_PP, e);
  }
  Ft = Ft + Tang;
  Fc = Ft;                 /* the kernel adds it: force += Force_PairA.trans (:709) */
  Tc = cross(RA, Ft);      /* torque = cross(R_Body_COMA, Force.trans) (:485) */
}

__device__ __forceinline__ void pair_ref(const DevParams& P, f3 pa, float ra, f3 va, float ma,
                                         f3 wa, int ta, f3 pb, float rb, f3 vb, float mb, f3 wb,
                                         int tb, f3& Fc, f3& Tc, float* tally = nullptr) {
  const f3 relPos = pa - pb;
  const f3 N = relPos / length(relPos);
  const float dis = fabsf(length(relPos) - (ra + rb)) + 0.000010f;
  const f3 RA = (-1.0f * N) * ra;
  const f3 RB = N * rb;
  pair_resp(P, va, ma, wa, ta, vb, mb, wb, tb, N, dis, RA, RB, Fc, Tc, tally);
}

/* ------------------------------------------------------------------------ */
/* History laws (extensions, textbook forms; same arithmetic as the oracle) */
/* ------------------------------------------------------------------------ */
__device__ __forceinline__ f3 rotate_spring(const DevParams& P, f3 xi_old, f3 n, f3 vt) {
  f3 xi = xi_old - dot(xi_old, n) * n;
  const float m_old = dot(xi_old, xi_old);
  const float m_new = dot(xi, xi);
  if (m_new > 0.0f) xi = xi * __builtin_sqrtf(m_old / m_new);
  return xi + vt * P.dt;
}

/* diss (may be null): the contact's dissipation rate for the tallies, this
 * library's definition for the history laws (the reference tallies only
 * its own law): normal damping force x |vn| while the normal force pushes,
 * or the elastic force x |vn| while damping has clipped it to zero (that
 * spring energy is not returned), plus the tangential damping force x |vt|
 * while the spring sticks, or the capped tangential force x |vt| while it
 * slides. */
__device__ __forceinline__ f3 history_law(const DevParams& P, f3 n, float delta, f3 vc,
                                          float rstar, float mstar, const LawPair& lp, f3& xi,
                                          float* diss = nullptr) {
  const float vn = dot(vc, n);
  const f3 vt = vc - vn * n;
  xi = rotate_spring(P, xi, n, vt);
  float fn, kt, gt, dn, fe;
  if (P.law == BDEM_LAW_LSD_COULOMB) {
    const float cn = -2.0f * lp.beta * __builtin_sqrtf(lp.kn * mstar);
    fe = lp.kn * delta;
    fn = fe - cn * vn;
    dn = cn * vn;
    kt = lp.kt;
    gt = 0.0f;
  } else {
    const float srd = __builtin_sqrtf(rstar * delta);
    const float sn = 2.0f * lp.estar * srd;
    const float st = 8.0f * lp.gstar * srd;
    const float gn = -1.8257418583505538f * lp.beta * __builtin_sqrtf(sn * mstar);
    gt = gn * lp.tg;   /* sqrt(St m*) = tg sqrt(Sn m*) */
    fe = (4.0f / 3.0f) * lp.estar * srd * delta;
    fn = fe - gn * vn;
    dn = gn * vn;
    kt = st;
  }
  const bool pushing = fn > 0.0f;
  if (fn < 0.0f) fn = 0.0f;
  f3 ft = -1.0f * kt * xi - gt * vt;
  const float ft2 = dot(ft, ft);
  const float fc = lp.mu * fn;
  bool slip = false;
  if (ft2 > fc * fc) {
    const float s = fc / __builtin_sqrtf(ft2);
    ft = ft * s;
    xi = (-1.0f / kt) * ft;
    slip = true;
  }
  if (diss) {
    float e = (pushing ? fabsf(dn) : fabsf(fe)) * fabsf(vn);
    const float vt2 = dot(vt, vt);
    if (slip) e = e + __builtin_sqrtf(dot(ft, ft)) * __builtin_sqrtf(vt2);
    else e = e + gt * vt2;
    *diss = e;
  }
  return fn * n + ft;
}

/* ------------------------------------------------------------------------ */
/* a8: walls                                                                 */
/* ------------------------------------------------------------------------ */
struct Coll { bool hit; float dis; f3 cp; f3 n; };

/* is_PointOnSurface, Device/Utilities.cuh:20-53 */
__device__ __forceinline__ bool point_on_surface(const bdem_Surface& s, const bdem_float3* vtx, f3 point) {
  float area = 0.0f;
  const int nv = s.num_vertex;
  for (int i = 0; i < nv - 1; ++i) {
    const f3 V1 = point - cvt(vtx[s.vertex_Order[i]]);
    const f3 V2 = point - cvt(vtx[s.vertex_Order[i + 1]]);
    area = area + 0.500000f * length(cross(V1, V2));
  }
  const f3 V1 = point - cvt(vtx[s.vertex_Order[nv - 1]]);
  const f3 V2 = point - cvt(vtx[s.vertex_Order[0]]);
  area = area + 0.500000f * length(cross(V1, V2));
  return (fabsf(area - s.area) / s.area) * 100.0000f <= s.Tol;
}
__device__ __forceinline__ bool point_on_surface(const bdem_Surface& s, const bdem_WOBJ& w, f3 point) {
  return point_on_surface(s, w.vertex, point);
}

/* Collision_Detection_Sphere_Surface / _Plane, Collision_Detection_Spheres.cuh:6-100 */
__device__ __forceinline__ Coll sphere_plane(f3 p, float r, f3 nrm, f3 cen, const bdem_Surface* s,
                                             const bdem_WOBJ* w) {
  Coll c{false, 0.0f, mk3(0.f, 0.f, 0.f), mk3(0.f, 0.f, 0.f)};
  const f3 Pt = p - r * nrm;
  float dis = dot(nrm, Pt - cen);
  if (dis < 0.000000f && dis > -r) {
    dis = fabsf(dis);
    if (s) {
      const f3 point = Pt + nrm * dis;
      if (!point_on_surface(*s, *w, point)) return c;
    }
    c.hit = true; c.dis = dis; c.cp = Pt; c.n = nrm;
  }
  return c;
}

/* Collision_Detection_Sphere_Cylinder, Collision_Detection_Spheres.cuh:234-265 */
__device__ __forceinline__ Coll sphere_cylinder(f3 p, float r, const bdem_Macro_Cylinder& cyl) {
  Coll c{false, 0.0f, mk3(0.f, 0.f, 0.f), mk3(0.f, 0.f, 0.f)};
  const f3 Pvec = p * cvt(cyl.AXIS) - cvt(cyl.center_bot_cap);
  /* inside by more than r + 10 um: no contact, decided without the square
   * root (the margin exceeds the rounding of both forms by far) */
  const float inner = cyl.radius - r - 0.001f;
  if (inner > 0.0f && dot(Pvec, Pvec) < inner * inner) return c;
  const float Pd = length(Pvec) + r;
  if (Pd > cyl.radius) {
    f3 n = -1.0f * Pvec;
    n = n * (1.0f / length(n));
    c.n = n;
    c.cp = p - n * r;
    c.dis = fabsf(Pd - cyl.radius);
    c.hit = true;
  }
  return c;
}

/* Conical-frustum wall (extension): closest point of the meridian segment. */
__device__ __forceinline__ Coll sphere_cone(f3 p, float r, const bdem_cone& cn) {
  Coll c{false, 0.0f, mk3(0.f, 0.f, 0.f), mk3(0.f, 0.f, 0.f)};
  const f3 ax = cvt(cn.axis);
  const f3 d = p - cvt(cn.base);
  const float s = dot(d, ax);
  /* the distance to the meridian segment is at least the axial gap to its
   * span [min(s0, s1), max(s0, s1)]: beyond r (plus a rounding margin) there
   * is no contact */
  const float slack = r + 0.001f + 4.0e-6f * fabsf(s);
  if (s > fmaxf(cn.s0, cn.s1) + slack || s < fminf(cn.s0, cn.s1) - slack) return c;
  const f3 q = d - s * ax;
  const float rho = length(q);
  if (!(rho > 0.0f)) return c;
  const float ds = cn.s1 - cn.s0, dr = cn.r1 - cn.r0;
  float t = ((s - cn.s0) * ds + (rho - cn.r0) * dr) / (ds * ds + dr * dr);
  t = fminf(fmaxf(t, 0.0f), 1.0f);
  const float sc = cn.s0 + t * ds, rc = cn.r0 + t * dr;
  const float es = s - sc, er = rho - rc;
  const float dist = __builtin_sqrtf(es * es + er * er);
  if (dist < r && dist > 0.0f) {
    const f3 qh = q / rho;
    const f3 n = (es / dist) * ax + (er / dist) * qh;
    c.hit = true;
    c.dis = r - dist;
    c.n = n;
    c.cp = p - n * r;
  }
  return c;
}

/* Collision_Response_Particle_StaticSurface, Collision_Response.cuh:4-109 */
__device__ __forceinline__ void resp_static(const DevParams& P, int t, f3 p, f3 v, f3 w, float m,
                                            const Coll& c, f3& F, f3& T, float* tally = nullptr) {
  const f3 Rc = c.cp - p;
  const f3 Vnormal = dot(v, c.n) * c.n;
  const float Fnormal = P.s_kn[t] * c.dis;
  const f3 Fdiss = P.s_cn[t] * Vnormal;
  f3 Ftr = Fnormal * c.n - Fdiss;
  const f3 Fdir = Ftr / length(Ftr);
  const bool attractive = dot(Fdir, c.n) < 0.0f;
  if (attractive) Ftr = Fnormal * c.n;
  f3 Vt = v - Vnormal;
  if (P.rotation) Vt = Vt + cross(w, Rc);
  float mag = dot(Vt, Vt);
  f3 Tang = mk3(0.f, 0.f, 0.f);
  if (mag > 0.0f) {
    mag = __builtin_sqrtf(mag);
    float Velf = dot(v, v);
    Velf = Velf > 0.0f ? __builtin_sqrtf(Velf) : 1.0f;
    const float tan_f = (0.050f * (m / P.dt)) * (P.s_fs[t] * (mag / Velf));
    Tang = (-1.0f * tan_f) * (Vt / mag);
  }
  wall_tally(P, tally, TALLY_SURF, attractive, Fdiss, Vnormal, Tang, mag);
  Ftr = Ftr + Tang;
  F = F + Ftr;
  if (P.rotation) T = T + cross(Rc, Ftr);
}

/* Collision_Response_Particle_MovingSurface, Collision_Response.cuh:116-233 */
__device__ __forceinline__ f3 resp_moving(const DevParams& P, int t, f3 p, f3 v, f3 w, float m,
                                          const Coll& c, f3 Vs, f3& Tout, float* tally = nullptr) {
  const f3 Rc = c.cp - p;
  const f3 RVel = v - Vs;
  const f3 Vn = dot(RVel, c.n) * c.n;
  const float Fnormal = P.s_kn[t] * c.dis;
  const f3 Fdiss = P.s_cn[t] * Vn;
  f3 Ftr = Fnormal * c.n - Fdiss;
  const f3 Fdir = Ftr / length(Ftr);
  const bool attractive = dot(Fdir, c.n) < 0.0f;
  if (attractive) Ftr = Fnormal * c.n;
  const f3 V_CPA = v + cross(w, Rc);
  const f3 V_CPNA = dot(V_CPA, c.n) * c.n;
  const f3 Vt_A = V_CPA - V_CPNA;
  const f3 Vt_B = Vs - dot(Vs, c.n) * c.n;
  const f3 Vt = Vt_A - Vt_B;
  float mag = dot(Vt, Vt);
  f3 Tang = mk3(0.f, 0.f, 0.f);
  if (mag > 0.0f) {
    mag = __builtin_sqrtf(mag);
    float Velf = dot(RVel, RVel);
    Velf = Velf > 0.0f ? __builtin_sqrtf(Velf) : 1.0f;
    const float tan_f = P.s_fk[t] * (mag / Velf);
    Tang = ((-1.0f * tan_f) * (Vt / mag)) * (m / P.dt) * 0.050f;
  }
  wall_tally(P, tally, TALLY_SURF, attractive, Fdiss, Vn, Tang, mag);
  Ftr = Ftr + Tang;
  Tout = P.rotation ? cross(Rc, Ftr) : mk3(0.f, 0.f, 0.f);
  return Ftr;
}

/* Collision_Response_EndPlate_Particle, Collision_Response.cuh:504-526 */
__device__ __forceinline__ f3 resp_endplate(const DevParams& P, int t, f3 v, const Coll& c) {
  const f3 Vn = dot(v, c.n) * c.n;
  const f3 Fdiss = P.s_cn[t] * Vn;
  const float Fn = P.s_kn[t] * c.dis;
  return Fn * c.n - Fdiss;
}

/* Drum surface velocity, DEM_Kernels_Compute.cuh:375-378, 449-452 */
__device__ __forceinline__ f3 drum_velocity(const DevParams& P, f3 cp) {
  const float Rd = P.wobj[0].cyl_geo.radius;
  return cross(mk3(0.0f, 0.0f, -P.mill_rot_vel), cp - mk3(Rd, Rd, 0.0f));
}

/* Look up a spring by id among `cnt` records, starting at `cursor` (the
 * partner order is nearly unchanged from step to step, so the first probe
 * usually hits). */
__device__ __forceinline__ f3 find_spring(const HistRec* __restrict__ recs, int cnt, int id,
                                          int& cursor) {
  const int start = (cursor >= 0 && cursor < cnt) ? cursor : 0;
  for (int k = 0; k < cnt; ++k) {
    int s = start + k;
    if (s >= cnt) s -= cnt;
    const HistRec r = recs[s];
    if (__float_as_int(r.w) == id) {
      cursor = s + 1;
      return mk3(r.x, r.y, r.z);
    }
  }
  return mk3(0.f, 0.f, 0.f);
}

/* find_spring with the first probe, recs[start(cursor)], already loaded by
 * the caller (issued beside the partner's state loads); the same probes in
 * the same order otherwise. */
__device__ __forceinline__ int spring_start(int cnt, int cursor) {
  return (cursor >= 0 && cursor < cnt) ? cursor : 0;
}
__device__ __forceinline__ f3 find_spring_first(const HistRec* __restrict__ recs, int cnt, int id, int& cursor,
                                                HistRec first) {
  const int start = spring_start(cnt, cursor);
  for (int k = 0; k < cnt; ++k) {
    int s = start + k;
    if (s >= cnt) s -= cnt;
    const HistRec r = k == 0 ? first : recs[s];
    if (__float_as_int(r.w) == id) {
      cursor = s + 1;
      return mk3(r.x, r.y, r.z);
    }
  }
  return mk3(0.f, 0.f, 0.f);
}

/* Wall force and torque on one particle.  The reference law keeps the
 * accumulate/overwrite pattern of DEM_Kernels_Compute.cuh:283-604 (plane hits
 * add; cylinder and cap hits assign the force; torques add).  History laws
 * sum every wall contact and keep one spring per touched wall. */
struct WallHist {
  const HistRec* in; int in_cnt;
  HistRec* out; int out_cnt;
};

__device__ __forceinline__ void wall_contact_hist(const DevParams& P, int t, f3 p, float r, f3 v,
                                                  f3 w, float m, const Coll& c, f3 vw, int wall_id,
                                                  f3& F, f3& T, WallHist& wh, float* tally = nullptr) {
  const f3 la = c.cp - p;
  const f3 vc = (v + cross(w, la)) - vw;
  int cur = 0;
  f3 xi = find_spring(wh.in, wh.in_cnt, wall_id, cur);
  float e = 0.0f;
  const f3 f = history_law(P, c.n, c.dis, vc, r, m, P.law_w[t], xi, tally ? &e : nullptr);
  if (tally) atomicAdd(tally + TALLY_SURF, e * P.dt);
  F = F + f;
  if (P.rotation) T = T + cross(la, f);
  if (wh.out_cnt < WMAX) {
    wh.out[wh.out_cnt] = make_float4(xi.x, xi.y, xi.z, __int_as_float(wall_id));
    ++wh.out_cnt;
  }
}

__device__ __forceinline__ void wall_forces(const DevParams& P, int t, f3 p, float r, f3 v, f3 w,
                                            float m, f3& F, f3& T, WallHist* wh, float* tally = nullptr) {
  F = mk3(0.f, 0.f, 0.f);
  T = mk3(0.f, 0.f, 0.f);
  const bool hist = P.law != BDEM_LAW_REFERENCE;
  const bool moving_planes = P.rotating_wobj && P.cyl_rotation;
  for (int k = 0; k < P.num_wobj; ++k) {
    const bdem_WOBJ& W = P.wobj[k];
    if (W.surface_type == BDEM_SURFACE_PLANE) {
      for (int s = 0; s < W.num_surfaces; ++s) {
        const bdem_Surface& S = W.surfaces[s];
        const Coll c = (S.area == -1.0f) ? sphere_plane(p, r, cvt(S.normal), cvt(S.centroid), nullptr, nullptr)
                                         : sphere_plane(p, r, cvt(S.normal), cvt(S.centroid), &S, &W);
        if (!c.hit) continue;
        if (hist) {
          const f3 vw = moving_planes ? drum_velocity(P, c.cp) : mk3(0.f, 0.f, 0.f);
          wall_contact_hist(P, t, p, r, v, w, m, c, vw, k * 16 + s, F, T, *wh, tally);
        } else if (moving_planes) {
          f3 tq;
          F = F + resp_moving(P, t, p, v, w, m, c, drum_velocity(P, c.cp), tq, tally);
          T = T + tq;
        } else {
          resp_static(P, t, p, v, w, m, c, F, T, tally);
        }
      }
    } else if (W.surface_type == BDEM_SURFACE_CYLINDER) {
      const bdem_Macro_Cylinder& cyl = W.cyl_geo;
      const Coll c = sphere_cylinder(p, r, cyl);
      if (c.hit) {
        if (hist) {
          const f3 vw = P.cyl_rotation ? drum_velocity(P, c.cp) : mk3(0.f, 0.f, 0.f);
          wall_contact_hist(P, t, p, r, v, w, m, c, vw, 128 + k, F, T, *wh, tally);
        } else if (P.cyl_rotation) {
          f3 tq;
          F = resp_moving(P, t, p, v, w, m, c, drum_velocity(P, c.cp), tq, tally);
          T = T + tq;
        } else {
          f3 f = mk3(0.f, 0.f, 0.f), tq = mk3(0.f, 0.f, 0.f);
          resp_static(P, t, p, v, w, m, c, f, tq, tally);
          F = f;
          T = T + tq;
        }
      }
      if (cyl.has_top_cap) {
        const Coll cc = sphere_plane(p, r, cvt(cyl.normal_top_cap), cvt(cyl.center_top_cap), nullptr, nullptr);
        if (cc.hit) {
          if (hist) wall_contact_hist(P, t, p, r, v, w, m, cc, mk3(0.f, 0.f, 0.f), 144 + k, F, T, *wh, tally);
          else F = resp_endplate(P, t, v, cc);
        }
      }
      if (cyl.has_bot_cap) {
        const Coll cc = sphere_plane(p, r, cvt(cyl.normal_bot_cap), cvt(cyl.center_bot_cap), nullptr, nullptr);
        if (cc.hit) {
          if (hist) wall_contact_hist(P, t, p, r, v, w, m, cc, mk3(0.f, 0.f, 0.f), 160 + k, F, T, *wh, tally);
          else F = resp_endplate(P, t, v, cc);
        }
      }
    }
  }
  for (int k = 0; k < P.num_cones; ++k) {
    const Coll c = sphere_cone(p, r, P.cone[k]);
    if (!c.hit) continue;
    if (hist) wall_contact_hist(P, t, p, r, v, w, m, c, mk3(0.f, 0.f, 0.f), 176 + k, F, T, *wh, tally);
    else resp_static(P, t, p, v, w, m, c, F, T, tally);
  }
}

/* ------------------------------------------------------------------------ */
/* a10: Integrate_Euler_NonSymmetry (Integration.cuh:6-184), spheres.         */
/* ------------------------------------------------------------------------ */
struct QuatF { float w, x, y, z; };

/* Quaternion product, Utilities/Device/Opertators.cuh:866-874 */
__device__ __forceinline__ QuatF qmul(QuatF a, QuatF b) {
  QuatF q;
  q.x = (a.x * b.w) + (a.y * b.z) - (a.z * b.y) + (a.w * b.x);
  q.y = -(a.x * b.z) + (a.y * b.w) + (a.z * b.x) + (a.w * b.y);
  q.z = (a.x * b.y) - (a.y * b.x) + (a.z * b.w) + (a.w * b.z);
  q.w = -(a.x * b.x) - (a.y * b.y) - (a.z * b.z) + (a.w * b.w);
  return q;
}
/* normaliseD, Utilities/Device/Functions.cuh:399-420 */
__device__ __forceinline__ QuatF qnormalise(QuatF a) {
  const float s = a.w * a.w + a.x * a.x + a.y * a.y + a.z * a.z;
  const float n = s > 0.0f ? 1.0f / __builtin_sqrtf(s) : 0.0f;
  return QuatF{a.w * n, a.x * n, a.y * n, a.z * n};
}

/* ------------------------------------------------------------------------ */
/* a9: lifters, VolumeObject_InteractionParticle (DEM_Kernels_Compute.cuh:12-246) */
/* ------------------------------------------------------------------------ */
struct DColl { bool hit; int type; float dis; f3 n; f3 cp; };

/* Collision_Detection_Sphere_DObject, Collision_Detection_Spheres.cuh:108-225:
 * faces first unless the previous lifter contact was an edge, then the edges
 * as half-lines from their first vertex. */
__device__ __forceinline__ DColl sphere_dobject(const bdem_DOBJ& D, f3 p, float r, int prev_type) {
  DColl c{false, 0, 0.0f, mk3(0.f, 0.f, 0.f), mk3(0.f, 0.f, 0.f)};
  if (prev_type != 2) {
    for (int j = 0; j < D.num_faces; ++j) {
      const bdem_Surface& S = D.faces[j];
      const f3 nrm = cvt(S.normal);
      const f3 Pt = p - r * nrm;
      float dis = dot(nrm, Pt - cvt(S.centroid));
      if (dis < 0.000000f && dis > -r * 0.50f) {
        dis = fabsf(dis);
        const f3 point = Pt + nrm * dis;
        if (point_on_surface(S, D.vertex, point)) {
          c.hit = true; c.type = 1; c.dis = dis; c.cp = Pt; c.n = nrm;
          return c;
        }
      }
    }
  }
  for (int e = 0; e < D.num_edges; ++e) {
    const f3 P0 = cvt(D.vertex[D.edge[e].point_vindex[0]]);
    const f3 dir = cvt(D.vertex[D.edge[e].point_vindex[1]]) - P0;
    const f3 Edir = dir / length(dir);
    const f3 vpc = p - P0;
    const float dVec = dot(Edir, vpc);
    if (dVec > 0.0f) {
      const f3 P2 = P0 + dVec * Edir;
      const f3 d2 = p - P2;
      const float d = length(d2);
      if (d < r) {
        const float dis = fabsf(d - r);
        if (dis < r) {
          f3 nn = p - P2;
          nn = nn * (1.0f / length(nn));
          c.hit = true; c.type = 2; c.dis = dis; c.n = nn; c.cp = mk3(0.f, 0.f, 0.f);
          return c;
        }
      }
    }
  }
  return c;
}

/* Collision_Response_Particle_VolumeObject, Collision_Response.cuh:241-363,
 * for a lifter without spin (R_Body_COMVOBJ = VelVOBJ_Ang = 0). */
__device__ __forceinline__ void resp_dobject(const DevParams& P, int t, f3 p, f3 v, f3 w, float m,
                                             const DColl& c, f3 Vl, f3& Ftr, f3& Tq, float* tally = nullptr) {
  const f3 Rc = c.cp - p;
  const f3 zero = mk3(0.f, 0.f, 0.f);
  const f3 RVel = v - Vl;
  const f3 Vnormal = dot(RVel, c.n) * c.n;
  const float Fnormal = P.d_kn[t] * c.dis;
  const f3 Fdiss = P.d_cn[t] * Vnormal;
  Ftr = Fnormal * c.n - Fdiss;
  const f3 Fdir = Ftr / length(Ftr);
  const bool attractive = dot(Fdir, c.n) < 0.0f;
  if (attractive) Ftr = Fnormal * c.n;
  const f3 V_CPA = v + cross(w, Rc);
  const f3 V_CPB = Vl + cross(zero, zero);
  const f3 V_CPNA = dot(V_CPA, c.n) * c.n;
  const f3 V_CPNB = dot(V_CPB, c.n) * c.n;
  const f3 Vt_A = V_CPA - V_CPNA;
  const f3 Vt_B = V_CPB - V_CPNB;
  const f3 Vt_Rel = Vt_A - Vt_B;
  float mag = dot(Vt_Rel, Vt_Rel);
  f3 Tang = zero;
  if (mag > 0.0f) {
    mag = __builtin_sqrtf(mag);
    float Velf = dot(RVel, RVel);
    Velf = Velf > 0.0f ? __builtin_sqrtf(Velf) : 1.0f;
    const float tan_f = (0.050f * (m / P.dt)) * (P.d_fk[t] * (mag / Velf));
    Tang = (-1.0f * tan_f) * (Vt_Rel / mag);
  }
  wall_tally(P, tally, TALLY_LIFT, attractive, Fdiss, Vnormal, Tang, mag);
  Ftr = Ftr + Tang;
  Tq = P.rotation ? cross(Rc, Ftr) : zero;
}

/* Lifter force and torque on one particle; returns the contact type kept
 * for the next step (0 none, 1 face, 2 edge).  Only the first lifter in
 * contact acts (`break`, DEM_Kernels_Compute.cuh:224). */
__device__ __forceinline__ int lifter_forces(const DevParams& P, int t, f3 p, float r, f3 v, f3 w,
                                             float m, int prev_type, f3& F, f3& T, float* tally = nullptr) {
  F = mk3(0.f, 0.f, 0.f);
  T = F;
  DColl res{false, 0, 0.0f, F, F};
  const bool check = P.dobj[0].is_translating || !P.mill_sim || P.rotating_wobj ||
                     length(mk3(p.x, p.y, 0.0f) - cvt(P.mill_cent)) > (P.dobj_bound_r - 2.0f * r);
  if (check) {
    for (int d = 0; d < P.num_dobj; ++d) {
      const bdem_DOBJ& D = P.dobj[d];
      const f3 BV = (p - cvt(D.COM)) * cvt(D.Axis);
      if (length(BV) < (D.boundR + r)) {
        res = sphere_dobject(D, p, r, prev_type);
        if (res.hit) {
          if (res.type == 2) res.cp = p;   /* no spin for an edge contact */
          f3 Vl = mk3(0.f, 0.f, 0.f);
          if (D.is_attached_DWorld) Vl = drum_velocity(P, res.cp);
          else if (D.is_translating) Vl = cvt(D.velocity);
          f3 Ftr, Tq;
          resp_dobject(P, t, p, v, w, m, res, Vl, Ftr, Tq, tally);
          F = F + Ftr;
          if (P.rotation) T = T + Tq;
          break;
        }
      }
    }
  }
  return res.type;
}

/* Get_Angular_Acc (Device/Utilities.cuh:486-608): IC = RQ I0, then
 * IN[i][j] = sum_k IC[k][j] RQT[i][k] = (RQ^T RQ I0)[i][j], which is I0 up to
 * rounding, applied to the torque; summation orders as written. */
__device__ __forceinline__ f3 angular_acc(const float* __restrict__ I0, QuatF Q, f3 tau) {
  float RQ[9];
  RQ[0] = (Q.w * Q.w + Q.x * Q.x - Q.y * Q.y - Q.z * Q.z);
  RQ[1] = (2.0f * Q.x * Q.y - 2.0f * Q.w * Q.z);
  RQ[2] = (2.0f * Q.x * Q.z + 2.0f * Q.w * Q.y);
  RQ[3] = (2.0f * Q.x * Q.y + 2.0f * Q.w * Q.z);
  RQ[4] = (Q.w * Q.w - Q.x * Q.x + Q.y * Q.y - Q.z * Q.z);
  RQ[5] = (2.0f * Q.y * Q.z - 2.0f * Q.w * Q.x);
  RQ[6] = (2.0f * Q.x * Q.z - 2.0f * Q.w * Q.y);
  RQ[7] = (2.0f * Q.y * Q.z + 2.0f * Q.w * Q.x);
  RQ[8] = (Q.w * Q.w - Q.x * Q.x - Q.y * Q.y + Q.z * Q.z);
  float IC[9], IN[9];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      float sum = 0.0f;
#pragma unroll
      for (int k = 0; k < 3; ++k) sum += (RQ[r * 3 + k] * I0[k * 3 + c]);
      IC[r * 3 + c] = sum;
    }
  const float RQT[9] = {RQ[0], RQ[3], RQ[6], RQ[1], RQ[4], RQ[7], RQ[2], RQ[5], RQ[8]};
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      float sum = 0.0f;
#pragma unroll
      for (int k = 0; k < 3; ++k) sum += ((IC[k * 3 + c]) * RQT[r * 3 + k]);
      IN[r * 3 + c] = sum;
    }
  const float f[3] = {tau.x, tau.y, tau.z};
  float out[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    float sum = 0.0f;
#pragma unroll
    for (int k = 0; k < 3; ++k) sum += (IN[r * 3 + k] * (f[k]));
    out[r] = sum;
  }
  return mk3(out[0], out[1], out[2]);
}

/* Integrate_Euler_NonSymmetry (Integration.cuh:6-184); the lifter terms are
 * added only when lifters exist (Integration.cuh:41-44, 92-95).  Polyhedra
 * (I0 given) take the tensor product of Get_Angular_Acc and always carry
 * their orientation. */
__device__ __forceinline__ void integrate(const DevParams& P, float inv_inertia, float m, f3 Fpp,
                                          f3 Tpp, f3 Fw, f3 Tw, f3 Fl, f3 Tl, f3& p, f3& v, f3& w,
                                          QuatF* q, const float* I0 = nullptr) {
  const float dt = P.dt;
  f3 acc = Fw + Fpp;
  if (P.num_dobj > 0) acc = acc + Fl;
  acc = acc * (1.0f / m);
  v = v + acc * dt;
  p = p + v * dt;
  if (P.mode2d) v.z = 0.0f;
  f3 fl = Tpp + Tw;
  if (P.num_dobj > 0) fl = fl + Tl;
  if (P.rotation) {
    if (dot(fl, fl) > 0.0f) {
      const f3 aa = I0 ? angular_acc(I0, *q, fl) : inv_inertia * fl;
      w = w + aa * dt;
      w = w - w * P.roll_res;
    }
    if ((P.sphere_orient || I0) && q) {
      float mo = dot(w, w);
      QuatF dQ{1.0f, 0.0f, 0.0f, 0.0f};
      if (mo > 0.000000f) {
        mo = __builtin_sqrtf(mo);
        const f3 dir = w / mo;
        dQ.w = cosf(0.5000f * mo * dt);
        const float sv = sinf(0.5000f * mo * dt);
        dQ.x = sv * dir.x; dQ.y = sv * dir.y; dQ.z = sv * dir.z;
      }
      const QuatF Q = qnormalise(qmul(dQ, *q));
      *q = (Q.w <= -1.00000f) ? dQ : Q;
      if (P.mode2d) { q->x = 0.0f; q->y = 0.0f; }
    }
    if (P.mode2d) { w.x = 0.0f; w.y = 0.0f; }
  }
}

}  // namespace bdem
