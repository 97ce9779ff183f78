/*
 * bdem_kernels.hpp - host-side launchers for the gfx950 step kernels
 * (internal to the library; the public surface is include/bdem_abi.h).
 */
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "bdem_device.hpp"

namespace bdem {

/* Buffers one step touches.  "cur" is the state in the order the previous
 * step left it; "srt" is scratch that receives the sorted gather. */
struct StepBuffers {
  int n;                               /* particle capacity (all slots processed) */
  float4 *pos_cur, *vel_cur, *omg_cur, *quat_cur;
  float4 *pos_srt, *vel_srt, *omg_srt, *quat_srt;
  uint32_t *keys_new;                  /* keys the step kernel (or k_hash) wrote, slot order */
  uint32_t *keys_sorted[2];            /* sorted keys, alternating per step (old/new) */
  int *idx_sorted;                     /* sorted order as original slots */
  uint8_t *changed;                    /* key changed in the last step kernel */
  unsigned long long *sbuf;            /* appended (key<<32 | slot) of changed slots */
  unsigned long long *skey;            /* n composites of scratch for oversized windows */
  uint32_t *meta;                      /* META_* words */
  uint32_t *rk[2];                     /* full radix sort ping-pong */
  int *rv[2];
  uint32_t *hist;                      /* full radix digit counters */
  uint2 *cells;                        /* (start, end) per cell, reference format */
  /* tangential history: records indexed by P_ID, double-buffered across
   * steps; their counts (pair | wall << 16) travel with the state in slot
   * order, so they are read and written coalesced */
  uint32_t *hwc_cur, *hwc_srt;
  HistRec *hrec[2];
  HistRec *wrec[2];
  int hist_cap;                        /* number of P_ID slots in the history tables */
  float *wall_force;                   /* optional n*3 output of the wall force */
  uint32_t *status;                    /* ST_* bits */
  int *kill_count;
  /* skin masks (spheres): which near candidates of each stencil row are
   * within ra + rb + skin; they travel with the state like the spring counts */
  uint64_t *mask_cur, *mask_srt;
  uint32_t *inv_map;                   /* previous sorted position -> new (written by the resort) */
  uint32_t *rowflag;                   /* per sorted position: 9 bits, rows to test in full this step */
  uint32_t *vmeta;                     /* VM_* words */
  float skin;                          /* 0: masks off (every candidate tested every step) */
};

enum { VM_DMAX = 1, VM_U = 3, VM_FULL = 5, VM_WORDS = 8 };

/* What the step kernel needs of the skin masks; par alternates per step. */
struct SkinArgs {
  const uint64_t* mask_in;             /* this step's masks, sorted order */
  uint64_t* mask_out;                  /* written with the state, for the next step */
  uint32_t* rowflag;
  uint32_t* vmeta;
  float skin;
  int par;
  int on;
};

/* Polyhedra: world-frame shapes, one record of `stride` float4 per particle
 * (pv vertex slots, pf face normals, pf face centroids), the bounding-sphere
 * contact list, its pair queue and one force/torque slot per listed pair. */
struct PolyLayout { int pv, pf, stride; };
/* The fields of a shape-table row that the pair tests read for every pair,
 * one 32-byte record per row (in the 3,148-byte row they lie at its start,
 * middle and end, three cache lines apart). */
struct ShapeHdr {
  int nv, nf, ne, pad0;
  float vbound, plane_slack, radius, pad1;
};
struct PolyBuffers {
  PolyLayout L;
  ShapeHdr* hdr;                       /* one per shape-table row */
  float4* local;                       /* per row, `stride` float4 in the rot layout, local frame: vertices,
                                          normals, centroid - COM; w = 1 where the slot is used */
  float4* rot;                         /* n * stride */
  uint32_t* list;                      /* n * 32 partner slots */
  uint32_t* nn;                        /* full contact count per particle */
  uint32_t* queue;                     /* (particle * 32 + k) work items */
  uint2* equeue;                       /* (slot, partner) of the pairs for the wave-per-pair kernel */
  uint32_t* qcount;                    /* [0] queue length, [1] edge queue length, [2] its chunk counter */
  float4 *pf, *pt;                     /* n * 32: pair force (w = 1 when in contact), torque */
  uint2* hint;                         /* n * 32 by P_ID and list slot: (partner P_ID, separating face code) */
};
constexpr int POLY_PAIR_BLOCKS = 2048;   /* a multiple of the 8 XCDs */

constexpr int SORT_THREADS = 256;
constexpr int SORT_ITEMS = 16;
constexpr int SORT_TILE = SORT_THREADS * SORT_ITEMS;
constexpr int RADIX = 256;
/* incremental resort: 1024 threads own 4096 consecutive slots; the changed
 * composites of a block's window are sorted in LDS up to RS_LCAP */
constexpr int RS_THREADS = 1024;
constexpr int RS_TILE = 4096;
constexpr int RS_LCAP = 4096;
/* meta words: two append counters used on alternate steps, the size of the
 * last changed set */
enum { META_APPEND = 0, META_COUNT = 2, META_WORDS = 4 };

inline __host__ __device__ int sort_blocks(int n) { return (n + SORT_TILE - 1) / SORT_TILE; }
int key_bits(uint32_t max_key);
int sort_passes(uint32_t max_key);

/* Everything the sort kernels touch for one step. */
struct SortArgs {
  int n;
  uint32_t C;
  int passes;
  const uint32_t* keys_new;
  const uint8_t* changed;
  const unsigned long long* sbuf;
  unsigned long long* skey;
  const uint32_t* okeys;               /* previous step's sorted keys */
  uint32_t* keys_sorted;
  int* idx_sorted;
  uint32_t* rk[2];
  int* rv[2];
  uint32_t* hist;
  uint32_t* meta;
  uint2* cells;
  uint32_t* status;
  const float4 *pos_cur, *vel_cur, *omg_cur, *quat_cur;
  float4 *pos_srt, *vel_srt, *omg_srt, *quat_srt;
  const uint32_t* hwc_cur;             /* history counts, gathered with the state when set */
  uint32_t* hwc_srt;
  uint32_t* inv_map;                   /* optional: previous position -> new position */
  const uint64_t* mask_cur;            /* skin masks, gathered with the state when set */
  uint64_t* mask_srt;
};

/* Individual stages (all asynchronous on `s`). */
void launch_hash(const DevParams* dP, const StepBuffers& b, hipStream_t s);
/* Sort + gather + cell table.  incremental: keys_new came from the step
 * kernel together with its changed list, counted in meta[META_APPEND + par];
 * otherwise the cooperative full radix sort runs.  Either way the counter
 * meta[META_APPEND + (par ^ 1)] is cleared for the next step kernel. */
hipError_t launch_sort(const SortArgs& a, bool incremental, int par, hipStream_t s);
/* Fused contact search + pair force + walls + integration, srt -> cur.
 * Writes next-step keys, the changed flags (against sorted_keys, this step's
 * sorted keys) and appends changed slots to sbuf, counted in
 * meta[META_APPEND + append_par]. */
void launch_step(const DevParams* dP, const StepBuffers& b, const uint32_t* sorted_keys, int law,
                 bool orient, int hist_par, int append_par, bool do_kill, bool want_wall_force, bool energy,
                 const SkinArgs& sk, uint32_t* dbg_list, uint32_t* dbg_nn, hipStream_t s);
/* Many simulations in one persistent launch (k_step_batch): params[s] is
 * simulation s's parameter block (its key_base offsets its keys and cell
 * table inside the shared sort), tile_sim[t] the simulation owning sorted
 * slots [64 t, 64 t + 64).  queue holds 2 x 8 tile counters, set par serving
 * this step (the persistent launch clears the other).  threads: 64, 128 or
 * 256 per workgroup; blocks: the persistent grid (0: one tile per wave, no
 * queue); grab: tiles taken per atomic. */
struct BatchLaunch {
  const DevParams* params;
  const uint32_t* tile_sim;
  uint32_t ntiles;
  uint32_t grab;
  uint32_t* queue;
  int par;
  const uint2* cells;
  int law;
  int threads;
  int blocks;
};
/* Keys of a batch (n a multiple of 64): tile_sim[i / 64] owns slot i. */
void launch_hash_batch(const DevParams* Pv, const uint32_t* tile_sim, const StepBuffers& b, hipStream_t s);
void launch_step_batch(const BatchLaunch& L, const StepBuffers& b, const uint32_t* sorted_keys, int hist_par,
                       int append_par, hipStream_t s);
/* Rotate_VolObject_Inc on the device: every lifter attached to the drum in
 * the parameter block turned by the angle whose cosine and sine are c, s
 * about (ax, ay, az), its faces re-derived.  One small launch. */
void launch_turn_lifters(DevParams* dP, double c, double s, float ax, float ay, float az, hipStream_t st);
/* Skin masks before a step: decide a new epoch (every row tested in full)
 * from the displacement bound, else flag the stencil rows whose cells a
 * particle entered or left.  Uses this step's cell table; par alternates. */
void launch_skin(const DevParams* dP, const StepBuffers& b, const uint32_t* old_keys, int par, bool full,
                 hipStream_t s);

/* Polyhedra step (rotate -> list + queue -> pairs -> finish), srt -> cur,
 * with the same outputs as launch_step; and the force stage alone. */
void launch_poly_step(const DevParams* dP, const bdem_poly_shape* dshape, const StepBuffers& b, const PolyBuffers& pb,
                      const uint32_t* sorted_keys, int append_par, bool do_kill, bool want_wall_force,
                      bool lifters, bool energy, hipStream_t s);
/* rot <- world-frame vertices, normals and centroids of the current state */
void launch_poly_world_shapes(const DevParams* dP, const bdem_poly_shape* dshape, const StepBuffers& b,
                              const PolyBuffers& pb, hipStream_t s);
void launch_poly_forces(const DevParams* dP, const bdem_poly_shape* dshape, const StepBuffers& b, const PolyBuffers& pb,
                        float* fpp, float* tpp, float* fw, float* tw, hipStream_t s);

/* Ensemble statistics of the current state (bdem_stats.hip): live
 * particles with y > min_y (count, sum of |v|*0.01, the fastest as
 * (speed bits << 32 | P_ID + 1)), and live centres inside [lo, hi) (count,
 * sum of 4/3 pi r^3 when want_volume).  `out` must be zeroed first. */
struct StatsWords {
  uint32_t count, box_count;
  double sum_speed, box_volume;
  unsigned long long best;
};
void launch_stats(const StepBuffers& b, float min_y, float3 lo, float3 hi, bool want_volume, StatsWords* out,
                  hipStream_t s);
/* Free-surface profile (bdem_stats.hip): per bin of width 1/inv_w along
 * axis 0 (x) or 2 (z) from lo, the largest y + r over the live particles, as
 * an order-preserving uint32 key of the float (0: empty).  `keys` (bins
 * words, bins <= PROFILE_MAX_BINS) must be zeroed first. */
constexpr int PROFILE_MAX_BINS = 4096;
void launch_profile(const StepBuffers& b, int axis, float lo, float inv_w, int bins, uint32_t* keys, hipStream_t s);
/* The allocation-free profile's scratch: PROFILE_MAX_BINS key words, then
 * PROFILE_MAX_BINS finished heights (float), in a file-scope device array
 * of bdem_stats.hip.  noalloc_scratch returns its address on `device`
 * (which must be current), taken once per device. */
constexpr int NOALLOC_SCRATCH_WORDS = 2 * PROFILE_MAX_BINS;
hipError_t noalloc_scratch(int device, uint32_t** out);
/* Clear the keys, bin, and write the heights to scr + PROFILE_MAX_BINS. */
void launch_profile_noalloc(const StepBuffers& b, int axis, float lo, float inv_w, int bins, uint32_t* scr,
                            hipStream_t s);

/* Debug stages on the sorted scratch state. */
void launch_contact_list(const DevParams* dP, const StepBuffers& b, uint32_t* list32, uint32_t* nn,
                         hipStream_t s);
void launch_forces(const DevParams* dP, const StepBuffers& b, int law, int hist_par, float* fpp,
                   float* tpp, float* fw, float* tw, hipStream_t s);
void launch_fill_u32(uint32_t* p, size_t count, uint32_t v, hipStream_t s);
void launch_copy_u4(uint4* dst, const uint4* src, size_t count, hipStream_t s);

}  // namespace bdem
