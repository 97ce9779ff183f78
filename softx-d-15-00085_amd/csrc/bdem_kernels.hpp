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
  uint32_t *keys[2];                   /* radix ping-pong keys */
  int *vals[2];                        /* radix ping-pong payload (original slot) */
  uint32_t *radix_hist;                /* 256 * ceil(n / tile) counters */
  /* incremental sort: changed flags from the step kernel, per-256 counts,
   * the unchanged run (rkeys/rvals), the changed set (sbuf) and
   * sort_meta = {|S|, fallback flag} */
  uint8_t *changed;
  uint32_t *blk_changed;
  uint32_t *rkeys;
  int *rvals;
  unsigned long long *sbuf;
  uint32_t *sort_meta;
  uint2 *cells;                        /* (start, end) per cell, reference format */
  /* tangential history, indexed by P_ID, double-buffered across steps */
  int *hcount[2]; HistRec *hrec[2];
  int *wcount[2]; HistRec *wrec[2];
  int hist_cap;                        /* number of P_ID slots in the history tables */
  float *wall_force;                   /* optional n*3 output of the wall force */
  uint32_t *status;                    /* ST_* bits */
  int *kill_count;
};

constexpr int SORT_THREADS = 256;
constexpr int SORT_ITEMS = 16;
constexpr int SORT_TILE = SORT_THREADS * SORT_ITEMS;
constexpr int RADIX = 256;

inline int sort_blocks(int n) { return (n + SORT_TILE - 1) / SORT_TILE; }
int key_bits(uint32_t max_key);
int sort_passes(uint32_t max_key);

/* Individual stages (all asynchronous on `s`). */
void launch_hash(const DevParams* dP, const StepBuffers& b, hipStream_t s);
/* Stable sort of keys[0] (payload: slot index) for keys <= max_key.
 * incremental: keys[0] came from the step kernel together with its changed
 * flags; otherwise a full radix sort runs.  Returns the index (0 or 1) of the
 * buffer pair holding the result. */
int launch_sort(const StepBuffers& b, uint32_t max_key, bool incremental, hipStream_t s);
void launch_gather(const StepBuffers& b, const int* idx, bool orient, hipStream_t s);
void launch_cells(const DevParams* dP, const StepBuffers& b, const uint32_t* sorted_keys,
                  uint32_t ncells, hipStream_t s);
/* Fused contact search + pair force + walls + integration, srt -> cur.
 * Writes next-step keys into keys[0] and the changed flags/counts against
 * sorted_keys (this step's sorted keys). */
void launch_step(const DevParams* dP, const StepBuffers& b, const uint32_t* sorted_keys, int law,
                 bool orient, int hist_par, bool do_kill, bool want_wall_force, hipStream_t s);

/* Debug stages on the sorted scratch state. */
void launch_contact_list(const DevParams* dP, const StepBuffers& b, uint32_t* list32, uint32_t* nn,
                         hipStream_t s);
void launch_forces(const DevParams* dP, const StepBuffers& b, int law, int hist_par, float* fpp,
                   float* tpp, float* fw, float* tw, hipStream_t s);
void launch_fill_u32(uint32_t* p, size_t count, uint32_t v, hipStream_t s);

}  // namespace bdem
