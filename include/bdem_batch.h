/*
 * bdem_batch.h - many independent DEM simulations advanced together on one
 * GPU (libbdem.so).  No reference counterpart: the reference keeps one
 * simulation in process-global state (Device/DeviceInterface.cu:52-186) and
 * launches its kernels once per simulation step (:1825-2513).
 *
 * A batch is built from configured single-simulation contexts: each
 * bdem_batch_add copies a context's parameters (world objects, materials,
 * contact law, grid, time step) and its current state.  bdem_batch_finalize
 * places every simulation in one set of device buffers (its slots padded to
 * whole 64-slot tiles, its cell keys offset into a private key range), after
 * which one step of every simulation is two launches: the shared
 * incremental re-sort, and one fused step launch over (simulation, tile)
 * work.  Steady-state steps replay a captured HIP graph.
 *
 * Each simulation evolves exactly as it would alone in a context with the
 * same SimulationInfo (bdem_batch_instance_sim returns the one in use): the
 * same float operations in the same order, so the states are identical.
 *
 * Restrictions: spheres only, one contact law for the whole batch, no
 * lifters, no discharge, no skin masks, no dissipation tallies.
 */
#ifndef BDEM_BATCH_H_
#define BDEM_BATCH_H_

#include "bdem_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bdem_batch bdem_batch;

/* Launch shape of the batched step.  threads: workgroup size (64, 128 or
 * 256).  waves_per_cu: 0 launches one wave per 64-slot tile (a plain grid);
 * otherwise a persistent grid of waves_per_cu waves per compute unit (1..16)
 * deals the tiles from per-XCD queues, grab consecutive tiles (1..64) per
 * queue atomic.  cell_factor (>= 1): every simulation's hash cell is its
 * context's cell size times this factor (applied at finalize).  use_graph:
 * replay steady-state steps from a captured HIP graph. */
typedef struct bdem_batch_tuning {
  int threads;
  int waves_per_cu;
  int grab;
  float cell_factor;
  int use_graph;
} bdem_batch_tuning;

int bdem_batch_create(bdem_batch** out, int device);
int bdem_batch_destroy(bdem_batch* b);
const char* bdem_batch_last_error(const bdem_batch* b);

/* Copy a configured sphere context (its parameters and current state: the
 * state must be as uploaded, before any step) into the batch as simulation
 * *index.  The context is not referenced afterwards. */
int bdem_batch_add(bdem_batch* b, bdem_ctx* ctx, int* index);
/* Launch shape; takes effect at the next bdem_batch_finalize (cell_factor)
 * or step (the rest).  NULL restores the defaults {256, 0, 1, 1.0, 1}. */
int bdem_batch_set_tuning(bdem_batch* b, const bdem_batch_tuning* t);
int bdem_batch_get_tuning(const bdem_batch* b, bdem_batch_tuning* t);
/* Allocate and upload every simulation's initial state (the state it had
 * when added).  Calling it again restarts every simulation from there. */
int bdem_batch_finalize(bdem_batch* b);
/* Advance every simulation by nsteps (asynchronous). */
int bdem_batch_step(bdem_batch* b, int nsteps);
int bdem_batch_sync(bdem_batch* b);
/* Try launch shapes and cell factors on the batch's own simulations (each
 * trial restarts them from their initial state, runs 2 untimed and `steps`
 * timed steps), keep the fastest and restart from the initial state with
 * it.  best (may be NULL) receives the choice, ms_per_step (may be NULL)
 * its time.  Candidates: cell factor 1 and 1.25; threads 64, 128, 256;
 * a plain grid, and persistent grids of 8 and 16 waves per CU with grabs 1
 * and 4. */
int bdem_batch_autotune(bdem_batch* b, int steps, bdem_batch_tuning* best, double* ms_per_step);

int bdem_batch_num_instances(const bdem_batch* b);
/* Particles of simulation index (its context's Num_Particles). */
int bdem_batch_num_particles(const bdem_batch* b, int index);
/* The SimulationInfo simulation index runs with (its context's, with the
 * cell size and cell counts of the cell factor). */
int bdem_batch_instance_sim(const bdem_batch* b, int index, bdem_SimulationInfo* sim);
/* State of simulation index (synchronous): its Num_Particles particles in
 * the batch's current order, P_IDs as the context numbered them.  Any
 * output may be NULL. */
int bdem_batch_get_state(bdem_batch* b, int index, bdem_float3* pos, bdem_float3* vel, bdem_float3* avel,
                         int* pid);
/* Kernel warnings accumulated since the last call (BDEM_WARN_* bits). */
int bdem_batch_get_status(bdem_batch* b, unsigned int* warnings);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* BDEM_BATCH_H_ */
