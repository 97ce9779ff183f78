/*
 * bdem_ctx.hpp - the context behind the opaque bdem_ctx handle (internal to
 * libbdem.so; shared by the single-simulation ABI and the batched ABI, which
 * reads a configured context's parameters and state).
 */
#pragma once
#include <hip/hip_runtime.h>
#include <string>
#include <vector>
#include "bdem_abi.h"
#include "bdem_kernels.hpp"

using bdem::DevParams;
using bdem::PolyBuffers;
using bdem::StatsWords;
using bdem::StepBuffers;

struct bdem_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  bool configured = false;

  bdem_SimulationInfo sim{};
  std::vector<bdem_WOBJ> wobj;
  std::vector<bdem_POBJ> pobj;
  std::vector<bdem_DOBJ> dobj;
  bdem_ext_params ext{};
  bool has_ext = false;
  int num_wobj_active = 0;
  bool cyl_rotation = false;
  bool record_wf = false;
  bool start_bin = false;
  float tallys[6] = {0, 0, 0, 0, 0, 0};
  float* d_tally = nullptr;  /* device tallies {particle, lifter, surface} the step kernel adds to */
  StatsWords* d_stats = nullptr;   /* bdem_particle_stats reduction words */
  uint32_t* d_profile = nullptr;   /* bdem_get_surface_profile bins */
  /* lifters turned on the device before every step (Rotate_VolObject_Inc) */
  bool lift_on = false;
  double lift_c = 1.0, lift_s = 0.0;
  float lift_about[3] = {0.f, 0.f, 0.f};
  bool dobj_dirty = false;   /* the device lifter table is ahead of c->dobj */

  DevParams hp{};
  DevParams* dp = nullptr;
  StepBuffers b{};
  std::vector<void*> allocs;
  bool orient = false;
  double vbound_scale = 1.0;  /* largest |q|^2 (>= 1) of the last orientation upload */
  bool poly = false;         /* particle_type 1 */
  /* polyhedra: the shape table the kernels read, one row per type (the
   * POBJ table's geometry) or, after bdem_set_poly_shapes, one per P_ID */
  std::vector<bdem_poly_shape> shapes;
  bool shape_by_pid = false;
  bdem_poly_shape* dshape = nullptr;
  PolyBuffers pb{};
  bool keys_valid = false;   /* keys_new holds the keys of the current state */
  bool flags_valid = false;  /* ... and the step kernel's changed flags for them */
  int hist_par = 0;
  int sorted_par = 0;        /* keys_sorted[sorted_par] = latest sorted keys */
  int append_par = 0;        /* meta[META_APPEND + append_par] counts the changed list */
  bool full_next = false;    /* a discharge check ran: many keys changed at once */
  bool lists_full = true;    /* every stencil row is tested in full at the next step */
  int vpar = 0;              /* skin-mask epoch words of the last step */
  float skin_req = -1.0f;    /* requested skin (cm); negative or 0: masks off */
  uint32_t* dbg_list = nullptr;  /* debug: the contacts the step kernel used */
  uint32_t* dbg_nn = nullptr;
  int stepc = 0;
  /* two steady-state steps captured as a HIP graph, one per parity of the
   * per-step alternations (see bdem_step) */
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  int graph_hist[2] = {0, 0}, graph_sorted[2] = {0, 0};
  bool graphs_on = true;
  /* stage timing: one event per stage boundary per recorded step */
  std::vector<hipEvent_t> tev;
  std::vector<char> t_lists;   /* per recorded step: the lists boundary was recorded */
  int t_cap = 0, t_used = 0;
};
