/*
 * bdem_batch.cpp - many independent simulations in one set of device
 * buffers (include/bdem_batch.h).
 *
 * Layout.  Simulation s owns the slot range [off_s, off_s + cap_s), cap_s
 * its particle count rounded up to a whole number of 64-slot tiles; the
 * padding slots are dead particles (P_ID -1) that every step copies through.
 * Its cell keys are its own hash plus key_base_s, the key ranges
 * [key_base_s, key_base_s + C_s + 2) (cells, then the out-of-grid and dead
 * sentinels) following one another in simulation order.  One stable sort of
 * all slots therefore keeps each simulation inside its slot range, in the
 * order its own sort would give, and each simulation reads the cell table at
 * cells + key_base_s with its own hash.  P_IDs become off_s + P_ID, so the
 * per-P_ID spring records of all simulations share one table.
 *
 * A step is the shared re-sort (incremental after the first step) and one
 * k_step_batch launch whose waves take 64-slot tiles, each with its
 * simulation's parameter block.  All per-step state alternates with period
 * two (sorted keys, append counters, spring buffers, tile queues), so two
 * consecutive steps are captured once in a HIP graph and replayed.
 */
#include <hip/hip_runtime.h>
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>
#include "bdem_batch.h"
#include "bdem_ctx.hpp"
#include "bdem_kernels.hpp"

using namespace bdem;

namespace {

struct Instance {
  DevParams hp{};                 /* the context's parameter block */
  bdem_SimulationInfo sim{};      /* the context's SimulationInfo */
  bdem_SimulationInfo sim_used{}; /* with the cell factor applied */
  int n = 0;
  std::vector<float4> P, V, O;    /* state when added, the context's slot order */
  int off = 0, cap = 0;
  uint32_t key_base = 0;
};

}  // namespace

struct bdem_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  std::vector<Instance> inst;
  int law = -1;
  bdem_batch_tuning tune{256, 0, 1, 1.0f, 1};
  float built_cell_factor = 0.0f;
  bool finalized = false;
  int cus = 256;

  StepBuffers b{};
  DevParams* dparams = nullptr;
  uint32_t* tile_sim = nullptr;
  uint32_t* queue = nullptr;      /* 2 x NUM_XCD tile counters */
  uint32_t C_tot = 0;
  int ntiles = 0;
  std::vector<void*> allocs;

  bool keys_valid = false;        /* false: the next step hashes and fully sorts */
  int hist_par = 0, sorted_par = 0, append_par = 0, qpar = 0;
  hipGraphExec_t graph[2] = {nullptr, nullptr};   /* two-step graph per parity */
  bdem_batch_tuning graph_tune[2]{};
};

namespace {

int bfail(bdem_batch* B, int code, const std::string& msg) {
  if (B) B->err = msg;
  return code;
}

#define BHIP(B, expr)                                                               \
  do {                                                                              \
    hipError_t e_ = (expr);                                                         \
    if (e_ != hipSuccess)                                                           \
      return bfail(B, BDEM_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <class T>
int balloc(bdem_batch* B, T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc((void**)p, count * sizeof(T));
  if (e != hipSuccess) return bfail(B, BDEM_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  B->allocs.push_back((void*)*p);
  return BDEM_OK;
}

void drop_graphs(bdem_batch* B) {
  for (auto& g : B->graph) {
    if (g) (void)hipGraphExecDestroy(g);
    g = nullptr;
  }
}

void free_buffers(bdem_batch* B) {
  drop_graphs(B);
  for (void* p : B->allocs) (void)hipFree(p);
  B->allocs.clear();
  B->b = StepBuffers{};
  B->dparams = nullptr;
  B->tile_sim = nullptr;
  B->queue = nullptr;
  B->finalized = false;
}

inline float bits_f(int v) { float f; std::memcpy(&f, &v, 4); return f; }
inline int f_bits(float f) { int v; std::memcpy(&v, &f, 4); return v; }

/* The grid of a simulation with its cell size times f: the same origin, the
 * cell counts that cover the context's grid. */
void apply_cell_factor(Instance& I, float f) {
  I.sim_used = I.sim;
  if (f == 1.0f) return;
  const bdem_float3 c0 = I.sim.cellSize;
  bdem_float3 c1{c0.x * f, c0.y * f, c0.z * f};
  auto cover = [](int n0, float cell0, float cell1) {
    const double span = double(n0) * double(cell0);
    int n1 = (int)std::ceil(span / double(cell1));
    return n1 < 1 ? 1 : n1;
  };
  I.sim_used.cellSize = c1;
  I.sim_used.num_NNCells.x = cover(I.sim.num_NNCells.x, c0.x, c1.x);
  I.sim_used.num_NNCells.y = cover(I.sim.num_NNCells.y, c0.y, c1.y);
  I.sim_used.num_NNCells.z = cover(I.sim.num_NNCells.z, c0.z, c1.z);
}

SortArgs sort_args(bdem_batch* B, bool incremental) {
  const StepBuffers& b = B->b;
  const int prev = B->sorted_par;
  const int next = incremental ? prev ^ 1 : prev;
  SortArgs a{};
  a.n = b.n;
  a.C = B->C_tot;
  a.passes = sort_passes(B->C_tot);
  a.keys_new = b.keys_new;
  a.changed = b.changed;
  a.sbuf = b.sbuf;
  a.skey = b.skey;
  a.okeys = b.keys_sorted[prev];
  a.keys_sorted = b.keys_sorted[next];
  a.idx_sorted = b.idx_sorted;
  a.rk[0] = b.rk[0]; a.rk[1] = b.rk[1];
  a.rv[0] = b.rv[0]; a.rv[1] = b.rv[1];
  a.hist = b.hist;
  a.meta = b.meta;
  a.cells = b.cells;
  a.status = b.status;
  a.pos_cur = b.pos_cur; a.vel_cur = b.vel_cur; a.omg_cur = b.omg_cur;
  a.pos_srt = b.pos_srt; a.vel_srt = b.vel_srt; a.omg_srt = b.omg_srt;
  a.hwc_cur = b.hwc_cur;
  a.hwc_srt = b.hwc_srt;
  return a;
}

/* One step's launches on the batch's stream (also used under capture). */
int launch_one(bdem_batch* B) {
  const bool incremental = B->keys_valid;
  if (!incremental) launch_hash_batch(B->dparams, B->tile_sim, B->b, B->stream);
  const SortArgs a = sort_args(B, incremental);
  hipError_t e = launch_sort(a, incremental, B->append_par, B->stream);
  if (e != hipSuccess) return bfail(B, BDEM_EHIP, std::string("sort launch: ") + hipGetErrorString(e));
  if (incremental) B->sorted_par ^= 1;
  B->append_par ^= 1;
  BatchLaunch L{};
  L.params = B->dparams;
  L.tile_sim = B->tile_sim;
  L.ntiles = (uint32_t)B->ntiles;
  L.grab = (uint32_t)B->tune.grab;
  L.queue = B->queue;
  L.par = B->qpar;
  L.cells = B->b.cells;
  L.law = B->law;
  L.threads = B->tune.threads;
  L.blocks = B->tune.waves_per_cu > 0 ? std::max(1, B->cus * B->tune.waves_per_cu / (B->tune.threads / 64)) : 0;
  launch_step_batch(L, B->b, B->b.keys_sorted[B->sorted_par], B->hist_par, B->append_par, B->stream);
  e = hipGetLastError();
  if (e != hipSuccess) return bfail(B, BDEM_EHIP, std::string("step launch: ") + hipGetErrorString(e));
  if (B->law != BDEM_LAW_REFERENCE) B->hist_par ^= 1;
  if (B->tune.waves_per_cu > 0) B->qpar ^= 1;
  B->keys_valid = true;
  return BDEM_OK;
}

bool same_launch(const bdem_batch_tuning& a, const bdem_batch_tuning& b) {
  return a.threads == b.threads && a.waves_per_cu == b.waves_per_cu && a.grab == b.grab;
}

/* Two steady-state steps, captured once per parity and launch shape. */
int graph_pair(bdem_batch* B) {
  const int key = B->append_par;
  if (B->graph[key] && !same_launch(B->graph_tune[key], B->tune)) {
    (void)hipGraphExecDestroy(B->graph[key]);
    B->graph[key] = nullptr;
  }
  if (!B->graph[key]) {
    /* the launches advance the host-side parities; capture from a copy of
     * them and check that two steps bring them back */
    const int hp = B->hist_par, sp = B->sorted_par, ap = B->append_par, qp = B->qpar;
    BHIP(B, hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal));
    int rc = launch_one(B);
    if (rc == BDEM_OK) rc = launch_one(B);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(B->stream, &g);
    if (rc != BDEM_OK) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return bfail(B, BDEM_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    if (B->hist_par != hp || B->sorted_par != sp || B->append_par != ap || B->qpar != qp) {
      (void)hipGraphDestroy(g);
      return bfail(B, BDEM_ESTATE, "internal: a step pair did not return to its parity");
    }
    e = hipGraphInstantiate(&B->graph[key], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      B->graph[key] = nullptr;
      return bfail(B, BDEM_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    B->graph_tune[key] = B->tune;
  }
  BHIP(B, hipGraphLaunch(B->graph[key], B->stream));
  return BDEM_OK;
}

int upload_initial(bdem_batch* B) {
  const int n = B->b.n;
  std::vector<float4> P(n), V(n), O(n);
  for (const Instance& I : B->inst) {
    for (int k = 0; k < I.cap; ++k) {
      const int s = I.off + k;
      if (k < I.n) {
        P[s] = I.P[k];
        V[s] = I.V[k];
        const int w = f_bits(I.O[k].w);
        const int pid = w >> 3, type = w & 7;
        O[s] = make_float4(I.O[k].x, I.O[k].y, I.O[k].z, bits_f(pack_id_type(pid < 0 ? -1 : I.off + pid, type)));
      } else {
        /* padding: an inert dead slot */
        P[s] = make_float4(0.f, 0.f, 0.f, I.P.empty() ? 1.f : I.P[0].w);
        V[s] = make_float4(0.f, 0.f, 0.f, I.V.empty() ? 1.f : I.V[0].w);
        O[s] = make_float4(0.f, 0.f, 0.f, bits_f(pack_id_type(-1, 0)));
      }
    }
  }
  BHIP(B, hipMemcpy(B->b.pos_cur, P.data(), n * sizeof(float4), hipMemcpyHostToDevice));
  BHIP(B, hipMemcpy(B->b.vel_cur, V.data(), n * sizeof(float4), hipMemcpyHostToDevice));
  BHIP(B, hipMemcpy(B->b.omg_cur, O.data(), n * sizeof(float4), hipMemcpyHostToDevice));
  if (B->b.hwc_cur) {
    launch_fill_u32(B->b.hwc_cur, n, 0u, B->stream);
    launch_fill_u32(B->b.hwc_srt, n, 0u, B->stream);
  }
  launch_fill_u32(B->b.status, 1, 0u, B->stream);
  launch_fill_u32(B->b.meta, META_WORDS, 0u, B->stream);
  launch_fill_u32(B->queue, 2 * NUM_XCD, 0u, B->stream);
  BHIP(B, hipStreamSynchronize(B->stream));
  B->keys_valid = false;
  B->hist_par = B->sorted_par = B->append_par = B->qpar = 0;
  drop_graphs(B);
  return BDEM_OK;
}

int build(bdem_batch* B) {
  if (B->inst.empty()) return bfail(B, BDEM_ESTATE, "no simulation added");
  BHIP(B, hipSetDevice(B->device));
  BHIP(B, hipStreamSynchronize(B->stream));
  free_buffers(B);
  /* placement */
  long long off = 0;
  unsigned long long kb = 0;
  std::vector<DevParams> hps;
  std::vector<uint32_t> tiles;
  for (size_t s = 0; s < B->inst.size(); ++s) {
    Instance& I = B->inst[s];
    apply_cell_factor(I, B->tune.cell_factor);
    DevParams P = I.hp;
    P.cell = I.sim_used.cellSize;
    P.nx = I.sim_used.num_NNCells.x;
    P.ny = I.sim_used.num_NNCells.y;
    P.nz = I.sim_used.num_NNCells.z;
    const unsigned long long C = (unsigned long long)P.nx * P.ny * P.nz;
    if (C + 2 > 0xFFFFFFF0ull) return bfail(B, BDEM_EINVAL, "grid too large for 32-bit keys (simulation " + std::to_string(s) + ")");
    P.C = (uint32_t)C;
    I.off = (int)off;
    I.cap = (I.n + 63) / 64 * 64;
    I.key_base = (uint32_t)kb;
    P.key_base = I.key_base;
    P.tally = nullptr;
    hps.push_back(P);
    for (int t = 0; t < I.cap / 64; ++t) tiles.push_back((uint32_t)s);
    off += I.cap;
    kb += C + 2;
    if (off >= (1ll << 28)) return bfail(B, BDEM_EINVAL, "more than 2^28 slots in one batch");
    if (kb > 0xFFFFFFF0ull) return bfail(B, BDEM_EINVAL, "the batch's cell keys exceed 32 bits");
  }
  const int n = (int)off;
  B->C_tot = (uint32_t)kb;
  B->ntiles = (int)tiles.size();
  StepBuffers& b = B->b;
  b.n = n;
  int rc;
#define BALLOC(p, cnt) if ((rc = balloc(B, &(p), (cnt))) != BDEM_OK) return rc
  BALLOC(B->dparams, hps.size());
  BALLOC(B->tile_sim, tiles.size());
  BALLOC(B->queue, 2 * NUM_XCD);
  BALLOC(b.pos_cur, n); BALLOC(b.vel_cur, n); BALLOC(b.omg_cur, n);
  BALLOC(b.pos_srt, n); BALLOC(b.vel_srt, n); BALLOC(b.omg_srt, n);
  BALLOC(b.keys_new, n); BALLOC(b.keys_sorted[0], n); BALLOC(b.keys_sorted[1], n);
  BALLOC(b.idx_sorted, n);
  BALLOC(b.changed, n);
  BALLOC(b.sbuf, n);
  BALLOC(b.skey, n);
  BALLOC(b.meta, META_WORDS);
  BALLOC(b.rk[0], n); BALLOC(b.rk[1], n); BALLOC(b.rv[0], n); BALLOC(b.rv[1], n);
  BALLOC(b.hist, (size_t)RADIX * sort_blocks(n));
  BALLOC(b.cells, (size_t)B->C_tot);
  BALLOC(b.status, 1); BALLOC(b.kill_count, 1);
  BALLOC(b.hwc_cur, n); BALLOC(b.hwc_srt, n);
  if (B->law != BDEM_LAW_REFERENCE) {
    b.hist_cap = n;
    for (int h = 0; h < 2; ++h) {
      BALLOC(b.hrec[h], (size_t)n * HMAX);
      BALLOC(b.wrec[h], (size_t)n * WMAX);
    }
  }
#undef BALLOC
  BHIP(B, hipMemcpy(B->dparams, hps.data(), hps.size() * sizeof(DevParams), hipMemcpyHostToDevice));
  BHIP(B, hipMemcpy(B->tile_sim, tiles.data(), tiles.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  launch_fill_u32(reinterpret_cast<uint32_t*>(b.kill_count), 1, 0u, B->stream);
  B->built_cell_factor = B->tune.cell_factor;
  B->finalized = true;
  return upload_initial(B);
}

int check_tuning(bdem_batch* B, const bdem_batch_tuning& t) {
  if (t.threads != 64 && t.threads != 128 && t.threads != 256)
    return bfail(B, BDEM_EINVAL, "threads must be 64, 128 or 256");
  if (t.waves_per_cu < 0 || t.waves_per_cu > 16) return bfail(B, BDEM_EINVAL, "waves_per_cu must be in [0, 16]");
  if (t.waves_per_cu > 0 && t.waves_per_cu % (t.threads / 64) != 0)
    return bfail(B, BDEM_EINVAL, "waves_per_cu must be a multiple of the waves per workgroup");
  if (t.grab < 1 || t.grab > 64) return bfail(B, BDEM_EINVAL, "grab must be in [1, 64]");
  if (!(t.cell_factor >= 1.0f && t.cell_factor <= 4.0f)) return bfail(B, BDEM_EINVAL, "cell_factor must be in [1, 4]");
  return BDEM_OK;
}

int require_final(bdem_batch* B) {
  if (!B) return BDEM_EINVAL;
  if (!B->finalized) return bfail(B, BDEM_ESTATE, "bdem_batch_finalize has not been called");
  return BDEM_OK;
}

}  // namespace

extern "C" {

int bdem_batch_create(bdem_batch** out, int device) {
  if (!out) return BDEM_EINVAL;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return BDEM_EHIP;
  if (device < 0 || device >= ndev) return BDEM_EINVAL;
  bdem_batch* B = new bdem_batch();
  B->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking) != hipSuccess) {
    delete B;
    return BDEM_EHIP;
  }
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0) B->cus = cus;
  *out = B;
  return BDEM_OK;
}

int bdem_batch_destroy(bdem_batch* B) {
  if (!B) return BDEM_EINVAL;
  (void)hipSetDevice(B->device);
  if (B->stream) (void)hipStreamSynchronize(B->stream);
  free_buffers(B);
  if (B->stream) (void)hipStreamDestroy(B->stream);
  delete B;
  return BDEM_OK;
}

const char* bdem_batch_last_error(const bdem_batch* B) { return B ? B->err.c_str() : "null batch"; }

int bdem_batch_add(bdem_batch* B, bdem_ctx* c, int* index) {
  if (!B) return BDEM_EINVAL;
  if (!c || !c->configured) return bfail(B, BDEM_EINVAL, "the context is not configured");
  if (c->poly || c->orient) return bfail(B, BDEM_ENOTSUP, "batches hold sphere simulations without orientations");
  if (!c->dobj.empty()) return bfail(B, BDEM_ENOTSUP, "batches hold simulations without lifters");
  if (c->sim.Kill_Particles) return bfail(B, BDEM_ENOTSUP, "batches hold simulations without discharge checks");
  if (c->sim.EnergyCalc) return bfail(B, BDEM_ENOTSUP, "batches keep no dissipation tallies");
  if (c->b.skin > 0.0f) return bfail(B, BDEM_ENOTSUP, "batches do not use skin masks");
  if (c->flags_valid) return bfail(B, BDEM_ESTATE, "the context has stepped since its state was uploaded");
  const int law = c->hp.law;
  if (B->law >= 0 && law != B->law) return bfail(B, BDEM_EINVAL, "every simulation of a batch uses one contact law");
  if (hipSetDevice(c->device) != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess)
    return bfail(B, BDEM_EHIP, "cannot reach the context's device");
  Instance I;
  I.hp = c->hp;
  I.sim = c->sim;
  I.n = c->b.n;
  I.P.resize(I.n); I.V.resize(I.n); I.O.resize(I.n);
  BHIP(B, hipMemcpy(I.P.data(), c->b.pos_cur, I.n * sizeof(float4), hipMemcpyDeviceToHost));
  BHIP(B, hipMemcpy(I.V.data(), c->b.vel_cur, I.n * sizeof(float4), hipMemcpyDeviceToHost));
  BHIP(B, hipMemcpy(I.O.data(), c->b.omg_cur, I.n * sizeof(float4), hipMemcpyDeviceToHost));
  for (int i = 0; i < I.n; ++i) {
    const int pid = f_bits(I.O[i].w) >> 3;
    if (pid >= I.n) return bfail(B, BDEM_EINVAL, "batched simulations need P_ID < Num_Particles");
  }
  B->law = law;
  B->inst.push_back(std::move(I));
  B->finalized = false;
  if (index) *index = (int)B->inst.size() - 1;
  return BDEM_OK;
}

int bdem_batch_set_tuning(bdem_batch* B, const bdem_batch_tuning* t) {
  if (!B) return BDEM_EINVAL;
  const bdem_batch_tuning d{256, 0, 1, 1.0f, 1};
  const bdem_batch_tuning v = t ? *t : d;
  int rc = check_tuning(B, v);
  if (rc) return rc;
  if (B->finalized) (void)hipStreamSynchronize(B->stream);
  B->tune = v;
  return BDEM_OK;
}

int bdem_batch_get_tuning(const bdem_batch* B, bdem_batch_tuning* t) {
  if (!B || !t) return BDEM_EINVAL;
  *t = B->tune;
  return BDEM_OK;
}

int bdem_batch_finalize(bdem_batch* B) {
  if (!B) return BDEM_EINVAL;
  if (B->finalized && B->built_cell_factor == B->tune.cell_factor) {
    BHIP(B, hipSetDevice(B->device));
    BHIP(B, hipStreamSynchronize(B->stream));
    return upload_initial(B);
  }
  return build(B);
}

int bdem_batch_step(bdem_batch* B, int nsteps) {
  int rc = require_final(B);
  if (rc) return rc;
  if (nsteps < 0) return bfail(B, BDEM_EINVAL, "nsteps must be >= 0");
  if (B->built_cell_factor != B->tune.cell_factor)
    return bfail(B, BDEM_ESTATE, "the cell factor changed: call bdem_batch_finalize");
  BHIP(B, hipSetDevice(B->device));
  int s = 0;
  while (s < nsteps) {
    if (B->keys_valid && B->tune.use_graph && nsteps - s >= 2) {
      if ((rc = graph_pair(B)) != BDEM_OK) return rc;
      s += 2;
    } else {
      if ((rc = launch_one(B)) != BDEM_OK) return rc;
      s += 1;
    }
  }
  return BDEM_OK;
}

int bdem_batch_sync(bdem_batch* B) {
  int rc = require_final(B);
  if (rc) return rc;
  BHIP(B, hipSetDevice(B->device));
  BHIP(B, hipStreamSynchronize(B->stream));
  return BDEM_OK;
}

int bdem_batch_autotune(bdem_batch* B, int steps, bdem_batch_tuning* best, double* ms_per_step) {
  if (!B) return BDEM_EINVAL;
  if (B->inst.empty()) return bfail(B, BDEM_ESTATE, "no simulation added");
  if (steps < 1 || steps > 100000) return bfail(B, BDEM_EINVAL, "steps must be in [1, 100000]");
  BHIP(B, hipSetDevice(B->device));
  const bdem_batch_tuning keep = B->tune;
  std::vector<bdem_batch_tuning> cand;
  for (float cf : {1.0f, 1.25f})
    for (int th : {64, 128, 256}) {
      cand.push_back(bdem_batch_tuning{th, 0, 1, cf, keep.use_graph});
      for (int w : {8, 16})
        for (int g : {1, 4}) cand.push_back(bdem_batch_tuning{th, w, g, cf, keep.use_graph});
    }
  hipEvent_t e0, e1;
  BHIP(B, hipEventCreate(&e0));
  BHIP(B, hipEventCreate(&e1));
  double best_ms = 1e300;
  bdem_batch_tuning best_t = keep;
  int rc = BDEM_OK;
  for (const bdem_batch_tuning& t : cand) {
    B->tune = t;
    if ((rc = bdem_batch_finalize(B)) != BDEM_OK) break;
    if ((rc = bdem_batch_step(B, 2)) != BDEM_OK) break;
    (void)hipEventRecord(e0, B->stream);
    if ((rc = bdem_batch_step(B, steps)) != BDEM_OK) break;
    (void)hipEventRecord(e1, B->stream);
    if (hipEventSynchronize(e1) != hipSuccess) { rc = bfail(B, BDEM_EHIP, "event wait failed"); break; }
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    const double per = double(ms) / steps;
    if (per < best_ms) { best_ms = per; best_t = t; }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != BDEM_OK) {
    B->tune = keep;
    return rc;
  }
  B->tune = best_t;
  if ((rc = bdem_batch_finalize(B)) != BDEM_OK) return rc;
  if (best) *best = best_t;
  if (ms_per_step) *ms_per_step = best_ms;
  return BDEM_OK;
}

int bdem_batch_num_instances(const bdem_batch* B) { return B ? (int)B->inst.size() : 0; }

int bdem_batch_num_particles(const bdem_batch* B, int index) {
  if (!B || index < 0 || index >= (int)B->inst.size()) return -1;
  return B->inst[(size_t)index].n;
}

int bdem_batch_instance_sim(const bdem_batch* B, int index, bdem_SimulationInfo* sim) {
  if (!B || !sim || index < 0 || index >= (int)B->inst.size()) return BDEM_EINVAL;
  bdem_batch* M = const_cast<bdem_batch*>(B);
  Instance& I = M->inst[(size_t)index];
  apply_cell_factor(I, B->finalized ? B->built_cell_factor : B->tune.cell_factor);
  *sim = I.sim_used;
  return BDEM_OK;
}

int bdem_batch_get_state(bdem_batch* B, int index, bdem_float3* pos, bdem_float3* vel, bdem_float3* avel, int* pid) {
  int rc = require_final(B);
  if (rc) return rc;
  if (index < 0 || index >= (int)B->inst.size()) return bfail(B, BDEM_EINVAL, "no such simulation");
  const Instance& I = B->inst[(size_t)index];
  BHIP(B, hipSetDevice(B->device));
  BHIP(B, hipStreamSynchronize(B->stream));
  std::vector<float4> P(I.cap), V(I.cap), O(I.cap);
  BHIP(B, hipMemcpy(P.data(), B->b.pos_cur + I.off, I.cap * sizeof(float4), hipMemcpyDeviceToHost));
  BHIP(B, hipMemcpy(V.data(), B->b.vel_cur + I.off, I.cap * sizeof(float4), hipMemcpyDeviceToHost));
  BHIP(B, hipMemcpy(O.data(), B->b.omg_cur + I.off, I.cap * sizeof(float4), hipMemcpyDeviceToHost));
  int k = 0;
  for (int s = 0; s < I.cap; ++s) {
    const int g = f_bits(O[s].w) >> 3;
    if (g < 0) continue;   /* padding */
    if (k >= I.n) return bfail(B, BDEM_ESTATE, "internal: more live slots than particles");
    if (pos) pos[k] = bdem_float3{P[s].x, P[s].y, P[s].z};
    if (vel) vel[k] = bdem_float3{V[s].x, V[s].y, V[s].z};
    if (avel) avel[k] = bdem_float3{O[s].x, O[s].y, O[s].z};
    if (pid) pid[k] = g - I.off;
    ++k;
  }
  if (k != I.n) return bfail(B, BDEM_ESTATE, "internal: a particle left its slot range");
  return BDEM_OK;
}

int bdem_batch_get_status(bdem_batch* B, unsigned int* w) {
  int rc = require_final(B);
  if (rc) return rc;
  BHIP(B, hipSetDevice(B->device));
  BHIP(B, hipStreamSynchronize(B->stream));
  uint32_t s = 0;
  BHIP(B, hipMemcpy(&s, B->b.status, sizeof(s), hipMemcpyDeviceToHost));
  launch_fill_u32(B->b.status, 1, 0u, B->stream);
  BHIP(B, hipStreamSynchronize(B->stream));
  if (w) *w = s;
  return BDEM_OK;
}

}  /* extern "C" */
