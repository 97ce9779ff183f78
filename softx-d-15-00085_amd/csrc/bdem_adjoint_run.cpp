/*
 * bdem_adjoint_run.cpp - checkpointed adjoint runs (include/bdem_adjoint.h).
 *
 * Forward: the context steps as bdem_update_sim does; every
 * checkpoint_every steps its state (three float4 streams in the context's
 * current order) is copied aside and the next step fully re-sorts (the
 * re-sort gives the same order and cell table as the incremental one, so a
 * segment replays exactly from its checkpoint).  After every step a kernel
 * writes per-block moments of the state, summed on the host in block order.
 *
 * Backward: for each segment, last first, the checkpoint is restored and
 * the segment re-run with the step kernel writing its contact lists; after
 * each step its sorted input state and sort permutation are copied into the
 * segment's store.  Then the adjoint kernels run over the segment's steps
 * in reverse, carrying the state cotangent (three float4 per slot) from
 * each step's output order to its input order.
 */
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>
#include "bdem_adjoint.h"
#include "bdem_adjoint_k.hpp"
#include "bdem_ctx.hpp"
#include "bdem_kernels.hpp"

using namespace bdem;

struct bdem_adjoint {
  bdem_ctx* c = nullptr;
  int device = 0;                  /* the context's device, for freeing without touching the context */
  std::string err;
  int n = 0, nsteps = 0, window = 0, every = 0, nseg = 0, nblk = 0;
  bool have_theta = false, have_forward = false;
  double theta[4] = {0, 0, 0, 0};
  double dcn_de = 0.0, dcn_dkn = 0.0;
  float4 *init = nullptr;          /* 3 n: initial state */
  float4 *ckpt = nullptr;          /* (nseg) x 3 n */
  float4 *spos = nullptr, *svel = nullptr, *somg = nullptr;   /* every x n: a segment's sorted inputs */
  uint32_t *slist = nullptr, *snn = nullptr;                   /* every x n x 32, every x n */
  int* sidx = nullptr;                                          /* every x n */
  double* moments = nullptr;       /* nsteps x nblk x 6 */
  float4 *lam_a = nullptr, *lam_b = nullptr, *lamF = nullptr, *lamIn = nullptr;
  double *part_rr = nullptr, *part_pp = nullptr;
  std::vector<double> M;           /* nsteps x 6 moments of the last forward run */
  std::vector<void*> allocs;
  double fwd_ms = 0.0, bwd_ms = 0.0;
};

namespace bdem {
int ctx_upload_params(bdem_ctx* c);
}

namespace {

int afail(bdem_adjoint* A, int code, const std::string& msg) {
  if (A) A->err = msg;
  return code;
}

#define AHIP(A, expr)                                                               \
  do {                                                                              \
    hipError_t e_ = (expr);                                                         \
    if (e_ != hipSuccess)                                                           \
      return afail(A, BDEM_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <class T>
int aalloc(bdem_adjoint* A, T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc((void**)p, count * sizeof(T));
  if (e != hipSuccess) return afail(A, BDEM_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  A->allocs.push_back((void*)*p);
  return BDEM_OK;
}

void free_all(bdem_adjoint* A) {
  for (void* p : A->allocs) (void)hipFree(p);
  A->allocs.clear();
}

const double kPi = 3.14159265358979323846;

/* Restore three float4 streams of n into the context and make the next step
 * hash and fully sort. */
int restore(bdem_adjoint* A, const float4* src) {
  bdem_ctx* c = A->c;
  const size_t bytes = (size_t)A->n * sizeof(float4);
  AHIP(A, hipMemcpyAsync(c->b.pos_cur, src, bytes, hipMemcpyDeviceToDevice, c->stream));
  AHIP(A, hipMemcpyAsync(c->b.vel_cur, src + A->n, bytes, hipMemcpyDeviceToDevice, c->stream));
  AHIP(A, hipMemcpyAsync(c->b.omg_cur, src + 2 * (size_t)A->n, bytes, hipMemcpyDeviceToDevice, c->stream));
  c->keys_valid = false;
  c->flags_valid = false;
  c->stepc = 0;
  return BDEM_OK;
}

int save(bdem_adjoint* A, float4* dst) {
  bdem_ctx* c = A->c;
  const size_t bytes = (size_t)A->n * sizeof(float4);
  AHIP(A, hipMemcpyAsync(dst, c->b.pos_cur, bytes, hipMemcpyDeviceToDevice, c->stream));
  AHIP(A, hipMemcpyAsync(dst + A->n, c->b.vel_cur, bytes, hipMemcpyDeviceToDevice, c->stream));
  AHIP(A, hipMemcpyAsync(dst + 2 * (size_t)A->n, c->b.omg_cur, bytes, hipMemcpyDeviceToDevice, c->stream));
  return BDEM_OK;
}

int step_ctx(bdem_adjoint* A) {
  const int rc = bdem_update_sim(A->c);
  if (rc != BDEM_OK) return afail(A, rc, std::string("step: ") + bdem_last_error(A->c));
  return BDEM_OK;
}

/* Observables of one step's moments: centre of mass x, y and the
 * inclination of the bed's principal axis. */
void observe(const double* m, double out[3]) {
  const double M = m[0];
  const double cx = m[1] / M, cy = m[2] / M;
  const double cxx = m[3] / M - cx * cx, cyy = m[4] / M - cy * cy, cxy = m[5] / M - cx * cy;
  out[0] = cx;
  out[1] = cy;
  out[2] = 0.5 * std::atan2(2.0 * cxy, cxx - cyy);
}

/* d(observables)/d(moments 1..5) contracted with the cotangent g[3]. */
LossCoef coef(const double* m, const double g[3], double scale) {
  const double M = m[0];
  const double Sx = m[1], Sy = m[2];
  const double cx = Sx / M, cy = Sy / M;
  const double a = (m[3] / M - cx * cx) - (m[4] / M - cy * cy);
  const double b = 2.0 * (m[5] / M - cx * cy);
  const double q = a * a + b * b;
  const double ta = q > 0.0 ? -0.5 * b / q : 0.0;   /* d theta / d a */
  const double tb = q > 0.0 ? 0.5 * a / q : 0.0;    /* d theta / d b */
  /* a = Sxx/M - Sx^2/M^2 - Syy/M + Sy^2/M^2, b = 2 Sxy/M - 2 Sx Sy/M^2 */
  const double dSx = g[0] / M + g[2] * (ta * (-2.0 * Sx / (M * M)) + tb * (-2.0 * Sy / (M * M)));
  const double dSy = g[1] / M + g[2] * (ta * (2.0 * Sy / (M * M)) + tb * (-2.0 * Sx / (M * M)));
  const double dSxx = g[2] * ta / M;
  const double dSyy = -g[2] * ta / M;
  const double dSxy = g[2] * tb * 2.0 / M;
  return LossCoef{float(dSx * scale), float(dSy * scale), float(dSxx * scale), float(dSyy * scale),
                  float(dSxy * scale)};
}

}  // namespace

extern "C" {

int bdem_adj_create(bdem_adjoint** out, bdem_ctx* c, int nsteps, int window, int every) {
  if (!out) return BDEM_EINVAL;
  *out = nullptr;
  if (!c || !c->configured) return BDEM_EINVAL;
  if (nsteps <= 0 || window <= 0 || every <= 0 || nsteps % window || nsteps % every) return BDEM_EINVAL;
  if (c->poly || c->orient) return BDEM_ENOTSUP;
  if (c->has_ext && (c->ext.contact_law != BDEM_LAW_REFERENCE || c->ext.num_cones > 0)) return BDEM_ENOTSUP;
  if (!c->dobj.empty() || c->pobj.size() != 1 || c->sim.Kill_Particles || c->b.skin > 0.0f) return BDEM_ENOTSUP;
  if (c->wobj.empty() || !c->cyl_rotation) return BDEM_ENOTSUP;
  for (const bdem_WOBJ& w : c->wobj)
    if (w.surface_type != BDEM_SURFACE_CYLINDER) return BDEM_ENOTSUP;
  if (c->flags_valid) return BDEM_ESTATE;
  bdem_adjoint* A = new bdem_adjoint();
  A->c = c;
  A->device = c->device;
  A->n = c->b.n;
  A->nsteps = nsteps;
  A->window = window;
  A->every = every;
  A->nseg = nsteps / every;
  A->nblk = (A->n + ADJ_THREADS - 1) / ADJ_THREADS;
  const size_t n = (size_t)A->n;
  int rc = BDEM_OK;
  if (hipSetDevice(c->device) != hipSuccess) rc = BDEM_EHIP;
#define AALLOC(p, cnt) if (rc == BDEM_OK) rc = aalloc(A, &(p), (cnt))
  AALLOC(A->init, 3 * n);
  AALLOC(A->ckpt, (size_t)A->nseg * 3 * n);
  AALLOC(A->spos, (size_t)every * n);
  AALLOC(A->svel, (size_t)every * n);
  AALLOC(A->somg, (size_t)every * n);
  AALLOC(A->slist, (size_t)every * n * LIST_STRIDE);
  AALLOC(A->snn, (size_t)every * n);
  AALLOC(A->sidx, (size_t)every * n);
  AALLOC(A->moments, (size_t)nsteps * A->nblk * ADJ_MOMENTS);
  AALLOC(A->lam_a, 3 * n);
  AALLOC(A->lam_b, 3 * n);
  AALLOC(A->lamF, 2 * n);
  AALLOC(A->lamIn, 3 * n);
  AALLOC(A->part_rr, (size_t)A->nblk);
  AALLOC(A->part_pp, (size_t)A->nblk * 3);
#undef AALLOC
  if (rc == BDEM_OK) rc = save(A, A->init);
  if (rc == BDEM_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = BDEM_EHIP;
  if (rc != BDEM_OK) {
    free_all(A);
    delete A;
    return rc;
  }
  *out = A;
  return BDEM_OK;
}

int bdem_adj_destroy(bdem_adjoint* A) {
  if (!A) return BDEM_EINVAL;
  /* the context may already be gone (a binding's finalizers can run in any
   * order): only the run's own buffers are touched; hipFree waits for the
   * device */
  (void)hipSetDevice(A->device);
  free_all(A);
  delete A;
  return BDEM_OK;
}

const char* bdem_adj_last_error(const bdem_adjoint* A) { return A ? A->err.c_str() : "null run"; }

int bdem_adj_num_observables(const bdem_adjoint* A) { return A ? 3 * (A->nsteps / A->window) : 0; }

int bdem_adj_set_params(bdem_adjoint* A, const double* th) {
  if (!A || !th) return BDEM_EINVAL;
  const double mu = th[0], e = th[1], kn = th[2], rr = th[3];
  if (!(mu >= 0.0) || !(e > 0.0 && e < 1.0) || !(kn > 0.0) || !(rr >= 0.0 && rr < 1.0))
    return afail(A, BDEM_EINVAL, "theta out of range (mu >= 0, 0 < e < 1, kn > 0, 0 <= rr < 1)");
  bdem_ctx* c = A->c;
  bdem_POBJ& p = c->pobj[0];
  /* Set_SpringConstantsParticleKN for the same type: reduced mass m/2, the
   * restitution shifted by 0.03808 inside [0.3, 0.9] */
  const double m = 1.0 / ((1.0 / p.mass) + (1.0 / p.mass));
  const double es = (e >= 0.30 && e <= 0.90) ? e - 0.038080 : e;
  p.PP[0].Kn = float(kn);
  const double K = p.PP[0].Kn;
  const double l = std::log(es), D = std::sqrt(l * l + kPi * kPi);
  const double cn = -2.0 * l * std::sqrt(K * m) / D;
  p.PP[0].Cn = float(cn);
  p.PP[0].Fric_kinetic = float(mu);
  c->sim.Roll_Res = float(rr);
  /* dCn/dKn = Cn / 2Kn; dCn/de = -2 sqrt(Kn m) pi^2 / (D^3 e') */
  A->dcn_dkn = cn / (2.0 * K);
  A->dcn_de = -2.0 * std::sqrt(K * m) * kPi * kPi / (D * D * D * es);
  std::memcpy(A->theta, th, sizeof(A->theta));
  A->have_theta = true;
  A->have_forward = false;
  const int rc = ctx_upload_params(c);
  if (rc != BDEM_OK) return afail(A, rc, bdem_last_error(c));
  return BDEM_OK;
}

int bdem_adj_forward(bdem_adjoint* A, double* obs) {
  if (!A) return BDEM_EINVAL;
  if (!A->have_theta) return afail(A, BDEM_ESTATE, "bdem_adj_set_params has not been called");
  bdem_ctx* c = A->c;
  AHIP(A, hipSetDevice(c->device));
  c->dbg_list = nullptr;
  c->dbg_nn = nullptr;
  int rc;
  if ((rc = restore(A, A->init)) != BDEM_OK) return rc;
  hipEvent_t e0, e1;
  AHIP(A, hipEventCreate(&e0));
  AHIP(A, hipEventCreate(&e1));
  (void)hipEventRecord(e0, c->stream);
  for (int s = 0; s < A->nsteps && rc == BDEM_OK; ++s) {
    if (s % A->every == 0) {
      rc = save(A, A->ckpt + (size_t)(s / A->every) * 3 * A->n);
      c->keys_valid = false;
      c->flags_valid = false;
    }
    if (rc == BDEM_OK) rc = step_ctx(A);
    if (rc == BDEM_OK)
      launch_adj_moments(A->n, c->b.pos_cur, c->b.vel_cur, c->b.omg_cur,
                         A->moments + (size_t)s * A->nblk * ADJ_MOMENTS, c->stream);
  }
  (void)hipEventRecord(e1, c->stream);
  hipError_t he = hipEventSynchronize(e1);
  float ms = 0.f;
  if (he == hipSuccess) (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != BDEM_OK) return rc;
  if (he != hipSuccess) return afail(A, BDEM_EHIP, std::string("forward: ") + hipGetErrorString(he));
  AHIP(A, hipGetLastError());
  A->fwd_ms = ms;
  std::vector<double> part((size_t)A->nsteps * A->nblk * ADJ_MOMENTS);
  AHIP(A, hipMemcpy(part.data(), A->moments, part.size() * sizeof(double), hipMemcpyDeviceToHost));
  A->M.assign((size_t)A->nsteps * ADJ_MOMENTS, 0.0);
  for (int s = 0; s < A->nsteps; ++s)
    for (int b = 0; b < A->nblk; ++b)
      for (int k = 0; k < ADJ_MOMENTS; ++k)
        A->M[(size_t)s * ADJ_MOMENTS + k] += part[((size_t)s * A->nblk + b) * ADJ_MOMENTS + k];
  const int nw = A->nsteps / A->window;
  for (int w = 0; w < nw; ++w) {
    double acc[3] = {0, 0, 0};
    for (int s = w * A->window; s < (w + 1) * A->window; ++s) {
      double o[3];
      observe(&A->M[(size_t)s * ADJ_MOMENTS], o);
      for (int k = 0; k < 3; ++k) acc[k] += o[k];
    }
    if (obs)
      for (int k = 0; k < 3; ++k) obs[3 * w + k] = acc[k] / A->window;
  }
  A->have_forward = true;
  return BDEM_OK;
}

int bdem_adj_backward(bdem_adjoint* A, const double* cot, double* dtheta) {
  if (!A || !cot) return BDEM_EINVAL;
  if (!A->have_forward) return afail(A, BDEM_ESTATE, "bdem_adj_forward has not been run with these parameters");
  bdem_ctx* c = A->c;
  AHIP(A, hipSetDevice(c->device));
  const size_t n = (size_t)A->n;
  std::vector<LossCoef> g((size_t)A->nsteps);
  for (int s = 0; s < A->nsteps; ++s) {
    const int w = s / A->window;
    g[(size_t)s] = coef(&A->M[(size_t)s * ADJ_MOMENTS], cot + 3 * w, 1.0 / A->window);
  }
  hipEvent_t e0, e1;
  AHIP(A, hipEventCreate(&e0));
  AHIP(A, hipEventCreate(&e1));
  (void)hipEventRecord(e0, c->stream);
  launch_fill_u32(reinterpret_cast<uint32_t*>(A->lam_a), 3 * n * 4, 0u, c->stream);
  launch_fill_u32(reinterpret_cast<uint32_t*>(A->part_rr), (size_t)A->nblk * 2, 0u, c->stream);
  launch_fill_u32(reinterpret_cast<uint32_t*>(A->part_pp), (size_t)A->nblk * 3 * 2, 0u, c->stream);
  float4* lam_cur = A->lam_a;
  float4* lam_nxt = A->lam_b;
  int rc = BDEM_OK;
  /* the step kernel writes its contact lists only while these are set */
  struct Unrecord {
    bdem_ctx* c;
    ~Unrecord() { c->dbg_list = nullptr; c->dbg_nn = nullptr; }
  } unrecord{c};
  for (int seg = A->nseg - 1; seg >= 0 && rc == BDEM_OK; --seg) {
    rc = restore(A, A->ckpt + (size_t)seg * 3 * n);
    for (int t = 0; t < A->every && rc == BDEM_OK; ++t) {
      c->dbg_list = A->slist + (size_t)t * n * LIST_STRIDE;
      c->dbg_nn = A->snn + (size_t)t * n;
      rc = step_ctx(A);
      if (rc != BDEM_OK) break;
      const size_t bytes = n * sizeof(float4);
      AHIP(A, hipMemcpyAsync(A->spos + (size_t)t * n, c->b.pos_srt, bytes, hipMemcpyDeviceToDevice, c->stream));
      AHIP(A, hipMemcpyAsync(A->svel + (size_t)t * n, c->b.vel_srt, bytes, hipMemcpyDeviceToDevice, c->stream));
      AHIP(A, hipMemcpyAsync(A->somg + (size_t)t * n, c->b.omg_srt, bytes, hipMemcpyDeviceToDevice, c->stream));
      AHIP(A, hipMemcpyAsync(A->sidx + (size_t)t * n, c->b.idx_sorted, n * sizeof(int), hipMemcpyDeviceToDevice,
                             c->stream));
    }
    c->dbg_list = nullptr;
    c->dbg_nn = nullptr;
    for (int t = A->every - 1; t >= 0 && rc == BDEM_OK; --t) {
      const int s = seg * A->every + t;
      AdjStep st{A->n, A->spos + (size_t)t * n, A->svel + (size_t)t * n, A->somg + (size_t)t * n,
                 A->slist + (size_t)t * n * LIST_STRIDE, A->snn + (size_t)t * n, A->sidx + (size_t)t * n};
      launch_adj_back(c->dp, st, lam_cur, A->lamF, A->lamIn, lam_nxt, g[(size_t)s], A->part_rr, A->part_pp,
                      c->stream);
      std::swap(lam_cur, lam_nxt);
    }
  }
  (void)hipEventRecord(e1, c->stream);
  hipError_t he = hipEventSynchronize(e1);
  float ms = 0.f;
  if (he == hipSuccess) (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != BDEM_OK) return rc;
  if (he != hipSuccess) return afail(A, BDEM_EHIP, std::string("backward: ") + hipGetErrorString(he));
  AHIP(A, hipGetLastError());
  A->bwd_ms = ms;
  unsigned int warn = 0;
  if ((rc = bdem_get_status(c, &warn)) != BDEM_OK) return afail(A, rc, bdem_last_error(c));
  if (warn & BDEM_WARN_LIST_OVERFLOW)
    return afail(A, BDEM_ENOTSUP, "a particle had more than 32 contacts: the recorded lists are incomplete");
  std::vector<double> prr((size_t)A->nblk), ppp((size_t)A->nblk * 3);
  AHIP(A, hipMemcpy(prr.data(), A->part_rr, prr.size() * sizeof(double), hipMemcpyDeviceToHost));
  AHIP(A, hipMemcpy(ppp.data(), A->part_pp, ppp.size() * sizeof(double), hipMemcpyDeviceToHost));
  double gkn = 0, gcn = 0, gfk = 0, grr = 0;
  for (int b = 0; b < A->nblk; ++b) {
    grr += prr[(size_t)b];
    gkn += ppp[(size_t)b * 3];
    gcn += ppp[(size_t)b * 3 + 1];
    gfk += ppp[(size_t)b * 3 + 2];
  }
  if (dtheta) {
    dtheta[0] = gfk;
    dtheta[1] = gcn * A->dcn_de;
    dtheta[2] = gkn + gcn * A->dcn_dkn;
    dtheta[3] = grr;
  }
  return BDEM_OK;
}

int bdem_adj_timing(const bdem_adjoint* A, double* f, double* b) {
  if (!A) return BDEM_EINVAL;
  if (f) *f = A->fwd_ms;
  if (b) *b = A->bwd_ms;
  return BDEM_OK;
}

}  /* extern "C" */
