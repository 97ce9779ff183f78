/*
 * bdem_oracle.cpp - extern "C" surface of the CPU restatement, for the
 * tests and the CPU-baseline leg of the benchmark (loaded with ctypes).
 *
 * TEST INFRASTRUCTURE ONLY: the product library never links this.
 *
 * Handles carry their precision (32 = float parity definition, 64 = double
 * accuracy reference).  Outputs are returned as double so that both
 * precisions share one calling convention; float values convert exactly.
 * Stage queries (hash/sorted/cells/contacts/forces) work on a copy and leave
 * the state untouched, like the library's debug entry points.
 */
#include "bdem_oracle.hpp"
#include <chrono>
#include <memory>

using namespace bdem_oracle;

namespace {
struct Handle {
  int bits;
  std::unique_ptr<Oracle<float>> f;
  std::unique_ptr<Oracle<double>> d;
};

template <class R>
void put3(const std::vector<V3<R>>& v, double* out) {
  for (size_t i = 0; i < v.size(); ++i) {
    out[3 * i] = double(v[i].x); out[3 * i + 1] = double(v[i].y); out[3 * i + 2] = double(v[i].z);
  }
}

template <class R>
int get_state(Oracle<R>& o, double* p, double* v, double* w, double* q, uint32_t* t, int* id) {
  if (p) put3(o.pos, p);
  if (v) put3(o.vel, v);
  if (w) put3(o.avel, w);
  for (int i = 0; i < o.N; ++i) {
    if (q) {
      q[4 * i] = double(o.quat[i].w); q[4 * i + 1] = double(o.quat[i].x);
      q[4 * i + 2] = double(o.quat[i].y); q[4 * i + 3] = double(o.quat[i].z);
    }
    if (t) t[i] = o.ptype[i];
    if (id) id[i] = o.pid[i];
  }
  return 0;
}
}  // namespace

#define WITH(h, ...)                                        \
  do {                                                      \
    Handle* H_ = static_cast<Handle*>(h);                   \
    if (!H_) return -1;                                     \
    if (H_->bits == 32) { auto& o = *H_->f; __VA_ARGS__; }  \
    else { auto& o = *H_->d; __VA_ARGS__; }                 \
  } while (0)

extern "C" {

void* bo_create(int bits, const bdem_SimulationInfo* sim, const bdem_WOBJ* w, int nw,
                const bdem_POBJ* p, int np, const bdem_ext_params* ext) {
  if (bits != 32 && bits != 64) return nullptr;
  Handle* h = new Handle();
  h->bits = bits;
  if (bits == 32) {
    h->f.reset(new Oracle<float>());
    h->f->configure(sim, w, nw, p, np, ext);
  } else {
    h->d.reset(new Oracle<double>());
    h->d->configure(sim, w, nw, p, np, ext);
  }
  return h;
}

void bo_destroy(void* h) { delete static_cast<Handle*>(h); }

int bo_set_state(void* h, int n, const float* pos, const float* vel, const float* avel,
                 const float* quat, const uint32_t* ptype, const int* pid, const float* radius,
                 const float* mass) {
  WITH(h, o.set_state(n, pos, vel, avel, quat, ptype, pid, radius, mass));
  return 0;
}

int bo_get_state(void* h, double* pos, double* vel, double* avel, double* quat, uint32_t* ptype,
                 int* pid) {
  WITH(h, return get_state(o, pos, vel, avel, quat, ptype, pid));
  return 0;
}

int bo_num_particles(void* h) { WITH(h, return o.N); return 0; }
long long bo_num_cells(void* h) { WITH(h, return (long long)o.C); return 0; }

int bo_hash(void* h, uint32_t* keys, int* idx) {
  WITH(h, {
    auto c = o;
    c.compute_hash();
    for (int i = 0; i < c.N; ++i) { keys[i] = c.keys[i]; idx[i] = c.idx[i]; }
  });
  return 0;
}

int bo_sorted(void* h, uint32_t* keys, int* idx) {
  WITH(h, {
    auto c = o;
    c.compute_hash();
    c.sort_by_key();
    for (int i = 0; i < c.N; ++i) { keys[i] = c.keys[i]; idx[i] = c.idx[i]; }
  });
  return 0;
}

int bo_cells(void* h, uint32_t* cstart, uint32_t* cend) {
  WITH(h, {
    auto c = o;
    c.prepare();
    for (uint32_t i = 0; i < c.C; ++i) { cstart[i] = c.cell_start[i]; cend[i] = c.cell_end[i]; }
  });
  return 0;
}

int bo_contacts(void* h, uint32_t* list32, uint32_t* nn) {
  WITH(h, {
    auto c = o;
    c.prepare();
    std::vector<uint32_t> l, n;
    c.contact_list(l, n);
    std::copy(l.begin(), l.end(), list32);
    std::copy(n.begin(), n.end(), nn);
  });
  return 0;
}

/* Forces on the sorted state: pair force (with gravity), pair torque, wall
 * force, wall torque; each N x 3.  Also returns the sorted P_ID order. */
int bo_forces(void* h, double* fpp, double* tpp, double* fw, double* tw, int* sorted_pid) {
  WITH(h, {
    auto c = o;
    c.prepare();
    c.compute_forces();
    put3(c.f_pp, fpp); put3(c.t_pp, tpp); put3(c.f_w, fw); put3(c.t_w, tw);
    if (sorted_pid) for (int i = 0; i < c.N; ++i) sorted_pid[i] = c.pid[i];
  });
  return 0;
}

/* Lifter force and torque of the current state (sorted order), as bo_forces. */
int bo_lifter_forces(void* h, double* fl, double* tl, int* sorted_pid) {
  WITH(h, {
    auto c = o;
    c.prepare();
    c.compute_forces();
    put3(c.f_l, fl); put3(c.t_l, tl);
    if (sorted_pid) for (int i = 0; i < c.N; ++i) sorted_pid[i] = c.pid[i];
  });
  return 0;
}

int bo_step(void* h, int nsteps) {
  WITH(h, { for (int s = 0; s < nsteps; ++s) o.step(); });
  return 0;
}

/* Wall-clock seconds for nsteps (the CPU-baseline leg). */
double bo_time_steps(void* h, int nsteps) {
  auto t0 = std::chrono::steady_clock::now();
  bo_step(h, nsteps);
  auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count();
}

int bo_set_active_world_objects(void* h, int n) { WITH(h, o.num_wobj_active = n); return 0; }
int bo_set_drum_rotation(void* h, int on) { WITH(h, o.cylinder_rotation = on != 0); return 0; }
int bo_set_pair_override(void* h, double kn, double cn, double fk, double rr) {
  WITH(h, {
    o.pp_override = true;
    o.ov_kn = kn; o.ov_cn = cn; o.ov_fk = fk; o.ov_rr = rr;
  });
  return 0;
}
/* Per-particle polyhedron shapes: row p for P_ID p (bdem_set_poly_shapes). */
int bo_set_poly_shapes(void* h, const bdem_poly_shape* t, int count) {
  WITH(h, o.set_poly_shapes(t, count));
  return 0;
}
/* Per-particle lifter contact type kept between steps (Contact_Info::cont_type),
 * in the current (sorted) slot order.  The setter exists for tests that start
 * a particle with a remembered contact. */
int bo_lifter_types(void* h, int* out) {
  WITH(h, for (int i = 0; i < o.N; ++i) out[i] = o.lifter_type[i]);
  return 0;
}
int bo_set_lifter_types(void* h, const int* in) {
  WITH(h, for (int i = 0; i < o.N; ++i) o.lifter_type[i] = in[i]);
  return 0;
}
/* The lifter table (Device_Update_DObjectPostions): nd objects, 0 removes them. */
int bo_set_dobj(void* h, const bdem_DOBJ* d, int nd) {
  WITH(h, o.dobj.assign(d, d + (d ? nd : 0)));
  return 0;
}

/* Polyhedra that touched a lifter in the last step. */
int bo_lifter_contacts(void* h) { WITH(h, return o.lifter_contacts); return 0; }

/* The lifter table as it stands (after any turning). */
int bo_get_dobj(void* h, bdem_DOBJ* out) {
  WITH(h, for (size_t k = 0; k < o.dobj.size(); ++k) out[k] = o.dobj[k]);
  return 0;
}

/* Rotate_VolObject_Inc once (Input/Processing_Mills.h:7-60). */
int bo_turn_lifters(void* h, double rad, float ax, float ay, float az) {
  WITH(h, o.turn_lifters(rad, bdem_float3{ax, ay, az}));
  return 0;
}

/* The mill loop of UpdateSim (main.cpp:114-121, 202-204): turn the lifters
 * by rad before each of nsteps steps (rad = 0: no turning). */
int bo_run(void* h, int nsteps, double rad, float ax, float ay, float az) {
  WITH(h, {
    for (int s = 0; s < nsteps; ++s) {
      if (rad != 0.0) o.turn_lifters(rad, bdem_float3{ax, ay, az});
      o.step();
    }
  });
  return 0;
}

/* Tallies {particle, lifter, surface}: scaled 1E-4 as GetEnergy does
 * (GPU_Memory_Transfer.cuh:603-619) into e (float), and the same
 * increments summed in double, unscaled, into exact (may be NULL). */
int bo_energy(void* h, float* e, double* exact) {
  WITH(h, for (int k = 0; k < 3; ++k) {
    e[k] = float(double(o.tally[k]) * 1E-4);
    if (exact) exact[k] = o.tally_exact[k];
  });
  return 0;
}

/* Device_Set_Tallys / Device_Reset_Energy for the three tallies. */
int bo_set_tallys(void* h, const float* t) {
  WITH(h, for (int k = 0; k < 3; ++k) { o.tally[k] = t[k]; o.tally_exact[k] = double(t[k]); });
  return 0;
}

/* Bit 0: live particle outside the grid; bit 1: history overflow. */
int bo_warnings(void* h) {
  WITH(h, return (o.warn_out_of_grid ? 1 : 0) | (o.warn_hist_overflow ? 2 : 0));
  return 0;
}

/* Pair-history records of the particle with P_ID p: count, then (id, xi) rows. */
int bo_history(void* h, int p, int* ids, double* xi) {
  WITH(h, {
    if (p < 0 || p >= int(o.hcount.size())) return -1;
    int c = o.hcount[p];
    for (int k = 0; k < c; ++k) {
      const auto& r = o.hist[size_t(p) * HMAX + k];
      ids[k] = r.id;
      xi[3 * k] = double(r.xi.x); xi[3 * k + 1] = double(r.xi.y); xi[3 * k + 2] = double(r.xi.z);
    }
    return c;
  });
  return 0;
}

/* Polyhedra pair probe on the current slot order: the contact of slot a with
 * slot b (Kernel_ParticleInteraction_Polyhedra_NonSymmetry's per-partner
 * body) and the force and torque it puts on a.  out[16] = hit, class (1 VF,
 * 2 EF, 3 FF, 4 EE, 5 SP; +8 contact-error clamp, +16 speed clamp), depth,
 * normal, contact point, F, T. */
int bo_poly_pair(void* h, int a, int b, double* out) {
  WITH(h, {
    if (!o.poly() || a < 0 || b < 0 || a >= o.N || b >= o.N) return -1;
    const auto A = o.rotate_poly(a);
    const auto c = o.poly_pair_contact(a, b, A);
    auto F = c.n, T = c.n;
    F.x = F.y = F.z = 0; T = F;
    o.pair_force_poly(a, b, A, F, T);
    const double v[16] = {double(c.hit), double(c.cls), double(c.dis), double(c.n.x), double(c.n.y),
                          double(c.n.z), double(c.cp.x), double(c.cp.y), double(c.cp.z), double(F.x),
                          double(F.y), double(F.z), double(T.x), double(T.y), double(T.z), 0.0};
    for (int k = 0; k < 16; ++k) out[k] = v[k];
    return 0;
  });
  return 0;
}

/* Get_Angular_Acc for slot i and torque tau: R(q) I0^-1 R(q)^T tau. */
int bo_angular_acc(void* h, int i, const double* tau, double* out) {
  WITH(h, {
    if (!o.poly() || i < 0 || i >= o.N) return -1;
    using RR = std::remove_reference_t<decltype(o.pos[0].x)>;
    const auto a = o.angular_acc(i, mk<RR>(RR(tau[0]), RR(tau[1]), RR(tau[2])));
    out[0] = double(a.x); out[1] = double(a.y); out[2] = double(a.z);
    return 0;
  });
  return 0;
}

/* Polyhedra wall probe: the wall force and torque on slot i. */
int bo_poly_walls(void* h, int i, double* out) {
  WITH(h, {
    if (!o.poly() || i < 0 || i >= o.N) return -1;
    const auto A = o.rotate_poly(i);
    auto F = A.V[0], T = A.V[0];
    o.wall_forces_poly(i, A, F, T);
    out[0] = double(F.x); out[1] = double(F.y); out[2] = double(F.z);
    out[3] = double(T.x); out[4] = double(T.y); out[5] = double(T.z);
    return 0;
  });
  return 0;
}

/* History-law probe: one evaluation of the normal/tangential law for the
 * pair constants of types (ta, tb) (wall = 0) or the wall constants of type
 * ta (wall = 1), with normal n, overlap delta, contact velocity vc,
 * effective radius and mass, and the spring xi (in: the stored spring; out:
 * the spring the law keeps).  F receives the force. */
int bo_history_law(void* h, int wall, int ta, int tb, const double* n, double delta, const double* vc,
                   double rstar, double mstar, double* xi, double* F) {
  WITH(h, {
    if (ta < 0 || ta >= 8 || tb < 0 || tb >= 8 || o.law_kind() == BDEM_LAW_REFERENCE) return -1;
    using RR = std::remove_reference_t<decltype(o.pos[0].x)>;
    auto v3 = [](const double* a) { return mk<RR>(RR(a[0]), RR(a[1]), RR(a[2])); };
    auto x = v3(xi);
    const auto& lp = wall ? o.wlaw[ta] : o.law[ta][tb];
    const auto f = o.history_law(v3(n), RR(delta), v3(vc), RR(rstar), RR(mstar), lp, x);
    xi[0] = double(x.x); xi[1] = double(x.y); xi[2] = double(x.z);
    F[0] = double(f.x); F[1] = double(f.y); F[2] = double(f.z);
    return 0;
  });
  return 0;
}

}  // extern "C"
