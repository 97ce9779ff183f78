/*
 * bdem_oracle.hpp - CPU restatement of the reference GPU step loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the HIP
 * library: only the tests, the smoke check and the CPU-baseline leg of the
 * benchmark load it.  Nothing in the product links or calls it.
 *
 * One template, instantiated for float (the parity definition, built with
 * -ffp-contract=off -fno-fast-math so every operation is a single IEEE
 * rounding in source order) and for double (the accuracy reference).
 *
 * Stage order follows Device_DEM_UpdateSim (Device/DeviceInterface.cu:1825-2513):
 *   hash -> stable sort -> reorder -> cell bounds -> contact list ->
 *   pair force (+gravity) -> wall force -> symplectic Euler -> discharge check.
 * Each function names the reference lines whose arithmetic it restates.
 *
 * Conventions shared with the HIP library (defined there, restated here):
 *   - keys: in-grid live particles get the reference hash; a live particle
 *     outside the grid gets KEY_OUT = C, a dead one (P_ID < 0) KEY_DEAD = C+1,
 *     so both sort behind every cell and are invisible to the cell tables.
 *   - the 27-cell stencil is clamped to the grid (identical to the reference
 *     for in-grid particles, Collision_Detection_BroadPhase.cuh:428-459).
 *   - every live particle is processed each step; dead ones are inert.
 *   - history laws keep tangential springs keyed by the partner's P_ID.
 */
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <numeric>
#include "bdem_types.h"

namespace bdem_oracle {

constexpr int LIST_STRIDE = 32;   /* Broad_List stride, Collision_Detection_BroadPhase.cuh:481 */
constexpr int HMAX = 16;          /* pair history records per particle */
constexpr int WMAX = 4;           /* wall history records per particle */

template <class R> struct V3 {
  R x, y, z;
};
template <class R> inline V3<R> mk(R x, R y, R z) { return V3<R>{x, y, z}; }
template <class R> inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> inline V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> inline V3<R> operator*(R s, V3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class R> inline V3<R> operator*(V3<R> a, V3<R> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class R> inline V3<R> operator/(V3<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }
/* Utilities/Device/Functions.cuh:198-201: x*x + y*y + z*z, left to right */
template <class R> inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
/* Utilities/Device/Functions.cuh:370-373 */
template <class R> inline V3<R> cross(V3<R> a, V3<R> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline float rsqrt_(float v) { return std::sqrt(v); }
inline double rsqrt_(double v) { return std::sqrt(v); }
/* Utilities/Device/Functions.cuh:241-244: sqrtf(dot(v,v)) */
template <class R> inline R length(V3<R> a) { return rsqrt_(dot(a, a)); }
template <class R> inline V3<R> cvt(bdem_float3 f) { return {R(f.x), R(f.y), R(f.z)}; }

template <class R> struct Q4 { R w, x, y, z; };

/* Quaternion product, Utilities/Device/Opertators.cuh:866-874 */
template <class R> inline Q4<R> qmul(Q4<R> a, Q4<R> b) {
  Q4<R> q;
  q.x = (a.x * b.w) + (a.y * b.z) - (a.z * b.y) + (a.w * b.x);
  q.y = -(a.x * b.z) + (a.y * b.w) + (a.z * b.x) + (a.w * b.y);
  q.z = (a.x * b.y) - (a.y * b.x) + (a.z * b.w) + (a.w * b.z);
  q.w = -(a.x * b.x) - (a.y * b.y) - (a.z * b.z) + (a.w * b.w);
  return q;
}
/* normaliseD, Utilities/Device/Functions.cuh:399-420 */
template <class R> inline Q4<R> qnormalise(Q4<R> a) {
  R s = a.w * a.w + a.x * a.x + a.y * a.y + a.z * a.z;
  R n = s > R(0) ? R(1) / rsqrt_(s) : R(0);
  return Q4<R>{a.w * n, a.x * n, a.y * n, a.z * n};
}

template <class R> struct PairHist { int id; V3<R> xi; };

/* Per type-pair constants for the history laws (host-derived, see derive_laws). */
template <class R> struct LawPair {
  R kn, kt, beta, mu;     /* LSD */
  R estar, gstar;         /* Hertz-Mindlin */
  R tg;                   /* Hertz-Mindlin: tangential over normal damping, sqrt(4 G* / E*) */
};

/* sqrt(St / Sn) = sqrt(8 G* / (2 E*)) from the float constants, in double */
inline float damping_ratio_t(float estar, float gstar) {
  return estar > 0.0f ? float(std::sqrt(4.0 * double(gstar) / double(estar))) : 0.0f;
}

/* The geometric fields of a POBJ row as a shape row (the per-type table). */
inline bdem_poly_shape shape_of_pobj(const bdem_POBJ& po) {
  bdem_poly_shape s;
  std::memset(&s, 0, sizeof(s));
  s.radius = po.radius;
  s.mass = po.mass;
  s.COM = po.COM;
  s.num_vertex = po.num_vertex;
  std::memcpy(s.Local_Vertex, po.Local_Vertex, sizeof(po.Local_Vertex));
  s.num_edges = po.num_edges;
  std::memcpy(s.edge, po.edge, sizeof(po.edge));
  s.num_faces = po.num_faces;
  std::memcpy(s.face, po.face, sizeof(po.face));
  std::memcpy(s.InertiaT, po.InertiaT, sizeof(po.InertiaT));
  return s;
}

template <class R>
class Oracle {
 public:
  /* --- configuration (the constant tables of GPU_Memory_Transfer.cuh:78-85) --- */
  bdem_SimulationInfo sim{};
  std::vector<bdem_WOBJ> wobj;
  std::vector<bdem_POBJ> pobj;
  /* polyhedra: geometry rows, one per type (the POBJ rows) or, after
   * set_poly_shapes, one per P_ID; the type still selects the constants */
  std::vector<bdem_poly_shape> shapes;
  bool shape_by_pid = false;
  std::vector<bdem_DOBJ> dobj;      /* lifters (DynamicObject[32]), host-rotated */
  bdem_ext_params ext{};
  bool has_ext = false;
  int num_wobj_active = 0;           /* Num_WorldObjects, decremented by Hatch */
  bool cylinder_rotation = false;    /* is_cylinder_rotation (drum_rotation) */
  /* Exact (double) pair constants of type pair (0, 0) and rolling
   * resistance, replacing the float table values when set: finite
   * differences of the FP64 restatement in these parameters need
   * perturbations finer than a float's spacing. */
  bool pp_override = false;
  double ov_kn = 0, ov_cn = 0, ov_fk = 0, ov_rr = 0;
  uint32_t C = 0;                    /* number of grid cells */
  LawPair<R> law[8][8]{};
  LawPair<R> wlaw[8]{};

  /* --- state, in current (sorted) order --- */
  int N = 0;
  std::vector<V3<R>> pos, vel, avel;
  std::vector<Q4<R>> quat;
  std::vector<R> radius, mass, inv_inertia;
  std::vector<uint32_t> ptype;
  std::vector<int> pid;
  std::vector<int> lifter_type;      /* Contact_Info::cont_type of the last lifter check, travels with the particle */
  /* history, indexed by P_ID */
  std::vector<int> hcount, wcount;
  std::vector<PairHist<R>> hist, whist;

  /* --- per-step work arrays --- */
  std::vector<uint32_t> keys;
  std::vector<int> idx;
  std::vector<uint32_t> cell_start, cell_end;   /* reference format, 0xFFFFFFFF empty */
  std::vector<V3<R>> f_pp, t_pp, f_w, t_w, f_l, t_l;
  int stepc = 0;
  long long steps_done = 0;
  /* energy-dissipation tallies {particle, lifter, surface}
   * (Tally_EnergyDis_PP / DisLift / DisSurf, GPU_Memory_Transfer.cuh:97-99):
   * every contact's increment is added on its own to a float total, as the
   * reference's atomicAdd does, here in particle order; `exact` sums the
   * same increments in double */
  enum { T_PP = 0, T_LIFT = 1, T_SURF = 2 };
  float tally[3] = {0.f, 0.f, 0.f};
  double tally_exact[3] = {0.0, 0.0, 0.0};
  mutable std::vector<std::vector<R>> incs[3];   /* this step's increments, per particle */
  int lifter_contacts = 0;           /* polyhedra in contact with a lifter in the last step */
  int warn_out_of_grid = 0;
  int warn_list_overflow = 0;
  int warn_hist_overflow = 0;

  void configure(const bdem_SimulationInfo* s, const bdem_WOBJ* w, int nw, const bdem_POBJ* p,
                 int np, const bdem_ext_params* e) {
    sim = *s;
    wobj.assign(w, w + nw);
    pobj.assign(p, p + np);
    shapes.clear();
    for (int t = 0; t < np; ++t) shapes.push_back(shape_of_pobj(p[t]));
    shape_by_pid = false;
    num_wobj_active = nw;
    has_ext = e != nullptr;
    if (e) ext = *e; else std::memset(&ext, 0, sizeof(ext));
    C = uint32_t(sim.num_NNCells.x) * uint32_t(sim.num_NNCells.y) * uint32_t(sim.num_NNCells.z);
    derive_laws();
  }

  int law_kind() const { return has_ext ? ext.contact_law : BDEM_LAW_REFERENCE; }

  /* The geometry row of particle i (a dead particle, P_ID -1, takes row 0). */
  const bdem_poly_shape& shp(int i) const {
    return shapes[shape_by_pid ? (pid[i] > 0 ? pid[i] : 0) : int(ptype[i])];
  }
  /* Per-particle shapes: row p for P_ID p.  Rewrites the radius and mass
   * of the current state's particles from their rows. */
  void set_poly_shapes(const bdem_poly_shape* t, int count) {
    shapes.assign(t, t + count);
    shape_by_pid = true;
    for (int i = 0; i < N; ++i) {
      radius[i] = R(shp(i).radius);
      mass[i] = R(shp(i).mass);
    }
  }

  /* Pair and wall constants of the history laws.  Restated from the
   * textbook forms (no reference counterpart):
   *   beta = ln e / sqrt(ln^2 e + pi^2)
   *   Hertz: 1/E* = (1-nu_i^2)/E_i + (1-nu_j^2)/E_j,
   *          1/G* = 2(2-nu_i)(1+nu_i)/E_i + 2(2-nu_j)(1+nu_j)/E_j
   * Pair values average e, mu, kn, kt/kn of the two types.  Computed in
   * double, stored in R (the HIP library stores the same float values). */
  void derive_laws() {
    const double pi = 3.14159265358979323846;
    for (int a = 0; a < 8; ++a) {
      for (int b = 0; b < 8; ++b) {
        const bdem_material& ma = ext.material[a];
        const bdem_material& mb = ext.material[b];
        double e = 0.5 * (double(ma.restitution) + double(mb.restitution));
        double beta = 0.0;
        if (e > 0.0 && e < 1.0) beta = std::log(e) / std::sqrt(std::log(e) * std::log(e) + pi * pi);
        double kn = 0.5 * (double(ma.kn) + double(mb.kn));
        double ktr = 0.5 * (double(ma.kt_over_kn) + double(mb.kt_over_kn));
        double mu = 0.5 * (double(ma.mu) + double(mb.mu));
        double ea = ma.youngs, eb = mb.youngs, na = ma.poisson, nb = mb.poisson;
        double estar = 0.0, gstar = 0.0;
        if (ea > 0 && eb > 0) {
          estar = 1.0 / ((1.0 - na * na) / ea + (1.0 - nb * nb) / eb);
          gstar = 1.0 / (2.0 * (2.0 - na) * (1.0 + na) / ea + 2.0 * (2.0 - nb) * (1.0 + nb) / eb);
        }
        law[a][b] = LawPair<R>{R(float(kn)), R(float(kn * ktr)), R(float(beta)), R(float(mu)),
                               R(float(estar)), R(float(gstar)),
                               R(damping_ratio_t(float(estar), float(gstar)))};
      }
      const bdem_material& ma = ext.material[a];
      double e = ma.restitution;
      double beta = 0.0;
      if (e > 0.0 && e < 1.0) beta = std::log(e) / std::sqrt(std::log(e) * std::log(e) + pi * pi);
      double ew = ext.wall_youngs > 0 ? ext.wall_youngs : ma.youngs;
      double nw = ext.wall_youngs > 0 ? ext.wall_poisson : ma.poisson;
      double estar = 0.0, gstar = 0.0;
      if (ma.youngs > 0 && ew > 0) {
        estar = 1.0 / ((1.0 - ma.poisson * ma.poisson) / ma.youngs + (1.0 - nw * nw) / ew);
        gstar = 1.0 / (2.0 * (2.0 - ma.poisson) * (1.0 + ma.poisson) / ma.youngs +
                       2.0 * (2.0 - nw) * (1.0 + nw) / ew);
      }
      wlaw[a] = LawPair<R>{R(ma.kn), R(float(double(ma.kn) * ma.kt_over_kn)), R(float(beta)),
                           R(ma.mu), R(float(estar)), R(float(gstar)),
                           R(damping_ratio_t(float(estar), float(gstar)))};
    }
  }

  /* Device_Set_System_State / Set_PSystem_State (GPU_Memory_Transfer.cuh:397-440).
   * radius_in/mass_in are the per-particle extension; without them the type
   * table supplies both, exactly as the reference does. */
  void set_state(int n, const float* p3, const float* v3, const float* w3, const float* q4,
                 const uint32_t* t, const int* id, const float* radius_in, const float* mass_in) {
    N = n;
    pos.resize(n); vel.resize(n); avel.resize(n); quat.resize(n);
    radius.resize(n); mass.resize(n); inv_inertia.resize(n); ptype.resize(n); pid.resize(n);
    lifter_type.assign(n, 0);
    for (int i = 0; i < n; ++i) {
      pos[i] = mk<R>(p3[3 * i], p3[3 * i + 1], p3[3 * i + 2]);
      vel[i] = v3 ? mk<R>(v3[3 * i], v3[3 * i + 1], v3[3 * i + 2]) : mk<R>(0, 0, 0);
      avel[i] = w3 ? mk<R>(w3[3 * i], w3[3 * i + 1], w3[3 * i + 2]) : mk<R>(0, 0, 0);
      quat[i] = q4 ? Q4<R>{R(q4[4 * i]), R(q4[4 * i + 1]), R(q4[4 * i + 2]), R(q4[4 * i + 3])}
                   : Q4<R>{1, 0, 0, 0};
      ptype[i] = t[i];
      pid[i] = id[i];
      const bdem_POBJ& po = pobj[t[i]];
      if (radius_in) {
        radius[i] = R(radius_in[i]);
        mass[i] = R(mass_in[i]);
        /* per-particle sphere inverse inertia, same float formula as
         * main.cpp:615-616 evaluated per particle */
        inv_inertia[i] = R(1) / (R(0.4f) * radius[i] * radius[i] * mass[i]);
      } else if (shape_by_pid) {
        const bdem_poly_shape& sh = shapes[id[i] > 0 ? id[i] : 0];
        radius[i] = R(sh.radius);
        mass[i] = R(sh.mass);
        inv_inertia[i] = R(sh.InertiaT[0]);
      } else {
        radius[i] = R(po.radius);
        mass[i] = R(po.mass);
        inv_inertia[i] = R(po.InertiaT[0]);
      }
    }
    int maxid = 0;
    for (int i = 0; i < n; ++i) maxid = std::max(maxid, pid[i] + 1);
    hcount.assign(std::max(maxid, n), 0);
    wcount.assign(std::max(maxid, n), 0);
    hist.assign(size_t(std::max(maxid, n)) * HMAX, PairHist<R>{-1, mk<R>(0, 0, 0)});
    whist.assign(size_t(std::max(maxid, n)) * WMAX, PairHist<R>{-1, mk<R>(0, 0, 0)});
    f_pp.assign(n, mk<R>(0, 0, 0)); t_pp = f_pp; f_w = f_pp; t_w = f_pp; f_l = f_pp; t_l = f_pp;
  }

  /* ------------------------------------------------------------------ */
  /* a1: grid position and hash                                          */
  /* SpatialDecomp_CalcGridPos  Collision_Detection_BroadPhase.cuh:20-29 */
  /* SpatialDecomp_CalcGridHash Collision_Detection_BroadPhase.cuh:42-50 */
  /* ------------------------------------------------------------------ */
  void grid_pos(V3<R> p, int g[3]) const {
    g[0] = int(std::floor((p.x - R(sim.worldOrigin.x)) / R(sim.cellSize.x)));
    g[1] = int(std::floor((p.y - R(sim.worldOrigin.y)) / R(sim.cellSize.y)));
    g[2] = int(std::floor((p.z - R(sim.worldOrigin.z)) / R(sim.cellSize.z)));
  }
  uint32_t hash_of(int gx, int gy, int gz) const {
    /* (x + z*Nz) + (y*Nx)*Nz in 32-bit two's-complement arithmetic */
    uint32_t nx = uint32_t(sim.num_NNCells.x), nz = uint32_t(sim.num_NNCells.z);
    return (uint32_t(gx) + uint32_t(gz) * nz) + (uint32_t(gy) * nx) * nz;
  }
  bool in_grid(const int g[3]) const {
    return g[0] >= 0 && g[1] >= 0 && g[2] >= 0 && g[0] < sim.num_NNCells.x &&
           g[1] < sim.num_NNCells.y && g[2] < sim.num_NNCells.z;
  }
  uint32_t key_of(int i) const {
    if (pid[i] < 0) return C + 1;                  /* KEY_DEAD */
    int g[3];
    grid_pos(pos[i], g);
    if (!in_grid(g)) return C;                     /* KEY_OUT */
    uint32_t h = hash_of(g[0], g[1], g[2]);
    return h < C ? h : C;                          /* aliasing past the table: KEY_OUT */
  }
  /* Kernel_SpatialDecomp_CalcPHash, Collision_Detection_BroadPhase.cuh:75-100 */
  void compute_hash() {
    keys.resize(N); idx.resize(N);
    for (int i = 0; i < N; ++i) {
      keys[i] = key_of(i);
      idx[i] = i;
      if (keys[i] == C) warn_out_of_grid = 1;
    }
  }

  /* a2: thrust::stable_sort_by_key (Collision_Detection_BroadPhase.cuh:110-117):
   * ascending key, ties keep ascending original index. */
  void sort_by_key() {
    std::vector<int> order(N);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return keys[a] < keys[b]; });
    std::vector<uint32_t> k2(N);
    std::vector<int> i2(N);
    for (int i = 0; i < N; ++i) { k2[i] = keys[order[i]]; i2[i] = idx[order[i]]; }
    keys.swap(k2); idx.swap(i2);
  }

  /* a3: Kernel_SortArrays + Kernel_SpatialDecomp_ReorderPDA
   * (Collision_Detection_BroadPhase.cuh:126-268): gather by sorted index. */
  template <class T> void permute(std::vector<T>& v) const {
    std::vector<T> t(N);
    for (int i = 0; i < N; ++i) t[i] = v[idx[i]];
    v.swap(t);
  }
  void reorder() {
    permute(pos); permute(vel); permute(avel); permute(quat);
    permute(radius); permute(mass); permute(inv_inertia); permute(ptype); permute(pid);
    permute(lifter_type);
  }

  /* a4: cudaMemset(0xFF) + Kernel_SpatialDecomp_BinData
   * (Device/DeviceInterface.cu:2097-2105, Collision_Detection_BroadPhase.cuh:279-325) */
  void bin_cells() {
    cell_start.assign(C, 0xFFFFFFFFu);
    cell_end.assign(C, 0xFFFFFFFFu);
    for (int i = 0; i < N; ++i) {
      uint32_t h = keys[i];
      if (h >= C) continue;
      if (i == 0 || keys[i - 1] != h) cell_start[h] = uint32_t(i);
      if (i == N - 1 || keys[i + 1] != h) cell_end[h] = uint32_t(i + 1);
    }
  }

  /* Stencil bounds, Collision_Detection_BroadPhase.cuh:428-459, clamped. */
  void stencil(const int g[3], int lo[3], int hi[3]) const {
    const int n[3] = {sim.num_NNCells.x, sim.num_NNCells.y, sim.num_NNCells.z};
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::max(g[a] - 1, 0);
      hi[a] = std::min(g[a] + 1, n[a] - 1);
    }
  }

  /* Overlap filter, CollisionDetection_Filter_SphereNNCell
   * (Collision_Detection_BroadPhase.cuh:369-383): length(pA-pB) - (rA+rB) < 0 */
  bool overlaps(int i, int j) const {
    return length(pos[i] - pos[j]) - (radius[i] + radius[j]) < R(0);
  }

  /* Visit the contacts of particle i in the reference order: y outer, z,
   * x inner, then ascending sorted index within a cell
   * (Collision_Detection_BroadPhase.cuh:466-489). */
  template <class F> int for_each_contact(int i, F&& fn) const {
    int g[3], lo[3], hi[3];
    grid_pos(pos[i], g);
    stencil(g, lo, hi);
    int count = 0;
    for (int y = lo[1]; y <= hi[1]; ++y)
      for (int z = lo[2]; z <= hi[2]; ++z)
        for (int x = lo[0]; x <= hi[0]; ++x) {
          uint32_t h = hash_of(x, y, z);
          if (h >= C) continue;
          uint32_t s = cell_start[h];
          if (s == 0xFFFFFFFFu) continue;
          uint32_t e = cell_end[h];
          for (uint32_t j = s; j < e; ++j) {
            if (int(j) == i) continue;
            if (overlaps(i, int(j))) { fn(int(j), count); ++count; }
          }
        }
    return count;
  }

  /* a5: Kernel_BroadCollisionDetection_NonSymmetry
   * (Collision_Detection_BroadPhase.cuh:402-513).  The list keeps the first
   * 32 entries per particle; nn holds the full count, as the reference's
   * NN_NUM does. */
  void contact_list(std::vector<uint32_t>& list, std::vector<uint32_t>& nn) {
    list.assign(size_t(N) * LIST_STRIDE, 0u);
    nn.assign(N, 0u);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i) {
      if (pid[i] < 0) continue;
      int c = for_each_contact(i, [&](int j, int k) {
        if (k < LIST_STRIDE) list[size_t(i) * LIST_STRIDE + k] = uint32_t(j);
      });
      nn[i] = uint32_t(c);
    }
  }

  /* ------------------------------------------------------------------ */
  /* a6: reference pair law                                              */
  /* Kernel_ParticleInteraction_Spheres_NonSymmetry DEM_Kernels_Compute.cuh:620-729 */
  /* Collision_Response_Particle_Particle Collision_Response.cuh:374-494 */
  /* ------------------------------------------------------------------ */
  void pair_force_ref(int a, int b, V3<R>& F, V3<R>& T) const {
    V3<R> relPos = pos[a] - pos[b];
    V3<R> N3 = relPos / length(relPos);
    R dis = std::fabs(length(relPos) - (radius[a] + radius[b])) + R(0.000010f);
    V3<R> RA = (R(-1.0f) * N3) * radius[a];
    V3<R> RB = N3 * radius[b];
    pair_response(a, b, N3, dis, RA, RB, F, T);
  }

  /* one contact's dissipation increment, kept per particle for this step's
   * ordered sum (the force-only debug pass keeps none) */
  void record(int k, int i, R e) const {
    if (sim.EnergyCalc && size_t(i) < incs[k].size()) incs[k][size_t(i)].push_back(e);
  }

  /* Collision_Response_Particle_Particle (Collision_Response.cuh:374-494):
   * normal N3 (towards a), depth dis, lever arms RA, RB from the centres. */
  void pair_response(int a, int b, V3<R> N3, R dis, V3<R> RA, V3<R> RB, V3<R>& F, V3<R>& T) const {
    const uint32_t ta = ptype[a], tb = ptype[b];
    V3<R> RVel = vel[a] - vel[b];
    const bdem_PP_Parms& pp = pobj[ta].PP[tb];
    const bool ov = pp_override && ta == 0 && tb == 0;
    V3<R> Vnormal = dot(RVel, N3) * N3;
    R Fnormal = (ov ? R(ov_kn) : R(pp.Kn)) * dis;
    V3<R> Fdiss = (ov ? R(ov_cn) : R(pp.Cn)) * Vnormal;
    V3<R> Ft = Fnormal * N3 - Fdiss;
    V3<R> V_CPA = vel[a] + cross(avel[a], RA);
    V3<R> V_CPB = vel[b] + cross(avel[b], RB);
    V3<R> V_CPNA = dot(V_CPA, N3) * N3;
    V3<R> V_CPNB = dot(V_CPB, N3) * N3;
    V3<R> Vt_A = V_CPA - V_CPNA;
    V3<R> Vt_B = V_CPB - V_CPNB;
    V3<R> Vt = Vt_A - Vt_B;
    R mag = dot(Vt, Vt);
    V3<R> Tang = mk<R>(0, 0, 0);
    if (mag > R(0)) {
      mag = rsqrt_(mag);
      R Velf = dot(RVel, RVel);
      Velf = Velf > R(0) ? rsqrt_(Velf) : R(1.0f);
      R tan_f = (((R(0.50f) * (mass[a] + mass[b])) / R(sim.InitalDelta_t)) * R(0.050f)) *
                (ov ? R(ov_fk) : R(pp.Fric_kinetic)) * (mag / Velf);
      Tang = (R(-1.0f) * tan_f) * (Vt / mag);
    }
    if (sim.EnergyCalc) {
      /* half the normal dissipation (each pair is visited from both sides),
       * the whole tangential one (:401-475) */
      R e = dot(Vnormal, Vnormal) > R(0) ? ((R(0.50f) * length(Fdiss)) * length(Vnormal)) * R(sim.InitalDelta_t)
                                         : R(0);
      if (mag > R(0)) e = e + ((length(Tang) * mag) * R(sim.InitalDelta_t));
      record(T_PP, a, e);
    }
    Ft = Ft + Tang;
    F = F + Ft;
    if (sim.Rotation) T = T + cross(RA, Ft);
  }

  /* ------------------------------------------------------------------ */
  /* History laws (extensions; textbook forms, no reference counterpart)  */
  /* ------------------------------------------------------------------ */
  /* Carry a tangential spring onto the current tangent plane, keeping its
   * length, then add vt*dt. */
  V3<R> rotate_spring(V3<R> xi_old, V3<R> n, V3<R> vt) const {
    V3<R> xi = xi_old - dot(xi_old, n) * n;
    R m_old = dot(xi_old, xi_old);
    R m_new = dot(xi, xi);
    if (m_new > R(0)) xi = xi * rsqrt_(m_old / m_new);
    return xi + vt * R(sim.InitalDelta_t);
  }

  /* Normal/tangential law shared by pairs and walls.
   *   n      unit normal pointing at particle a
   *   delta  overlap (>0)
   *   vc     contact-point velocity of a relative to the partner
   *   rstar, mstar  effective radius and mass
   *   lp     per type-pair constants; xi in/out tangential spring
   * Returns the force on a. */
  /* diss (may be null): the contact's dissipation rate, the library's own
   * definition for the history laws (the reference tallies only its own
   * law): normal damping force x |vn| while the normal force pushes, or the
   * elastic force x |vn| while damping has clipped it to zero (that spring
   * energy is not returned), plus the tangential damping force x |vt| while
   * the spring sticks, or the capped tangential force x |vt| while it slides. */
  V3<R> history_law(V3<R> n, R delta, V3<R> vc, R rstar, R mstar, const LawPair<R>& lp,
                    V3<R>& xi, R* diss = nullptr) const {
    R vn = dot(vc, n);
    V3<R> vt = vc - vn * n;
    xi = rotate_spring(xi, n, vt);
    R fn, kt, gt, dn, fe;
    if (law_kind() == BDEM_LAW_LSD_COULOMB) {
      R cn = R(-2.0f) * lp.beta * rsqrt_(lp.kn * mstar);
      fe = lp.kn * delta;
      fn = fe - cn * vn;
      dn = cn * vn;
      kt = lp.kt;
      gt = R(0);
    } else {
      R srd = rsqrt_(rstar * delta);
      R sn = R(2.0f) * lp.estar * srd;
      R st = R(8.0f) * lp.gstar * srd;
      R gn = R(-1.8257418583505538f) * lp.beta * rsqrt_(sn * mstar);   /* 2 sqrt(5/6) */
      gt = gn * lp.tg;                                                   /* sqrt(St m*) = tg sqrt(Sn m*) */
      fe = (R(4.0f) / R(3.0f)) * lp.estar * srd * delta;
      fn = fe - gn * vn;
      dn = gn * vn;
      kt = st;
    }
    const bool pushing = fn > R(0);
    if (fn < R(0)) fn = R(0);
    V3<R> ft = R(-1.0f) * kt * xi - gt * vt;
    R ft2 = dot(ft, ft);
    R fc = lp.mu * fn;
    bool slip = false;
    if (ft2 > fc * fc) {
      R s = fc / rsqrt_(ft2);
      ft = ft * s;
      xi = (R(-1.0f) / kt) * ft;
      slip = true;
    }
    if (diss) {
      R e = (pushing ? std::fabs(dn) : std::fabs(fe)) * std::fabs(vn);
      const R vt2 = dot(vt, vt);
      if (slip) e = e + rsqrt_(dot(ft, ft)) * rsqrt_(vt2);
      else e = e + gt * vt2;
      *diss = e;
    }
    return fn * n + ft;
  }

  /* Find the spring for partner id in the previous records of particle
   * `owner` (pair or wall table), 0 if absent. */
  V3<R> find_spring(const std::vector<PairHist<R>>& tab, const std::vector<int>& cnt,
                    int stride, int owner, int id) const {
    if (owner < 0 || owner >= int(cnt.size())) return mk<R>(0, 0, 0);
    int c = cnt[owner];
    for (int k = 0; k < c; ++k) {
      const PairHist<R>& h = tab[size_t(owner) * stride + k];
      if (h.id == id) return h.xi;
    }
    return mk<R>(0, 0, 0);
  }

  void pair_force_hist(int a, int b, V3<R>& F, V3<R>& T, std::vector<PairHist<R>>& out,
                       const std::vector<PairHist<R>>& hin, const std::vector<int>& hcin) const {
    V3<R> rel = pos[a] - pos[b];
    R d = length(rel);
    V3<R> n = rel * (R(1.0f) / d);
    R delta = (radius[a] + radius[b]) - d;
    V3<R> la = (R(-1.0f) * n) * radius[a];
    V3<R> lb = n * radius[b];
    V3<R> vc = (vel[a] + cross(avel[a], la)) - (vel[b] + cross(avel[b], lb));
    /* R* = ra rb / (ra + rb) and m* = ma mb / (ma + mb) from one division */
    R sr = radius[a] + radius[b], sm = mass[a] + mass[b];
    R q = R(1.0f) / (sr * sm);
    R rstar = ((radius[a] * radius[b]) * sm) * q;
    R mstar = ((mass[a] * mass[b]) * sr) * q;
    V3<R> xi = find_spring(hin, hcin, HMAX, pid[a], pid[b]);
    R e = R(0);
    V3<R> f = history_law(n, delta, vc, rstar, mstar, law[ptype[a]][ptype[b]], xi, &e);
    /* half the pair's dissipation from each side, as the reference law's
     * tally splits its normal work (Collision_Response.cuh:466-474) */
    record(T_PP, a, (R(0.5f) * e) * R(sim.InitalDelta_t));
    F = F + f;
    if (sim.Rotation) T = T + cross(la, f);
    if (int(out.size()) < HMAX) out.push_back(PairHist<R>{pid[b], xi});
  }

  /* ------------------------------------------------------------------ */
  /* a8: walls (WorldObject_InteractionParticle DEM_Kernels_Compute.cuh:254-610) */
  /* ------------------------------------------------------------------ */
  struct Coll { bool hit; R dis; V3<R> cp; V3<R> n; };

  /* is_PointOnSurface, Device/Utilities.cuh:20-53 (vertices of a world or
   * dynamic object) */
  bool point_on_surface(const bdem_Surface& s, const bdem_float3* vtx, V3<R> point) const {
    R area = R(0);
    for (int i = 0; i < s.num_vertex - 1; ++i) {
      V3<R> V1 = point - cvt<R>(vtx[s.vertex_Order[i]]);
      V3<R> V2 = point - cvt<R>(vtx[s.vertex_Order[i + 1]]);
      area = area + R(0.500000f) * length(cross(V1, V2));
    }
    V3<R> V1 = point - cvt<R>(vtx[s.vertex_Order[s.num_vertex - 1]]);
    V3<R> V2 = point - cvt<R>(vtx[s.vertex_Order[0]]);
    area = area + R(0.500000f) * length(cross(V1, V2));
    return (std::fabs(area - R(s.area)) / R(s.area)) * R(100.0000f) <= R(s.Tol);
  }
  bool point_on_surface(const bdem_Surface& s, const bdem_WOBJ& w, V3<R> point) const {
    return point_on_surface(s, w.vertex, point);
  }

  /* Collision_Detection_Sphere_Surface / _Plane (Collision_Detection_Spheres.cuh:6-100) */
  Coll sphere_plane(int i, V3<R> nrm, V3<R> cen, const bdem_Surface* s, const bdem_WOBJ* w) const {
    Coll c{false, R(0), mk<R>(0, 0, 0), mk<R>(0, 0, 0)};
    R r = radius[i];
    V3<R> Pt = pos[i] - r * nrm;
    R dis = dot(nrm, Pt - cen);
    if (dis < R(0.000000f) && dis > -r) {
      dis = std::fabs(dis);
      if (s) {
        V3<R> point = Pt + nrm * dis;
        if (!point_on_surface(*s, *w, point)) return c;
      }
      c.hit = true; c.dis = dis; c.cp = Pt; c.n = nrm;
    }
    return c;
  }

  /* Collision_Detection_Sphere_Cylinder (Collision_Detection_Spheres.cuh:234-265) */
  Coll sphere_cylinder(int i, const bdem_Macro_Cylinder& cyl) const {
    Coll c{false, R(0), mk<R>(0, 0, 0), mk<R>(0, 0, 0)};
    V3<R> Pvec = pos[i] * cvt<R>(cyl.AXIS) - cvt<R>(cyl.center_bot_cap);
    R Pd = length(Pvec) + radius[i];
    if (Pd > R(cyl.radius)) {
      V3<R> n = R(-1.0f) * Pvec;
      n = n * (R(1.0f) / length(n));
      c.n = n;
      c.cp = pos[i] - n * radius[i];
      c.dis = std::fabs(Pd - R(cyl.radius));
      c.hit = true;
    }
    return c;
  }

  /* Conical frustum wall (extension).  Closest point of the meridian
   * segment (s0,r0)-(s1,r1) to the particle centre; contact when closer than
   * the radius.  Normal points from the wall to the centre. */
  Coll sphere_cone(int i, const bdem_cone& cn) const {
    Coll c{false, R(0), mk<R>(0, 0, 0), mk<R>(0, 0, 0)};
    V3<R> ax = cvt<R>(cn.axis);
    V3<R> d = pos[i] - cvt<R>(cn.base);
    R s = dot(d, ax);
    V3<R> q = d - s * ax;
    R rho = length(q);
    if (!(rho > R(0))) return c;
    R ds = R(cn.s1) - R(cn.s0), dr = R(cn.r1) - R(cn.r0);
    R t = ((s - R(cn.s0)) * ds + (rho - R(cn.r0)) * dr) / (ds * ds + dr * dr);
    t = std::min(std::max(t, R(0)), R(1));
    R sc = R(cn.s0) + t * ds, rc = R(cn.r0) + t * dr;
    R es = s - sc, er = rho - rc;
    R dist = rsqrt_(es * es + er * er);
    if (dist < radius[i] && dist > R(0)) {
      V3<R> qh = q / rho;
      V3<R> n = (es / dist) * ax + (er / dist) * qh;
      c.hit = true;
      c.dis = radius[i] - dist;
      c.n = n;
      c.cp = pos[i] - n * radius[i];
    }
    return c;
  }

  /* The wall and lifter tallies (Collision_Response.cuh:25-89, 127-214,
   * 266-344): the damping work |Fdiss| |Vn| dt unless the force was clipped
   * as attractive, plus the tangential work |Ft| |Vt| dt. */
  void wall_tally(int k, int i, bool attractive, V3<R> Fdiss, V3<R> Vn, V3<R> Tang, R mag) const {
    if (!sim.EnergyCalc) return;
    R e = (!attractive && dot(Vn, Vn) > R(0)) ? (length(Fdiss) * length(Vn)) * R(sim.InitalDelta_t) : R(0);
    if (mag > R(0)) e = e + ((length(Tang) * mag) * R(sim.InitalDelta_t));
    record(k, i, e);
  }

  /* Collision_Response_Particle_StaticSurface (Collision_Response.cuh:4-109) */
  void resp_static(int i, const Coll& c, V3<R>& F, V3<R>& T) const {
    const bdem_POBJ& P = pobj[ptype[i]];
    V3<R> PVel = vel[i];
    V3<R> Rc = c.cp - pos[i];
    V3<R> Vnormal = dot(PVel, c.n) * c.n;
    R Fnormal = R(P.surface_Kn) * c.dis;
    V3<R> Fdiss = R(P.surface_Cn) * Vnormal;
    V3<R> Ftr = Fnormal * c.n - Fdiss;
    V3<R> Fdir = Ftr / length(Ftr);
    const bool attractive = dot(Fdir, c.n) < R(0);
    if (attractive) Ftr = Fnormal * c.n;
    V3<R> Vt = PVel - Vnormal;
    if (sim.Rotation) Vt = Vt + cross(avel[i], Rc);
    R mag = dot(Vt, Vt);
    V3<R> Tang = mk<R>(0, 0, 0);
    if (mag > R(0)) {
      mag = rsqrt_(mag);
      R Velf = dot(PVel, PVel);
      Velf = Velf > R(0) ? rsqrt_(Velf) : R(1.0f);
      R tan_f = (R(0.050f) * (mass[i] / R(sim.InitalDelta_t))) *
                (R(P.surface_Fric_static) * (mag / Velf));
      Tang = (R(-1.0f) * tan_f) * (Vt / mag);
    }
    wall_tally(T_SURF, i, attractive, Fdiss, Vnormal, Tang, mag);
    Ftr = Ftr + Tang;
    F = F + Ftr;
    if (sim.Rotation) T = T + cross(Rc, Ftr);
  }

  /* Collision_Response_Particle_MovingSurface (Collision_Response.cuh:116-233) */
  V3<R> resp_moving(int i, const Coll& c, V3<R> Vs, V3<R>& Tout) const {
    const bdem_POBJ& P = pobj[ptype[i]];
    V3<R> Rc = c.cp - pos[i];
    V3<R> RVel = vel[i] - Vs;
    V3<R> Vn = dot(RVel, c.n) * c.n;
    R Fnormal = R(P.surface_Kn) * c.dis;
    V3<R> Fdiss = R(P.surface_Cn) * Vn;
    V3<R> Ftr = Fnormal * c.n - Fdiss;
    V3<R> Fdir = Ftr / length(Ftr);
    const bool attractive = dot(Fdir, c.n) < R(0);
    if (attractive) Ftr = Fnormal * c.n;
    V3<R> V_CPA = vel[i] + cross(avel[i], Rc);
    V3<R> V_CPNA = dot(V_CPA, c.n) * c.n;
    V3<R> Vt_A = V_CPA - V_CPNA;
    V3<R> Vt_B = Vs - dot(Vs, c.n) * c.n;
    V3<R> Vt = Vt_A - Vt_B;
    R mag = dot(Vt, Vt);
    V3<R> Tang = mk<R>(0, 0, 0);
    if (mag > R(0)) {
      mag = rsqrt_(mag);
      R Velf = dot(RVel, RVel);
      Velf = Velf > R(0) ? rsqrt_(Velf) : R(1.0f);
      R tan_f = R(P.surface_Fric_kinetic) * (mag / Velf);
      Tang = ((R(-1.0f) * tan_f) * (Vt / mag)) * (mass[i] / R(sim.InitalDelta_t)) * R(0.050f);
    }
    wall_tally(T_SURF, i, attractive, Fdiss, Vn, Tang, mag);
    Ftr = Ftr + Tang;
    Tout = sim.Rotation ? cross(Rc, Ftr) : mk<R>(0, 0, 0);
    return Ftr;
  }

  /* Collision_Response_EndPlate_Particle (Collision_Response.cuh:504-526) */
  V3<R> resp_endplate(int i, const Coll& c) const {
    const bdem_POBJ& P = pobj[ptype[i]];
    V3<R> Vn = dot(vel[i], c.n) * c.n;
    V3<R> Fdiss = R(P.surface_Cn) * Vn;
    R Fn = R(P.surface_Kn) * c.dis;
    return Fn * c.n - Fdiss;
  }

  /* Drum surface velocity, DEM_Kernels_Compute.cuh:375-378 and 449-452 */
  V3<R> drum_velocity(V3<R> cp) const {
    R Rd = R(wobj[0].cyl_geo.radius);
    return cross(mk<R>(R(0.0), R(0.0), -R(sim.MillRotVel)), cp - mk<R>(Rd, Rd, R(0.0f)));
  }

  /* History-law wall contact: effective radius r, mass m, wall velocity vw. */
  void wall_hist(int i, const Coll& c, V3<R> vw, int wall_id, V3<R>& F, V3<R>& T,
                 std::vector<PairHist<R>>& wout) const {
    V3<R> la = c.cp - pos[i];
    V3<R> vc = (vel[i] + cross(avel[i], la)) - vw;
    V3<R> xi = find_spring(whist, wcount, WMAX, pid[i], wall_id);
    R e = R(0);
    V3<R> f = history_law(c.n, c.dis, vc, radius[i], mass[i], wlaw[ptype[i]], xi, &e);
    record(T_SURF, i, e * R(sim.InitalDelta_t));
    F = F + f;
    if (sim.Rotation) T = T + cross(la, f);
    if (int(wout.size()) < WMAX) wout.push_back(PairHist<R>{wall_id, xi});
  }

  /* Wall force and torque on particle i.  Reference law reproduces the
   * accumulate/overwrite pattern of DEM_Kernels_Compute.cuh:283-604: plane
   * hits add, cylinder and cap hits assign the force (torques add). */
  void wall_forces(int i, V3<R>& F, V3<R>& T, std::vector<PairHist<R>>* wout) const {
    F = mk<R>(0, 0, 0);
    T = mk<R>(0, 0, 0);
    const bool use_hist = law_kind() != BDEM_LAW_REFERENCE;
    const bool moving_planes = sim.Rotating_WObject && cylinder_rotation;
    for (int k = 0; k < num_wobj_active; ++k) {
      const bdem_WOBJ& w = wobj[k];
      if (w.surface_type == BDEM_SURFACE_PLANE) {
        for (int s = 0; s < w.num_surfaces; ++s) {
          const bdem_Surface& S = w.surfaces[s];
          Coll c = (S.area == -1.0f) ? sphere_plane(i, cvt<R>(S.normal), cvt<R>(S.centroid), nullptr, nullptr)
                                     : sphere_plane(i, cvt<R>(S.normal), cvt<R>(S.centroid), &S, &w);
          if (!c.hit) continue;
          if (use_hist) {
            V3<R> vw = moving_planes ? drum_velocity(c.cp) : mk<R>(0, 0, 0);
            wall_hist(i, c, vw, k * 16 + s, F, T, *wout);
          } else if (moving_planes) {
            V3<R> tq;
            F = F + resp_moving(i, c, drum_velocity(c.cp), tq);
            T = T + tq;
          } else {
            resp_static(i, c, F, T);
          }
        }
      } else if (w.surface_type == BDEM_SURFACE_CYLINDER) {
        const bdem_Macro_Cylinder& cyl = w.cyl_geo;
        Coll c = sphere_cylinder(i, cyl);
        if (c.hit) {
          if (use_hist) {
            V3<R> vw = cylinder_rotation ? drum_velocity(c.cp) : mk<R>(0, 0, 0);
            wall_hist(i, c, vw, 128 + k, F, T, *wout);
          } else if (cylinder_rotation) {
            V3<R> tq;
            F = resp_moving(i, c, drum_velocity(c.cp), tq);
            T = T + tq;
          } else {
            V3<R> f = mk<R>(0, 0, 0), tq = mk<R>(0, 0, 0);
            resp_static(i, c, f, tq);
            F = f;
            T = T + tq;
          }
        }
        if (cyl.has_top_cap) {
          Coll cc = sphere_plane(i, cvt<R>(cyl.normal_top_cap), cvt<R>(cyl.center_top_cap), nullptr, nullptr);
          if (cc.hit) {
            if (use_hist) wall_hist(i, cc, mk<R>(0, 0, 0), 144 + k, F, T, *wout);
            else F = resp_endplate(i, cc);
          }
        }
        if (cyl.has_bot_cap) {
          Coll cc = sphere_plane(i, cvt<R>(cyl.normal_bot_cap), cvt<R>(cyl.center_bot_cap), nullptr, nullptr);
          if (cc.hit) {
            if (use_hist) wall_hist(i, cc, mk<R>(0, 0, 0), 160 + k, F, T, *wout);
            else F = resp_endplate(i, cc);
          }
        }
      }
    }
    if (has_ext) {
      for (int k = 0; k < ext.num_cones; ++k) {
        Coll c = sphere_cone(i, ext.cone[k]);
        if (!c.hit) continue;
        if (use_hist) {
          wall_hist(i, c, mk<R>(0, 0, 0), 176 + k, F, T, *wout);
        } else {
          resp_static(i, c, F, T);
        }
      }
    }
  }

  /* ------------------------------------------------------------------ */
  /* a7: polyhedra (Collision_Detection_Polyhedra.cuh, Device/Utilities.cuh) */
  /* ------------------------------------------------------------------ */
  bool poly() const { return sim.particle_type == 1; }

  struct PolyRot { int nv, nf; V3<R> V[32]; V3<R> Fn[32]; V3<R> Fc[32]; };
  struct PColl { bool hit; R dis; V3<R> cp; V3<R> n; int cls; };
  enum { C_UNDEF = 0, C_VF = 1, C_EF = 2, C_FF = 3, C_EE = 4, C_SP = 5 };

  V3<R> qrot(Q4<R> Q, Q4<R> QT, V3<R> v) const {
    Q4<R> p{R(0.0f), v.x, v.y, v.z};
    Q4<R> r = qmul(qmul(Q, p), QT);
    return mk<R>(r.x, r.y, r.z);
  }

  /* RotatePolyhedra (Device/Utilities.cuh:420-479) */
  PolyRot rotate_poly(int i) const {
    PolyRot P;
    const bdem_poly_shape& po = shp(i);
    P.nv = po.num_vertex;
    P.nf = po.num_faces;
    const V3<R> c = pos[i];
    const Q4<R> Q = quat[i];
    const Q4<R> QT{Q.w, -Q.x, -Q.y, -Q.z};
    const int m = std::max(P.nf, P.nv);
    for (int k = 0; k < m; ++k) {
      if (sim.Rotation) {
        if (k < P.nv) P.V[k] = qrot(Q, QT, cvt<R>(po.Local_Vertex[k])) + c;
        if (k < P.nf) {
          P.Fn[k] = qrot(Q, QT, cvt<R>(po.face[k].normal));
          P.Fc[k] = qrot(Q, QT, cvt<R>(po.face[k].centroid) - cvt<R>(po.COM)) + c;
        }
      } else {
        if (k < P.nv) P.V[k] = cvt<R>(po.Local_Vertex[k]) + c;
        if (k < P.nf) {
          P.Fn[k] = cvt<R>(po.face[k].normal);
          P.Fc[k] = cvt<R>(po.face[k].centroid) - cvt<R>(po.COM) + c;
        }
      }
    }
    return P;
  }

  /* is_PointOnSurface over a polyhedron's rotated vertices (Device/Utilities.cuh:20-53) */
  bool point_on_poly_face(const bdem_Surface& f, const V3<R>* V, V3<R> point, R tol) const {
    R area = R(0.000000f);
    for (int q = 0; q < f.num_vertex - 1; ++q) {
      V3<R> V1 = point - V[f.vertex_Order[q]];
      V3<R> V2 = point - V[f.vertex_Order[q + 1]];
      area = area + R(0.500000f) * length(cross(V1, V2));
    }
    V3<R> V1 = point - V[f.vertex_Order[f.num_vertex - 1]];
    V3<R> V2 = point - V[f.vertex_Order[0]];
    area = area + R(0.500000f) * length(cross(V1, V2));
    return (std::fabs(area - R(f.area)) / R(f.area)) * R(100.0000f) <= tol;
  }

  /* is_SP_Poly_Poly (Device/Utilities.cuh:105-156): a face of A with B's
   * centre in front and every vertex of B in front separates them. */
  bool sp_poly_poly(const PolyRot& A, V3<R> PosB, const PolyRot& B) const {
    for (int f = 0; f < A.nf; ++f) {
      if (double(dot(A.Fn[f], PosB - A.Fc[f])) > 0.0) {
        bool contact = false;
        for (int v = 0; v < B.nv; ++v) {
          if (dot(A.Fn[f], B.V[v] - A.Fc[f]) <= R(0.00000f)) { contact = true; break; }
        }
        if (!contact) return true;
      }
    }
    return false;
  }

  /* isPointinPoly (Device/Utilities.cuh:63-98) */
  bool point_in_poly(const PolyRot& A, V3<R> point) const {
    for (int f = 0; f < A.nf; ++f)
      if (!(dot(A.Fn[f], point - A.Fc[f]) <= R(0.000100f))) return false;
    return true;
  }

  struct InPoly { bool inside; int face[3]; R dist[3]; };
  /* Collision_Detection_Vertex_InPoly_3Faces (Collision_Detection_Polyhedra.cuh:287-375) */
  InPoly vertex_in_poly(V3<R> PosP, V3<R> v, const PolyRot& B) const {
    InPoly r{false, {-1, -1, -1}, {R(10.0f), R(10.0f), R(10.0f)}};
    for (int f = 0; f < B.nf; ++f) {
      bool inside_plane = false;
      const R vd = dot(B.Fn[f], v - B.Fc[f]);
      if (vd < R(-0.0000100f)) {
        inside_plane = true;
        if (double(dot(B.Fn[f], PosP - B.Fc[f])) > 0.0) {
          const R a = std::fabs(vd);
          if (a < r.dist[0]) {
            if (r.face[0] > -1) {
              if (r.face[1] > -1) { r.dist[2] = r.dist[1]; r.face[2] = r.face[1]; }
              r.dist[1] = r.dist[0]; r.face[1] = r.face[0];
            }
            r.dist[0] = a; r.face[0] = f;
          } else if (a < r.dist[1]) {
            if (r.face[1] > -1) { r.dist[2] = r.dist[1]; r.face[2] = r.face[1]; }
            r.dist[1] = a; r.face[1] = f;
          } else if (a < r.dist[2]) {
            r.dist[2] = a; r.face[2] = f;
          }
        }
      }
      if (!inside_plane) return r;
    }
    r.inside = true;
    return r;
  }

  /* Face f of B; f = -1 (no face had A's centre in front: A's centre is
   * inside B) reads as a zero plane.  The reference indexes its face array
   * with -1 there, which is undefined; a zero normal makes the ray tests fail
   * and the depth 0, the same on every path of this build. */
  V3<R> fnorm(const PolyRot& B, int f) const { return f >= 0 ? B.Fn[f] : mk<R>(R(0), R(0), R(0)); }
  V3<R> fcent(const PolyRot& B, int f) const { return f >= 0 ? B.Fc[f] : mk<R>(R(0), R(0), R(0)); }

  /* The ray from PosA through `through` meets face f of B inside its polygon. */
  bool ray_hits_face(V3<R> PosA, V3<R> through, const PolyRot& B, const bdem_poly_shape& pb, int f, R tol) const {
    if (f < 0) return false;
    V3<R> Ray = through - PosA;
    Ray = Ray / length(Ray);
    const R dir = dot(Ray, B.Fn[f]);
    if (!(std::fabs(dir) > R(0))) return false;
    const R d = dot((B.Fc[f] - PosA), B.Fn[f]) / dir;
    return point_on_poly_face(pb.face[f], B.V, PosA + d * Ray, tol);
  }

  /* Collision_Detection_Polyhedara_Polyhedra_VertexFaces
   * (Collision_Detection_Polyhedra.cuh:1154-1630): vertices of A inside B;
   * one vertex -> vertex-face, two -> edge-face, more -> face-face, with a
   * ray test choosing among the (up to) three closest faces. */
  PColl vertex_faces(V3<R> PosA, const PolyRot& A, V3<R> PosB, const PolyRot& B, const bdem_poly_shape& pb) const {
    PColl res{false, R(0), mk<R>(0, 0, 0), mk<R>(0, 0, 0), C_UNDEF};
    int np = 0;
    int valid[8];
    InPoly sel[8];
    for (int v = 0; v < A.nv; ++v) {
      InPoly t = vertex_in_poly(PosA, A.V[v], B);
      if (t.inside) {
        if (np < 8) { sel[np] = t; valid[np] = v; }
        ++np;
      }
    }
    if (np == 0) return res;
    np = std::min(np, 8);
    res.hit = true;
    const InPoly& s0 = sel[0];
    if (np == 1) {
      res.cls = C_VF;
      res.dis = s0.dist[0];
      res.n = fnorm(B, s0.face[0]);
      res.cp = A.V[valid[0]];
      if (s0.face[1] == -1) return res;
      if (ray_hits_face(PosA, A.V[valid[0]], B, pb, s0.face[0], R(0.10f))) return res;
      if (s0.face[1] > -1 && ray_hits_face(PosA, A.V[valid[0]], B, pb, s0.face[1], R(0.10f))) {
        res.n = fnorm(B, s0.face[1]);
        res.dis = s0.dist[1];
        return res;
      }
      if (s0.face[2] > -1 && ray_hits_face(PosA, A.V[valid[0]], B, pb, s0.face[2], R(0.10f))) {
        res.n = fnorm(B, s0.face[2]);
        res.dis = s0.dist[2];
        return res;
      }
      return res;   /* the reference's "VF logic" fall-through keeps face 0 */
    }
    auto depth = [&](int f) { return std::fabs(dot(fnorm(B, f), (res.cp - fcent(B, f)))); };
    if (np == 2) {
      res.cls = C_EF;
      res.cp = (A.V[valid[0]] + A.V[valid[1]]) * R(0.50f);
      res.dis = depth(s0.face[0]);
      res.n = fnorm(B, s0.face[0]);
      if (s0.face[1] == -1 && sel[1].face[1] == -1) return res;
      for (int q = 0; q < 3; ++q) {
        if (q > 0 && !(s0.face[q] > -1)) continue;
        if (ray_hits_face(PosA, res.cp, B, pb, s0.face[q], R(0.10f))) {
          res.dis = depth(s0.face[q]);
          res.n = fnorm(B, s0.face[q]);
          return res;
        }
      }
      return res;
    }
    res.cls = C_FF;
    V3<R> cp = mk<R>(R(0.0f), R(0.0f), R(0.0f));
    for (int q = 0; q < np; ++q) cp = cp + A.V[valid[q]];
    res.cp = cp / R(float(np));
    const R tols[3] = {R(0.10f), R(0.010f), R(0.10f)};
    for (int q = 0; q < 3; ++q) {
      if (q > 0 && !(s0.face[q] > -1)) continue;
      if (ray_hits_face(PosA, res.cp, B, pb, s0.face[q], tols[q])) {
        res.dis = depth(s0.face[q]);
        res.n = fnorm(B, s0.face[q]);
        return res;
      }
    }
    return res;   /* "Face Logic" fall-through: normal and depth left unset */
  }

  /* Collision_Detection_Polyhedra_Polyhedra_Edges (Collision_Detection_Polyhedra.cuh:537-707) */
  PColl edges(const PolyRot& A, const bdem_poly_shape& pa, const PolyRot& B, const bdem_poly_shape& pb, V3<R> RelPos,
              V3<R> PosB) const {
    PColl res{false, R(0), mk<R>(R(0.0f), R(0.0f), R(0.0f)), mk<R>(0, 0, 0), C_SP};
    R min_edgedis = R(5.0f);
    V3<R> selA = mk<R>(0, 0, 0), selB = selA;
    bool found = false;
    for (int i = 0; i < pa.num_edges; ++i) {
      const V3<R> PA = A.V[pa.edge[i].point_vindex[0]];
      const V3<R> DA = A.V[pa.edge[i].point_vindex[1]] - PA;
      const R DEAD = dot(DA, DA);
      for (int j = 0; j < pb.num_edges; ++j) {
        const V3<R> PB = B.V[pb.edge[j].point_vindex[0]];
        const V3<R> DB = B.V[pb.edge[j].point_vindex[1]] - PB;
        const V3<R> dir = cross(DA, DB);
        const R mag = dot(dir, dir);
        if (!(mag > R(0.00000f))) continue;
        V3<R> DV = PA - PB;
        const R DEBD = dot(DB, DB);
        const R DEABD = dot(DA, DB);
        const R DEADAB = dot(DA, DV);
        const R DEBDAB = dot(DB, DV);
        const R detJinv = R(1.00000f) / (DEABD * DEABD - DEAD * DEBD);
        const R dA = detJinv * (DEBD * DEADAB - DEABD * DEBDAB);
        if (!(dA >= R(-0.010f) && dA <= R(1.0100f))) continue;
        const R dB = detJinv * (DEABD * DEADAB - DEAD * DEBDAB);
        if (!(dB >= R(-0.0100f) && dB <= R(1.010f))) continue;
        const V3<R> pA = PA + DA * dA;
        const V3<R> pB = PB + DB * dB;
        DV = (pA - pB);
        R dis = dot(DV, DV);
        if (dis > R(0.00000f)) {
          dis = rsqrt_(dis);
          if (dis < min_edgedis) {
            found = true;
            min_edgedis = dis;
            res.cp = pA;
            res.n = cross(DA, DB);
            selA = pA;
            selB = pB;
          }
        }
      }
    }
    if (!found) return res;
    bool valid = true;
    res.n = res.n * (R(1.0f) / length(res.n));
    if (length(PosB - selB) < length(PosB - selA)) {
      valid = false;
      if (point_in_poly(B, selA)) valid = true;
    }
    if (!valid) return res;
    res.hit = true;
    res.cls = C_EE;
    const R dn = dot(RelPos, res.n);
    if (std::fabs(dn) > R(0.0f)) {
      res.n = (dot(RelPos, res.n) / std::fabs(dot(RelPos, res.n))) * res.n;
      res.dis = std::fabs(dot(res.n, (selB - selA))) + R(0.000010f);
    } else {
      res.n = RelPos / length(RelPos);
      res.cp = PosB;
      res.dis = std::fabs(dot(res.n, (selB - selA))) + R(0.000010f);
    }
    return res;
  }

  /* Kernel_ParticleInteraction_Polyhedra_NonSymmetry (DEM_Kernels_Compute.cuh:736-992):
   * the contact of a with listed partner b (A's rotated shape given), after
   * the reference's depth clamps. */
  PColl poly_pair_contact(int a, int b, const PolyRot& A) const {
    const bdem_poly_shape& pa = shp(a);
    const bdem_poly_shape& pb = shp(b);
    const V3<R> PosA = pos[a], PosB = pos[b];
    const V3<R> RVel = vel[a] - vel[b];
    const PolyRot B = rotate_poly(b);
    const V3<R> RelPos = (PosA - PosB) / length(PosA - PosB);
    bool sp = sp_poly_poly(A, PosB, B);
    if (!sp) sp = sp_poly_poly(B, PosA, A);
    PColl res{false, R(0), mk<R>(0, 0, 0), mk<R>(0, 0, 0), C_SP};
    if (!sp) {
      res = vertex_faces(PosA, A, PosB, B, pb);
      if (!res.hit) {
        res = vertex_faces(PosB, B, PosA, A, pa);
        res.n = res.n * R(-1.0f);   /* B into A: normal reversed */
      }
      if (!res.hit) {
        res = edges(A, pa, B, pb, RelPos, PosB);
        if (res.hit) {
          res.cls = C_EE;
          res.dis = std::min(R(0.025f) * R(pa.radius), res.dis);
        }
      }
    }
    if (!res.hit) return res;
    const R rA = R(pa.radius);
    if (res.dis > R(0.1500f) * rA) {   /* contact error: fixed small overlap along the centres */
      res.n = RelPos / length(RelPos);
      res.dis = R(0.025f) * rA;
      res.cp = PosA;
      res.cls |= 8;
    }
    if (res.dis > R(0.0500f) * rA) {   /* large overlap: limited by speed (first branch wins, as written) */
      res.n = RelPos / length(RelPos);
      const double sp_rel = double(length(RVel));
      if (sp_rel > 50.0) res.dis = R(0.010f) * rA;
      else if (sp_rel > 150.0) res.dis = R(0.015f) * rA;
      else if (sp_rel > 250.0) res.dis = R(0.020f) * rA;
      else res.dis = R(0.0030f) * rA;
      res.cp = PosA;
      res.cls |= 16;
    }
    return res;
  }

  void pair_force_poly(int a, int b, const PolyRot& A, V3<R>& F, V3<R>& T) const {
    const PColl res = poly_pair_contact(a, b, A);
    if (!res.hit) return;
    pair_response(a, b, res.n, res.dis, res.cp - pos[a], res.cp - pos[b], F, T);
  }

  /* Collision_Detection_Vertcies_Surface (Collision_Detection_Polyhedra.cuh:85-185) */
  PColl vertices_surface(const PolyRot& A, const bdem_Surface& S, const bdem_float3* svtx) const {
    PColl res{false, R(0), mk<R>(R(0.0f), R(0.0f), R(0.0f)), mk<R>(0, 0, 0), C_UNDEF};
    R dis_max = R(0.0);
    int cnt = 0;
    const V3<R> n = cvt<R>(S.normal), c = cvt<R>(S.centroid);
    for (int v = 0; v < A.nv; ++v) {
      const R vd = dot(n, (A.V[v] - c));
      if (vd < R(0.000000f)) {
        const V3<R> hit = A.V[v] + std::fabs(vd) * n;
        if (point_on_surface(S, svtx, hit)) {
          ++cnt;
          res.cp = res.cp + hit;
          if (std::fabs(vd) > dis_max) { res.hit = true; dis_max = std::fabs(vd); }
          if (cnt > 2) {
            res.cls = C_FF; res.n = n; res.dis = dis_max;
            res.cp = mk<R>(R(0.0f), R(0.0f), R(0.0f));
            return res;
          }
        }
      }
    }
    if (!res.hit) { res.cls = C_SP; return res; }
    res.cls = cnt == 1 ? C_VF : C_EF;
    res.n = n;
    res.dis = dis_max;
    res.cp = res.cp / R(float(cnt));
    return res;
  }

  /* Collision_Detection_Polyhedra_Surface (Collision_Detection_Polyhedra.cuh:489-529) */
  PColl poly_surface(int i, const PolyRot& A, const bdem_Surface& S, const bdem_float3* svtx) const {
    PColl res{false, R(0), mk<R>(0, 0, 0), mk<R>(0, 0, 0), C_UNDEF};
    const V3<R> n = cvt<R>(S.normal), c = cvt<R>(S.centroid);
    if (dot(n, (pos[i] - c)) > R(0.000000f)) {
      if (dot(n, (pos[i] - c)) <= R(shp(i).radius)) {
        res = vertices_surface(A, S, svtx);
        if (res.cls == C_FF) res.cp = pos[i];   /* no moment for face-face */
        return res;
      }
    }
    return res;
  }

  /* Collision_Detection_Vertcies_Plane (Collision_Detection_Polyhedra.cuh:6-84) */
  PColl vertices_plane(const PolyRot& A, V3<R> n, V3<R> c) const {
    PColl res{false, R(0), mk<R>(R(0.0f), R(0.0f), R(0.0f)), mk<R>(0, 0, 0), C_SP};
    R dis_max = R(0.0);
    int cnt = 0;
    for (int v = 0; v < A.nv; ++v) {
      const R vd = dot(n, (A.V[v] - c));
      if (vd < R(0.000000f)) {
        ++cnt;
        res.cp = res.cp + A.V[v];
        if (std::fabs(vd) > dis_max) dis_max = std::fabs(vd);
      }
    }
    if (cnt == 0) return res;
    res.hit = true;
    res.cls = cnt == 2 ? C_EF : C_FF;   /* as written: one vertex also ends as face-face */
    res.cp = res.cp / R(float(cnt));
    res.n = n;
    res.dis = dis_max;
    return res;
  }

  /* Collision_Detection_Polyhedra_Cylinder (Collision_Detection_Polyhedra.cuh:2010-2132) */
  PColl poly_cylinder(int i, const PolyRot& A, const bdem_Macro_Cylinder& cyl) const {
    PColl res{false, R(0), mk<R>(0, 0, 0), mk<R>(0, 0, 0), C_UNDEF};
    const V3<R> AX = cvt<R>(cyl.AXIS), cb = cvt<R>(cyl.center_bot_cap);
    const V3<R> Pvec = pos[i] * AX - cb;
    const R Pvec_dis = length(Pvec) + R(shp(i).radius);
    if (!(Pvec_dis > R(cyl.radius))) return res;
    R dis = R(0.0);
    int cnt = 0;
    V3<R> s_point = mk<R>(R(0.0f), R(0.0f), R(0.0f));
    auto scale = [&](R x, int k) { return R(double(x) * (1.0 / double(float(k)))); };
    for (int v = 0; v < A.nv; ++v) {
      const V3<R> vp = (A.V[v]) * AX - cb;
      const R rc = length(vp);
      if (rc > R(cyl.radius)) {
        const R vd = rc - R(cyl.radius);
        ++cnt;
        s_point = s_point + A.V[v];
        dis = dis + vd;
        if (cnt > 2) {
          dis = scale(dis, cnt);
          res.hit = true;
          res.cls = C_FF;
          res.n = R(-1.0f) * Pvec;
          res.n = res.n * (R(1.0f) / length(res.n));
          res.dis = std::min(R(0.00100f), std::fabs(dis));
          res.cp = pos[i];
          return res;
        }
      }
    }
    if (cnt == 0) return res;
    if (cnt == 2) {
      dis = scale(dis, cnt);
      res.hit = true;
      res.cls = C_EF;
      res.cp = s_point / R(float(cnt));
      res.n = R(-1.0f) * Pvec;
      res.n = res.n * (R(1.0f) / length(res.n));
      res.dis = std::fabs(dis);
      return res;
    }
    dis = scale(dis, cnt);
    res.cp = s_point;
    res.n = R(-1.0f) * Pvec;
    res.n = res.n * (R(1.0f) / length(res.n));
    res.dis = dis;
    res.hit = true;
    res.cls = C_VF;
    return res;
  }

  /* WorldObject_InteractionParticle, polyhedra branches (DEM_Kernels_Compute.cuh:254-610).
   * Infinite planes have no polyhedra method in the reference: the previous
   * surface's result is reused there, as written. */
  void wall_forces_poly(int i, const PolyRot& A, V3<R>& F, V3<R>& T) const {
    F = mk<R>(0, 0, 0);
    T = mk<R>(0, 0, 0);
    const bool moving_planes = sim.Rotating_WObject && cylinder_rotation;
    PColl res{false, R(0), mk<R>(0, 0, 0), mk<R>(0, 0, 0), C_UNDEF};
    for (int k = 0; k < num_wobj_active; ++k) {
      const bdem_WOBJ& w = wobj[k];
      if (w.surface_type == BDEM_SURFACE_PLANE) {
        for (int s = 0; s < w.num_surfaces; ++s) {
          const bdem_Surface& S = w.surfaces[s];
          if (S.area != -1.0f) res = poly_surface(i, A, S, w.vertex);
          if (!res.hit) continue;
          Coll c{true, res.dis, res.cp, res.n};
          V3<R> ft = mk<R>(0, 0, 0), tq = ft;
          if (moving_planes) ft = resp_moving(i, c, drum_velocity(res.cp), tq);
          else resp_static(i, c, ft, tq);
          F = F + ft;
          if (sim.Rotation && res.cls != C_FF) T = T + tq;
        }
      } else if (w.surface_type == BDEM_SURFACE_CYLINDER) {
        const bdem_Macro_Cylinder& cyl = w.cyl_geo;
        res = poly_cylinder(i, A, cyl);
        if (res.hit) {
          Coll c{true, res.dis, res.cp, res.n};
          V3<R> ft = mk<R>(0, 0, 0), tq = ft;
          if (cylinder_rotation) ft = resp_moving(i, c, drum_velocity(res.cp), tq);
          else resp_static(i, c, ft, tq);
          F = ft;
          if (sim.Rotation && res.cls != C_FF) T = T + tq;
        }
        if (cyl.has_top_cap) {
          res = vertices_plane(A, cvt<R>(cyl.normal_top_cap), cvt<R>(cyl.center_top_cap));
          if (res.hit) F = resp_endplate(i, Coll{true, res.dis, res.cp, res.n});
        }
        if (cyl.has_bot_cap) {
          res = vertices_plane(A, cvt<R>(cyl.normal_bot_cap), cvt<R>(cyl.center_bot_cap));
          if (res.hit) F = resp_endplate(i, Coll{true, res.dis, res.cp, res.n});
        }
      }
    }
  }

  /* Get_Angular_Acc (Device/Utilities.cuh:486-608).  IC = RQ I0 (:511-524),
   * then IN[i][j] = sum_k IC[k][j] RQT[i][k] (:546-560), which is RQ^T RQ I0:
   * the rotation cancels up to rounding, so the body-frame tensor I0 acts on
   * the world-frame torque, as written. */
  V3<R> angular_acc(int i, V3<R> tau) const {
    const Q4<R> Q = quat[i];
    const float* I0 = shp(i).InertiaT;
    R RQ[9];
    RQ[0] = (Q.w * Q.w + Q.x * Q.x - Q.y * Q.y - Q.z * Q.z);
    RQ[1] = (R(2.0f) * Q.x * Q.y - R(2.0f) * Q.w * Q.z);
    RQ[2] = (R(2.0f) * Q.x * Q.z + R(2.0f) * Q.w * Q.y);
    RQ[3] = (R(2.0f) * Q.x * Q.y + R(2.0f) * Q.w * Q.z);
    RQ[4] = (Q.w * Q.w - Q.x * Q.x + Q.y * Q.y - Q.z * Q.z);
    RQ[5] = (R(2.0f) * Q.y * Q.z - R(2.0f) * Q.w * Q.x);
    RQ[6] = (R(2.0f) * Q.x * Q.z - R(2.0f) * Q.w * Q.y);
    RQ[7] = (R(2.0f) * Q.y * Q.z + R(2.0f) * Q.w * Q.x);
    RQ[8] = (Q.w * Q.w - Q.x * Q.x - Q.y * Q.y + Q.z * Q.z);
    R IC[9], IN[9];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        R sum = R(0);
        for (int k = 0; k < 3; ++k) sum += (RQ[r * 3 + k] * R(I0[k * 3 + c]));
        IC[r * 3 + c] = sum;
      }
    const R RQT[9] = {RQ[0], RQ[3], RQ[6], RQ[1], RQ[4], RQ[7], RQ[2], RQ[5], RQ[8]};
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        R sum = R(0);
        for (int k = 0; k < 3; ++k) sum += ((IC[k * 3 + c]) * RQT[r * 3 + k]);
        IN[r * 3 + c] = sum;
      }
    const R f[3] = {tau.x, tau.y, tau.z};
    R out[3];
    for (int r = 0; r < 3; ++r) {
      R sum = R(0);
      for (int k = 0; k < 3; ++k) sum += (IN[r * 3 + k] * (f[k]));
      out[r] = sum;
    }
    return mk<R>(out[0], out[1], out[2]);
  }

  /* ------------------------------------------------------------------ */
  /* a9: lifters (VolumeObject_InteractionParticle, DEM_Kernels_Compute.cuh:12-246) */
  /* ------------------------------------------------------------------ */
  struct DColl { bool hit; int type; R dis; V3<R> n; V3<R> cp; };

  /* Collision_Detection_Sphere_DObject (Collision_Detection_Spheres.cuh:108-225):
   * faces first unless the previous lifter contact was an edge, then the
   * edges as half-lines from their first vertex. */
  DColl sphere_dobject(int i, const bdem_DOBJ& D, int prev_type) const {
    DColl c{false, 0, R(0), mk<R>(0, 0, 0), mk<R>(0, 0, 0)};
    const R r = radius[i];
    const V3<R> p = pos[i];
    if (prev_type != 2) {
      for (int j = 0; j < D.num_faces; ++j) {
        const bdem_Surface& S = D.faces[j];
        const V3<R> nrm = cvt<R>(S.normal);
        const V3<R> Pt = p - r * nrm;
        R dis = dot(nrm, Pt - cvt<R>(S.centroid));
        if (dis < R(0.000000f) && dis > -r * R(0.50f)) {
          dis = std::fabs(dis);
          const V3<R> point = Pt + nrm * dis;
          if (point_on_surface(S, D.vertex, point)) {
            c.hit = true; c.type = 1; c.dis = dis; c.cp = Pt; c.n = nrm;
            return c;
          }
        }
      }
    }
    for (int e = 0; e < D.num_edges; ++e) {
      const V3<R> P0 = cvt<R>(D.vertex[D.edge[e].point_vindex[0]]);
      const V3<R> dir = cvt<R>(D.vertex[D.edge[e].point_vindex[1]]) - P0;
      const V3<R> Edir = dir / length(dir);
      const V3<R> vpc = p - P0;
      const R dVec = dot(Edir, vpc);
      if (dVec > R(0)) {
        const V3<R> P2 = P0 + dVec * Edir;
        const V3<R> d2 = p - P2;
        const R d = length(d2);
        if (d < r) {
          const R dis = std::fabs(d - r);
          if (dis < r) {
            V3<R> nn = p - P2;
            nn = nn * (R(1.0f) / length(nn));
            c.hit = true; c.type = 2; c.dis = dis; c.n = nn; c.cp = mk<R>(0, 0, 0);
            return c;
          }
        }
      }
    }
    return c;
  }

  /* Collision_Response_Particle_VolumeObject (Collision_Response.cuh:241-363),
   * lifter without spin (R_Body_COMVOBJ = VelVOBJ_Ang = 0). */
  void resp_dobject(int i, const DColl& c, V3<R> Vl, V3<R>& Ftr, V3<R>& Tq) const {
    const bdem_POBJ& P = pobj[ptype[i]];
    const V3<R> Rc = c.cp - pos[i];
    const V3<R> zero = mk<R>(0, 0, 0);
    const V3<R> RVel = vel[i] - Vl;
    const V3<R> Vnormal = dot(RVel, c.n) * c.n;
    const R Fnormal = R(P.Dsurface_Kn) * c.dis;
    const V3<R> Fdiss = R(P.Dsurface_Cn) * Vnormal;
    Ftr = Fnormal * c.n - Fdiss;
    const V3<R> Fdir = Ftr / length(Ftr);
    const bool attractive = dot(Fdir, c.n) < R(0);
    if (attractive) Ftr = Fnormal * c.n;
    const V3<R> V_CPA = vel[i] + cross(avel[i], Rc);
    const V3<R> V_CPB = Vl + cross(zero, zero);
    const V3<R> V_CPNA = dot(V_CPA, c.n) * c.n;
    const V3<R> V_CPNB = dot(V_CPB, c.n) * c.n;
    const V3<R> Vt_A = V_CPA - V_CPNA;
    const V3<R> Vt_B = V_CPB - V_CPNB;
    const V3<R> Vt_Rel = Vt_A - Vt_B;
    R mag = dot(Vt_Rel, Vt_Rel);
    V3<R> Tang = zero;
    if (mag > R(0)) {
      mag = rsqrt_(mag);
      R Velf = dot(RVel, RVel);
      Velf = Velf > R(0) ? rsqrt_(Velf) : R(1.0f);
      const R tan_f = (R(0.050f) * (mass[i] / R(sim.InitalDelta_t))) *
                      (R(P.Dsurface_Fric_kinetic) * (mag / Velf));
      Tang = (R(-1.0f) * tan_f) * (Vt_Rel / mag);
    }
    wall_tally(T_LIFT, i, attractive, Fdiss, Vnormal, Tang, mag);
    Ftr = Ftr + Tang;
    Tq = sim.Rotation ? cross(Rc, Ftr) : zero;
  }

  /* ---- polyhedra against lifters (VolumeObject_InteractionParticle,
   * DEM_Kernels_Compute.cuh:100-224) ---- */

  /* is_SP_VolumeObject_Poly (Device/Utilities.cuh:272-309): a lifter face
   * with the particle's centre in front and every particle vertex in front. */
  bool sp_dobj_poly(const bdem_DOBJ& D, V3<R> PosP, const PolyRot& A) const {
    for (int f = 0; f < D.num_faces; ++f) {
      const V3<R> n = cvt<R>(D.faces[f].normal), c = cvt<R>(D.faces[f].centroid);
      if (double(dot(n, PosP - c)) > 0.0) {
        bool contact = false;
        for (int v = 0; v < A.nv; ++v)
          if (dot(n, A.V[v] - c) <= R(0.00000f)) { contact = true; break; }
        if (!contact) return true;
      }
    }
    return false;
  }

  /* is_SP_Poly_VolumeObject (Device/Utilities.cuh:212-247): a particle face
   * with the lifter's COM in front and every lifter vertex in front. */
  bool sp_poly_dobj(const PolyRot& A, V3<R> PosD, const bdem_DOBJ& D) const {
    for (int f = 0; f < A.nf; ++f) {
      if (double(dot(A.Fn[f], PosD - A.Fc[f])) > 0.0) {
        bool contact = false;
        for (int v = 0; v < D.num_vertex; ++v)
          if (dot(A.Fn[f], cvt<R>(D.vertex[v]) - A.Fc[f]) <= R(0.00000f)) { contact = true; break; }
        if (!contact) return true;
      }
    }
    return false;
  }

  /* Collision_Detection_Vertex_InDObject_3Faces (Collision_Detection_Polyhedra.cuh:383-476):
   * behind (or within 1e-5 of) every lifter face; the three closest faces
   * that have the particle's centre in front. */
  InPoly vertex_in_dobj(V3<R> PosP, V3<R> v, const bdem_DOBJ& D) const {
    InPoly r{false, {-1, -1, -1}, {R(10.0f), R(10.0f), R(10.0f)}};
    for (int f = 0; f < D.num_faces; ++f) {
      const V3<R> n = cvt<R>(D.faces[f].normal), c = cvt<R>(D.faces[f].centroid);
      const R vd = dot(n, v - c);
      if (!(vd <= R(0.000010f))) return r;
      if (dot(n, PosP - c) > R(0.0f)) {
        const R a = std::fabs(vd);
        if (a < r.dist[0]) {
          if (r.face[0] > -1) {
            if (r.face[1] > -1) { r.dist[2] = r.dist[1]; r.face[2] = r.face[1]; }
            r.dist[1] = r.dist[0]; r.face[1] = r.face[0];
          }
          r.dist[0] = a; r.face[0] = f;
        } else if (a < r.dist[1]) {
          if (r.face[1] > -1) { r.dist[2] = r.dist[1]; r.face[2] = r.face[1]; }
          r.dist[1] = a; r.face[1] = f;
        } else if (a < r.dist[2]) {
          r.dist[2] = a; r.face[2] = f;
        }
      }
    }
    r.inside = true;
    return r;
  }

  /* Lifter face f; f = -1 (no face had the centre in front) reads as a zero
   * plane, as for polyhedra. */
  V3<R> dnorm(const bdem_DOBJ& D, int f) const { return f >= 0 ? cvt<R>(D.faces[f].normal) : mk<R>(0, 0, 0); }
  V3<R> dcent(const bdem_DOBJ& D, int f) const { return f >= 0 ? cvt<R>(D.faces[f].centroid) : mk<R>(0, 0, 0); }

  bool ray_hits_dface(V3<R> PosP, V3<R> through, const bdem_DOBJ& D, const V3<R>* DV, int f, R tol) const {
    if (f < 0) return false;
    V3<R> Ray = through - PosP;
    Ray = Ray / length(Ray);
    const R dir = dot(Ray, dnorm(D, f));
    if (!(std::fabs(dir) > R(0))) return false;
    const R d = dot((dcent(D, f) - PosP), dnorm(D, f)) / dir;
    return point_on_poly_face(D.faces[f], DV, PosP + d * Ray, tol);
  }

  /* Collision_Detection_Polyhedara_VolumeObject_VertexFace
   * (Collision_Detection_Polyhedra.cuh:1638-1990): particle vertices inside
   * the lifter; one -> vertex-face, two -> edge-face (ray tests against the
   * three closest faces), more -> the face every inside vertex has among its
   * three closest.  The first eight inside vertices count. */
  PColl vertex_faces_dobj(V3<R> PosP, const PolyRot& A, const bdem_DOBJ& D) const {
    PColl res{false, R(0), mk<R>(0, 0, 0), mk<R>(0, 0, 0), C_UNDEF};
    V3<R> DV[8];
    for (int k = 0; k < D.num_vertex && k < 8; ++k) DV[k] = cvt<R>(D.vertex[k]);
    int np = 0, valid[8];
    InPoly sel[8];
    for (int v = 0; v < A.nv; ++v) {
      InPoly t = vertex_in_dobj(PosP, A.V[v], D);
      if (t.inside) {
        if (np < 8) { sel[np] = t; valid[np] = v; }
        ++np;
      }
    }
    if (np == 0) return res;
    np = std::min(np, 8);
    res.hit = true;
    const InPoly& s0 = sel[0];
    if (np == 1) {
      res.cls = C_VF;
      res.dis = s0.dist[0];
      res.n = dnorm(D, s0.face[0]);
      res.cp = A.V[valid[0]];
      if (s0.face[1] == -1) return res;
      if (ray_hits_dface(PosP, A.V[valid[0]], D, DV, s0.face[0], R(0.1f))) return res;
      for (int q = 1; q < 3; ++q) {
        if (s0.face[q] > -1 && ray_hits_dface(PosP, A.V[valid[0]], D, DV, s0.face[q], R(0.10f))) {
          res.n = dnorm(D, s0.face[q]);
          res.dis = s0.dist[q];
          return res;
        }
      }
      return res;
    }
    auto depth = [&](int f) { return std::fabs(dot(dnorm(D, f), (res.cp - dcent(D, f)))); };
    if (np == 2) {
      res.cls = C_EF;
      res.cp = (A.V[valid[0]] + A.V[valid[1]]) * R(0.50f);
      res.n = dnorm(D, s0.face[0]);
      res.dis = depth(s0.face[0]);
      if (s0.face[1] == -1 && sel[1].face[1] == -1) return res;
      for (int q = 0; q < 3; ++q) {
        if (q > 0 && !(s0.face[q] > -1)) continue;
        if (ray_hits_dface(PosP, res.cp, D, DV, s0.face[q], R(0.10f))) {
          res.dis = depth(s0.face[q]);
          res.n = dnorm(D, s0.face[q]);
          return res;
        }
      }
      return res;
    }
    res.cls = C_FF;
    V3<R> cp = mk<R>(R(0.0f), R(0.0f), R(0.0f));
    for (int q = 0; q < np; ++q) cp = cp + A.V[valid[q]];
    res.cp = cp / R(float(np));
    /* count, per lifter face, the inside vertices that have it among their
     * three closest, with the smallest distance */
    int common[3] = {0, 0, 0};
    R cdis[3] = {R(10.0), R(10.0), R(10.0)};
    for (int q = 0; q < np; ++q)
      for (int j = 0; j < 3; ++j) {
        const int f = sel[q].face[j];
        if (f > -1 && f < 3) {
          common[f]++;
          if (sel[q].dist[j] < cdis[f]) cdis[f] = sel[q].dist[j];
        }
      }
    for (int f = 0; f < D.num_faces && f < 3; ++f)
      if (common[f] == np) {
        res.dis = cdis[f];
        res.n = dnorm(D, f);
        return res;
      }
    return res;   /* no common face: normal and depth left unset (zero here) */
  }

  /* Collision_Detection_DObject_Polyhedra_Edges (Collision_Detection_Polyhedra.cuh:716-890):
   * closest crossing of a lifter edge (A) and a particle edge (B) where the
   * lifter's point is nearer the particle's centre.  The reference leaves
   * the normal unset when no such crossing is within 0.1 cm of the other
   * point; that case is no contact here. */
  PColl edges_dobj(const bdem_DOBJ& D, const PolyRot& B, const bdem_poly_shape& pb, V3<R> PosB, V3<R> RelPos) const {
    PColl res{false, R(0), mk<R>(R(0.0f), R(0.0f), R(0.0f)), mk<R>(0, 0, 0), C_SP};
    R min_edgedis = R(0.10f);
    V3<R> selA = mk<R>(0, 0, 0), selB = selA, nsel = selA, cpsum = selA;
    int count = 0;
    bool found = false, have_min = false;
    for (int i = 0; i < D.num_edges; ++i) {
      const V3<R> PA = cvt<R>(D.vertex[D.edge[i].point_vindex[0]]);
      const V3<R> DA = cvt<R>(D.vertex[D.edge[i].point_vindex[1]]) - PA;
      const R DEAD = dot(DA, DA);
      for (int j = 0; j < pb.num_edges; ++j) {
        const V3<R> PB = B.V[pb.edge[j].point_vindex[0]];
        const V3<R> DB = B.V[pb.edge[j].point_vindex[1]] - PB;
        const V3<R> dir = cross(DA / length(DA), DB / length(DB));
        const R mag = dot(dir, dir);
        if (!(mag > R(0.00000f))) continue;
        V3<R> DV = PA - PB;
        const R DEBD = dot(DB, DB);
        const R DEABD = dot(DA, DB);
        const R DEADAB = dot(DA, DV);
        const R DEBDAB = dot(DB, DV);
        const R detJinv = R(1.00000f) / (DEABD * DEABD - DEAD * DEBD);
        const R dA = detJinv * (DEBD * DEADAB - DEABD * DEBDAB);
        if (!(dA >= R(-0.010f) && dA <= R(1.0100f))) continue;
        const R dB = detJinv * (DEABD * DEADAB - DEAD * DEBDAB);
        if (!(dB >= R(-0.0100f) && dB <= R(1.010f))) continue;
        const V3<R> pA = PA + DA * dA;
        const V3<R> pB = PB + DB * dB;
        DV = (pA - pB);
        R dis = dot(DV, DV);
        if (!(dis > R(0.00000f))) continue;
        dis = rsqrt_(dis);
        if (length(PosB - pB) > length(PosB - pA)) {
          ++count;
          cpsum = cpsum + pA;
          found = true;
          if (dis <= min_edgedis) {
            min_edgedis = dis;
            nsel = cross(DA, DB);
            selA = pA;
            selB = pB;
            have_min = true;
          }
        }
      }
    }
    if (!found || !have_min) return res;
    V3<R> n = nsel * (R(1.0f) / length(nsel));
    res.hit = true;
    res.cls = C_EE;
    const R dn = dot(RelPos, n);
    if (std::fabs(dn) > R(0.0f)) {
      res.n = R(-1.0f) * ((dot(RelPos, n) / std::fabs(dot(RelPos, n))) * n);
      res.dis = std::fabs(dot(res.n, (selB - selA)));
      res.cp = cpsum / R(float(count));
    } else {
      res.n = R(-1.0f) * RelPos / length(RelPos);
      res.cp = PosB;
      res.dis = std::fabs(dot(res.n, (selB - selA))) + R(0.000010f);
    }
    return res;
  }

  /* Lifter force and torque on polyhedron i (the particle_type 1 branch of
   * VolumeObject_InteractionParticle): first lifter in contact acts. */
  void lifter_forces_poly(int i, const PolyRot& A, V3<R>& F, V3<R>& T) const {
    F = mk<R>(0, 0, 0);
    T = mk<R>(0, 0, 0);
    const bdem_poly_shape& pa = shp(i);
    const R r = R(pa.radius);
    const V3<R> P = pos[i];
    const bool check = dobj[0].is_translating || !sim.isMillSim || sim.Rotating_WObject ||
                       length(mk<R>(P.x, P.y, R(0.0)) - cvt<R>(sim.cent)) >
                           (R(sim.DObject_CylinderBoundR) - R(2.0f) * r);
    if (!check) return;
    for (size_t d = 0; d < dobj.size(); ++d) {
      const bdem_DOBJ& D = dobj[d];
      const V3<R> BV = (P - cvt<R>(D.COM)) * cvt<R>(D.Axis);
      if (!(length(BV) < (R(D.boundR) + r))) continue;
      PColl res{false, R(0), mk<R>(0, 0, 0), mk<R>(0, 0, 0), C_SP};
      bool sp = sp_dobj_poly(D, P, A);
      if (!sp) sp = sp_poly_dobj(A, cvt<R>(D.COM), D);
      if (!sp) {
        V3<R> PosD = cvt<R>(D.COM);
        PosD.z = P.z;
        res = vertex_faces_dobj(P, A, D);
        if (!res.hit) {
          const V3<R> RelPos = (PosD - P) / length(PosD - P);
          res = edges_dobj(D, A, pa, P, RelPos);
        }
      }
      if (!res.hit) continue;
      V3<R> Vl = mk<R>(0, 0, 0);
      if (D.is_attached_DWorld) Vl = drum_velocity(res.cp);
      else if (D.is_translating) Vl = cvt<R>(D.velocity);
      const DColl c{true, 0, res.dis, res.n, res.cp};
      V3<R> Ftr, Tq;
      resp_dobject(i, c, Vl, Ftr, Tq);
      F = F + Ftr;
      if (sim.Rotation) T = T + Tq;
      break;
    }
  }

  /* Lifter force and torque on particle i; returns the contact type kept
   * for the next step (0 none, 1 face, 2 edge).  Only the first lifter in
   * contact acts (`break`, DEM_Kernels_Compute.cuh:224). */
  int lifter_forces(int i, V3<R>& F, V3<R>& T) const {
    F = mk<R>(0, 0, 0);
    T = mk<R>(0, 0, 0);
    if (dobj.empty()) return 0;
    const R r = radius[i];
    const V3<R> p = pos[i];
    DColl res{false, 0, R(0), mk<R>(0, 0, 0), mk<R>(0, 0, 0)};
    const bool check = dobj[0].is_translating || !sim.isMillSim || sim.Rotating_WObject ||
                       length(mk<R>(p.x, p.y, R(0.0)) - cvt<R>(sim.cent)) > (R(sim.DObject_CylinderBoundR) - R(2.0f) * r);
    if (check) {
      for (size_t d = 0; d < dobj.size(); ++d) {
        const bdem_DOBJ& D = dobj[d];
        const V3<R> BV = (p - cvt<R>(D.COM)) * cvt<R>(D.Axis);
        if (length(BV) < (R(D.boundR) + r)) {
          res = sphere_dobject(i, D, lifter_type[i]);
          if (res.hit) {
            if (res.type == 2) res.cp = p;   /* no spin for an edge contact */
            V3<R> Vl = mk<R>(0, 0, 0);
            if (D.is_attached_DWorld) Vl = drum_velocity(res.cp);
            else if (D.is_translating) Vl = cvt<R>(D.velocity);
            V3<R> Ftr, Tq;
            resp_dobject(i, res, Vl, Ftr, Tq);
            F = F + Ftr;
            if (sim.Rotation) T = T + Tq;
            break;
          }
        }
      }
    }
    return res.type;
  }

  /* Pair forces (including gravity, DEM_Kernels_Compute.cuh:654) and wall
   * forces of every live particle on the sorted state. */
  void compute_forces() {
    f_pp.assign(N, mk<R>(0, 0, 0)); t_pp = f_pp; f_w = f_pp; t_w = f_pp; f_l = f_pp; t_l = f_pp;
    const bool use_hist = law_kind() != BDEM_LAW_REFERENCE;
    std::vector<int> new_ltype(lifter_type);
    std::vector<int> hc_new, wc_new;
    std::vector<PairHist<R>> h_new, w_new;
    if (use_hist) {
      hc_new.assign(hcount.size(), 0);
      wc_new.assign(wcount.size(), 0);
      h_new.assign(hist.size(), PairHist<R>{-1, mk<R>(0, 0, 0)});
      w_new.assign(whist.size(), PairHist<R>{-1, mk<R>(0, 0, 0)});
    }
    int hov = 0;
    if (poly()) {
      int lov = 0, lhit = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(| : lov) reduction(+ : lhit)
      for (int i = 0; i < N; ++i) {
        if (pid[i] < 0) continue;
        V3<R> F = cvt<R>(sim.Force_Field) * mass[i];
        V3<R> T = mk<R>(0, 0, 0);
        const PolyRot A = rotate_poly(i);
        int n = for_each_contact(i, [&](int j, int k) {
          if (k < LIST_STRIDE) pair_force_poly(i, j, A, F, T);   /* the 32-slot list */
        });
        if (n > LIST_STRIDE) lov = 1;
        f_pp[i] = F;
        t_pp[i] = sim.Rotation ? T : mk<R>(0, 0, 0);
        V3<R> Fw, Tw;
        wall_forces_poly(i, A, Fw, Tw);
        f_w[i] = Fw;
        t_w[i] = Tw;
        if (!dobj.empty()) {
          V3<R> Fl, Tl;
          lifter_forces_poly(i, A, Fl, Tl);
          f_l[i] = Fl;
          t_l[i] = Tl;
          if (dot(Fl, Fl) > R(0)) lhit = lhit + 1;
        }
      }
      if (lov) warn_list_overflow = 1;
      lifter_contacts = lhit;
      return;
    }
#pragma omp parallel for schedule(static) reduction(| : hov)
    for (int i = 0; i < N; ++i) {
      if (pid[i] < 0) continue;
      V3<R> F = cvt<R>(sim.Force_Field) * mass[i];
      V3<R> T = mk<R>(0, 0, 0);
      std::vector<PairHist<R>> out, wout;
      int n = for_each_contact(i, [&](int j, int) {
        if (use_hist) pair_force_hist(i, j, F, T, out, hist, hcount);
        else pair_force_ref(i, j, F, T);
      });
      if (use_hist && n > HMAX) hov = 1;
      if (!sim.Rotation) T = mk<R>(0, 0, 0);
      f_pp[i] = F;
      t_pp[i] = T;
      V3<R> Fw, Tw;
      wall_forces(i, Fw, Tw, &wout);
      f_w[i] = Fw;
      t_w[i] = Tw;
      if (!dobj.empty()) {
        V3<R> Fl, Tl;
        new_ltype[i] = lifter_forces(i, Fl, Tl);
        f_l[i] = Fl;
        t_l[i] = Tl;
      }
      if (use_hist) {
        int p = pid[i];
        hc_new[p] = int(out.size());
        for (size_t k = 0; k < out.size(); ++k) h_new[size_t(p) * HMAX + k] = out[k];
        wc_new[p] = int(wout.size());
        for (size_t k = 0; k < wout.size(); ++k) w_new[size_t(p) * WMAX + k] = wout[k];
      }
    }
    if (hov) warn_hist_overflow = 1;
    lifter_type.swap(new_ltype);
    if (use_hist) {
      hcount.swap(hc_new); hist.swap(h_new);
      wcount.swap(wc_new); whist.swap(w_new);
    }
  }

  /* a10: Integrate_Euler_NonSymmetry (Integration.cuh:6-184), spheres. */
  void integrate() {
    const R dt = R(sim.InitalDelta_t);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i) {
      if (pid[i] < 0) continue;
      V3<R> acc = f_w[i] + f_pp[i];
      if (!dobj.empty()) acc = acc + f_l[i];   /* Integration.cuh:41-44 */
      acc = acc * (R(1.0f) / mass[i]);
      vel[i] = vel[i] + acc * dt;
      pos[i] = pos[i] + vel[i] * dt;
      if (sim.Mode2D) vel[i].z = R(0);
      V3<R> fl = t_pp[i] + t_w[i];
      if (!dobj.empty()) fl = fl + t_l[i];     /* Integration.cuh:92-95 */
      if (sim.Rotation) {
        if (dot(fl, fl) > R(0)) {
          V3<R> aa = poly() ? angular_acc(i, fl) : inv_inertia[i] * fl;
          avel[i] = avel[i] + aa * dt;
          avel[i] = avel[i] - avel[i] * (pp_override ? R(ov_rr) : R(sim.Roll_Res));
        }
        if (poly() || sim.sphere_orient) {
          V3<R> om = avel[i];
          R mo = dot(om, om);
          Q4<R> dQ{1, 0, 0, 0};
          if (mo > R(0.000000f)) {
            mo = rsqrt_(mo);
            V3<R> dir = om / mo;
            dQ.w = std::cos(R(0.5000f) * mo * dt);
            R sv = std::sin(R(0.5000f) * mo * dt);
            dQ.x = sv * dir.x; dQ.y = sv * dir.y; dQ.z = sv * dir.z;
          }
          Q4<R> Q = qnormalise(qmul(dQ, quat[i]));
          quat[i] = (Q.w <= R(-1.00000f)) ? dQ : Q;
          if (sim.Mode2D) { quat[i].x = R(0); quat[i].y = R(0); }
        }
        if (sim.Mode2D) { avel[i].x = R(0); avel[i].y = R(0); }
      }
    }
  }

  /* a11: Remove_Silo_Discharged_Particles (GPU_Memory_Transfer.cuh:650-664),
   * every kill_check_interval steps (Device/DeviceInterface.cu:2466-2499). */
  int kill_discharged() {
    int killed = 0;
    for (int i = 0; i < N; ++i) {
      if (pid[i] == -1) continue;
      if (pos[i].y < R(sim.Silo_Kill_Height)) {
        pos[i] = cvt<R>(sim.max_size);
        pid[i] = -1;
        ++killed;
      }
    }
    return killed;
  }

  /* Hash, sort, reorder and bin: the broad-phase front half of a step. */
  void prepare() {
    compute_hash();
    sort_by_key();
    reorder();
    bin_cells();
  }

  /* Rotate_VolObject_Inc (Input/Processing_Mills.h:7-60): every lifter
   * attached to the drum turned by `rad` about `about` in the xy plane --
   * cos/sin of the double angle, products and sum in double, stored in
   * float -- and its faces re-derived in float (quadrilateral form). */
  void turn_lifters(double rad, bdem_float3 about) {
    const double c = std::cos(rad), s = std::sin(rad);
    auto turn = [&](bdem_float3 v) {   /* x, y replaced; z kept; then += about */
      const float px = v.x - about.x, py = v.y - about.y;
      bdem_float3 o;
      o.x = float(c * double(px) + s * double(py)) + about.x;
      o.y = float(-s * double(px) + c * double(py)) + about.y;
      o.z = v.z + about.z;
      return o;
    };
    for (bdem_DOBJ& d : dobj) {
      if (!d.is_attached_DWorld) continue;
      for (int k = 0; k < d.num_vertex; ++k) d.vertex[k] = turn(d.vertex[k]);
      for (int j = 0; j < d.num_faces; ++j) {
        bdem_Surface& f = d.faces[j];
        V3<float> sum = mk<float>(0.f, 0.f, 0.f);
        for (int q = 0; q < f.num_vertex; ++q) sum = sum + cvt<float>(d.vertex[f.vertex_Order[q]]);
        const V3<float> cen = (1.0f / float(f.num_vertex)) * sum;
        const V3<float> v0 = cvt<float>(d.vertex[f.vertex_Order[0]]), v1 = cvt<float>(d.vertex[f.vertex_Order[1]]);
        const V3<float> v2 = cvt<float>(d.vertex[f.vertex_Order[2]]), v3 = cvt<float>(d.vertex[f.vertex_Order[3]]);
        const V3<float> V1 = v1 - v0, V2 = v2 - v1, V3_ = v3 - v2, V4 = v0 - v3;
        const V3<float> n = cross(V2, V1);
        const V3<float> nn = n / length(n);
        f.centroid = bdem_float3{cen.x, cen.y, cen.z};
        f.normal = bdem_float3{nn.x, nn.y, nn.z};
        f.area = 0.500000f * length(n) + 0.500000f * length(cross(V3_, V4));
      }
      d.COM = turn(d.COM);
    }
  }

  /* One Device_DEM_UpdateSim. */
  void step() {
    prepare();
    if (sim.EnergyCalc)
      for (int k = 0; k < 3; ++k) incs[k].assign(size_t(N), std::vector<R>());
    compute_forces();
    if (sim.EnergyCalc)
      for (int i = 0; i < N; ++i)
        for (int k = 0; k < 3; ++k)
          for (R e : incs[k][size_t(i)]) {
            tally[k] = float(tally[k] + float(e));
            tally_exact[k] += double(e);
          }
    integrate();
    if (sim.Kill_Particles) {
      ++stepc;
      int interval = (has_ext && ext.kill_check_interval > 0) ? ext.kill_check_interval : 1000;
      if (stepc >= interval) {
        kill_discharged();
        stepc = 0;
      }
    }
    ++steps_done;
  }
};

}  // namespace bdem_oracle
