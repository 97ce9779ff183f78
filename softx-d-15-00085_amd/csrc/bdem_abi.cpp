/*
 * bdem_abi.cpp - context, buffer management and the C ABI (include/bdem_abi.h).
 *
 * One context = one simulation on one HIP stream.  All device buffers are
 * allocated once in bdem_set_simulation_data; a step launches
 *   [hash] -> radix sort (3 x hist/scan/scatter) -> gather -> fill+bin -> step
 * with no host synchronisation and no allocation.
 */
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>
#include "bdem_abi.h"
#include "bdem_ctx.hpp"
#include "bdem_kernels.hpp"

using namespace bdem;


/* Stage boundaries recorded per step when timing is on:
 * 0 start | sort (hash if needed, resort or full sort, gather, cell table)
 * | lists (skin-list marking and rebuilding) | step (fused kernel) */
static const int BDEM_NSTAGE = 3;

namespace {

inline float bits_f(int v) { float f; std::memcpy(&f, &v, 4); return f; }
inline int f_bits(float f) { int v; std::memcpy(&v, &f, 4); return v; }

int fail(bdem_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPCHK(ctx, expr)                                                          \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess)                                                          \
      return fail(ctx, BDEM_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <class T>
int dalloc(bdem_ctx* c, T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc((void**)p, count * sizeof(T));
  if (e != hipSuccess) return fail(c, BDEM_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  c->allocs.push_back((void*)*p);
  return BDEM_OK;
}

/* Frees one allocation made by dalloc. */
template <class T>
void dfree(bdem_ctx* c, T*& p) {
  if (!p) return;
  auto it = std::find(c->allocs.begin(), c->allocs.end(), (void*)p);
  if (it != c->allocs.end()) c->allocs.erase(it);
  (void)hipFree((void*)p);
  p = nullptr;
}

void drop_graphs(bdem_ctx* c) {
  if ((c->graph[0] || c->graph[1]) && c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto& g : c->graph) {
    if (g) (void)hipGraphExecDestroy(g);
    g = nullptr;
  }
}

void free_all(bdem_ctx* c) {
  drop_graphs(c);
  for (void* p : c->allocs) (void)hipFree(p);
  c->allocs.clear();
  c->b = StepBuffers{};
  c->pb = PolyBuffers{};
  c->dshape = nullptr;
  c->shapes.clear();
  c->shape_by_pid = false;
  c->dp = nullptr;
  c->configured = false;
}

/* Host derivation of the history-law constants; the oracle restates the
 * same formulas (beta = ln e / sqrt(ln^2 e + pi^2), Hertz E*, G*). */
void derive_laws(const bdem_ext_params& ext, DevParams& P) {
  const double pi = 3.14159265358979323846;
  auto beta_of = [&](double e) {
    return (e > 0.0 && e < 1.0) ? std::log(e) / std::sqrt(std::log(e) * std::log(e) + pi * pi) : 0.0;
  };
  /* Hertz-Mindlin tangential damping over normal: sqrt(St / Sn) = sqrt(4 G* / E*) */
  auto tg_of = [](float estar, float gstar) {
    return estar > 0.0f ? float(std::sqrt(4.0 * double(gstar) / double(estar))) : 0.0f;
  };
  for (int a = 0; a < MAX_TYPES; ++a) {
    const bdem_material& ma = ext.material[a];
    for (int b = 0; b < MAX_TYPES; ++b) {
      const bdem_material& mb = ext.material[b];
      const double e = 0.5 * (double(ma.restitution) + double(mb.restitution));
      const double kn = 0.5 * (double(ma.kn) + double(mb.kn));
      const double ktr = 0.5 * (double(ma.kt_over_kn) + double(mb.kt_over_kn));
      const double mu = 0.5 * (double(ma.mu) + double(mb.mu));
      const double ea = ma.youngs, eb = mb.youngs, na = ma.poisson, nb = mb.poisson;
      double estar = 0.0, gstar = 0.0;
      if (ea > 0 && eb > 0) {
        estar = 1.0 / ((1.0 - na * na) / ea + (1.0 - nb * nb) / eb);
        gstar = 1.0 / (2.0 * (2.0 - na) * (1.0 + na) / ea + 2.0 * (2.0 - nb) * (1.0 + nb) / eb);
      }
      P.law_pp[a][b] = LawPair{float(kn), float(kn * ktr), float(beta_of(e)), float(mu),
                               float(estar), float(gstar), tg_of(float(estar), float(gstar)), 0.f};
    }
    const double ew = ext.wall_youngs > 0 ? ext.wall_youngs : ma.youngs;
    const double nw = ext.wall_youngs > 0 ? ext.wall_poisson : ma.poisson;
    double estar = 0.0, gstar = 0.0;
    if (ma.youngs > 0 && ew > 0) {
      estar = 1.0 / ((1.0 - double(ma.poisson) * ma.poisson) / ma.youngs + (1.0 - nw * nw) / ew);
      gstar = 1.0 / (2.0 * (2.0 - double(ma.poisson)) * (1.0 + double(ma.poisson)) / ma.youngs +
                     2.0 * (2.0 - nw) * (1.0 + nw) / ew);
    }
    P.law_w[a] = LawPair{ma.kn, float(double(ma.kn) * ma.kt_over_kn), float(beta_of(ma.restitution)),
                         ma.mu, float(estar), float(gstar), tg_of(float(estar), float(gstar)), 0.f};
  }
}

void build_params(bdem_ctx* c) {
  DevParams& P = c->hp;
  std::memset(&P, 0, sizeof(P));
  const bdem_SimulationInfo& s = c->sim;
  P.origin = s.worldOrigin;
  P.cell = s.cellSize;
  P.nx = s.num_NNCells.x; P.ny = s.num_NNCells.y; P.nz = s.num_NNCells.z;
  P.C = uint32_t(P.nx) * uint32_t(P.ny) * uint32_t(P.nz);
  P.dt = s.InitalDelta_t;
  P.gravity = s.Force_Field;
  P.roll_res = s.Roll_Res;
  P.rotation = s.Rotation ? 1 : 0;
  P.mode2d = s.Mode2D ? 1 : 0;
  P.sphere_orient = s.sphere_orient ? 1 : 0;
  P.poly = s.particle_type == 1 ? 1 : 0;
  P.law = c->has_ext ? c->ext.contact_law : BDEM_LAW_REFERENCE;
  P.per_particle = (c->has_ext && c->ext.per_particle_radius) ? 1 : 0;
  P.num_wobj = c->num_wobj_active;
  P.cyl_rotation = c->cyl_rotation ? 1 : 0;
  P.rotating_wobj = s.Rotating_WObject ? 1 : 0;
  P.mill_rot_vel = s.MillRotVel;
  P.num_cones = c->has_ext ? c->ext.num_cones : 0;
  P.kill_height = s.Silo_Kill_Height;
  P.max_size = s.max_size;
  P.kill_enabled = s.Kill_Particles ? 1 : 0;
  P.one_type = c->pobj.size() == 1 ? 1 : 0;   /* uploads reject types >= num_pobj */
  P.vbound_scale = float(c->vbound_scale);
  P.shape_by_pid = c->shape_by_pid ? 1 : 0;
  for (size_t t = 0; t < c->pobj.size() && t < size_t(MAX_TYPES); ++t) {
    const bdem_POBJ& po = c->pobj[t];
    P.t_radius[t] = po.radius;
    P.t_mass[t] = po.mass;
    P.t_inertia[t] = po.InertiaT[0];
    for (size_t u = 0; u < c->pobj.size() && u < size_t(MAX_TYPES); ++u) {
      /* POBJ::PP has 6 entries (KSimulationObjects.h:290) */
      if (u < 6) {
        P.pp_kn[t][u] = po.PP[u].Kn;
        P.pp_cn[t][u] = po.PP[u].Cn;
        P.pp_fk[t][u] = po.PP[u].Fric_kinetic;
      }
    }
    P.s_kn[t] = po.surface_Kn;
    P.s_cn[t] = po.surface_Cn;
    P.s_fs[t] = po.surface_Fric_static;
    P.s_fk[t] = po.surface_Fric_kinetic;
  }
  if (c->has_ext) {
    derive_laws(c->ext, P);
    for (int k = 0; k < c->ext.num_cones && k < BDEM_MAX_CONES; ++k) P.cone[k] = c->ext.cone[k];
  }
  for (size_t k = 0; k < c->wobj.size() && k < size_t(MAX_WOBJ); ++k) P.wobj[k] = c->wobj[k];
  P.num_dobj = int(c->dobj.size() < size_t(MAX_DOBJ) ? c->dobj.size() : size_t(MAX_DOBJ));
  for (int k = 0; k < P.num_dobj; ++k) P.dobj[k] = c->dobj[k];
  P.mill_sim = s.isMillSim ? 1 : 0;
  P.tally = c->d_tally;
  P.mill_cent = s.cent;
  P.dobj_bound_r = s.DObject_CylinderBoundR;
  for (size_t t = 0; t < c->pobj.size() && t < size_t(MAX_TYPES); ++t) {
    P.d_kn[t] = c->pobj[t].Dsurface_Kn;
    P.d_cn[t] = c->pobj[t].Dsurface_Cn;
    P.d_fk[t] = c->pobj[t].Dsurface_Fric_kinetic;
  }
}

/* The lifters the device has turned since the last upload, back into the
 * host table (so a later upload does not undo the turning). */
int pull_dobj(bdem_ctx* c) {
  if (!c->dobj_dirty || c->dobj.empty() || !c->dp) return BDEM_OK;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const size_t nd = c->dobj.size() < size_t(MAX_DOBJ) ? c->dobj.size() : size_t(MAX_DOBJ);
  HIPCHK(c, hipMemcpy(c->dobj.data(), reinterpret_cast<char*>(c->dp) + offsetof(DevParams, dobj),
                      nd * sizeof(bdem_DOBJ), hipMemcpyDeviceToHost));
  c->dobj_dirty = false;
  return BDEM_OK;
}

int upload_params(bdem_ctx* c) {
  drop_graphs(c);   /* captured launches may depend on what changes here */
  int rc0 = pull_dobj(c);
  if (rc0 != BDEM_OK) return rc0;
  build_params(c);
  HIPCHK(c, hipMemcpyAsync(c->dp, &c->hp, sizeof(DevParams), hipMemcpyHostToDevice, c->stream));
  std::vector<ShapeHdr> hdr;
  std::vector<float4> local;
  if (c->dshape) {
    HIPCHK(c, hipMemcpyAsync(c->dshape, c->shapes.data(), c->shapes.size() * sizeof(bdem_poly_shape),
                             hipMemcpyHostToDevice, c->stream));
    for (const bdem_poly_shape& s : c->shapes)
      hdr.push_back(ShapeHdr{s.num_vertex, s.num_faces, s.num_edges, 0, s.vbound, s.plane_slack, s.radius, 0.0f});
    HIPCHK(c, hipMemcpyAsync(c->pb.hdr, hdr.data(), hdr.size() * sizeof(ShapeHdr), hipMemcpyHostToDevice,
                             c->stream));
    const PolyLayout& L = c->pb.L;
    local.assign(c->shapes.size() * (size_t)L.stride, float4{0.f, 0.f, 0.f, 0.f});
    for (size_t r = 0; r < c->shapes.size(); ++r) {
      const bdem_poly_shape& s = c->shapes[r];
      float4* row = local.data() + r * (size_t)L.stride;
      for (int v = 0; v < s.num_vertex; ++v)
        row[v] = float4{s.Local_Vertex[v].x, s.Local_Vertex[v].y, s.Local_Vertex[v].z, 1.0f};
      for (int f = 0; f < s.num_faces; ++f) {
        const bdem_Surface& F = s.face[f];
        row[L.pv + f] = float4{F.normal.x, F.normal.y, F.normal.z, 1.0f};
        /* the float differences the turned centroid is made from */
        const float dx = F.centroid.x - s.COM.x, dy = F.centroid.y - s.COM.y, dz = F.centroid.z - s.COM.z;
        row[L.pv + L.pf + f] = float4{dx, dy, dz, 1.0f};
      }
    }
    HIPCHK(c, hipMemcpyAsync(c->pb.local, local.data(), local.size() * sizeof(float4), hipMemcpyHostToDevice,
                             c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return BDEM_OK;
}

bool need_history(const bdem_ctx* c) {
  return c->has_ext && c->ext.contact_law != BDEM_LAW_REFERENCE;
}

/* The per-particle word carried with the state: spring counts (history
 * laws) and the last lifter contact type (Contact_Info::cont_type). */
int ensure_counts(bdem_ctx* c) {
  if (c->b.hwc_cur || (!need_history(c) && c->dobj.empty())) return BDEM_OK;
  int rc;
  if ((rc = dalloc(c, &c->b.hwc_cur, c->b.n)) != BDEM_OK) return rc;
  if ((rc = dalloc(c, &c->b.hwc_srt, c->b.n)) != BDEM_OK) return rc;
  launch_fill_u32(c->b.hwc_cur, c->b.n, 0u, c->stream);
  launch_fill_u32(c->b.hwc_srt, c->b.n, 0u, c->stream);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return BDEM_OK;
}

/* DOBJ tables (KSimulationObjects.h:318-342): index ranges the kernels use. */
int check_dobj(bdem_ctx* c, const bdem_DOBJ* d, int nd) {
  for (int k = 0; k < nd; ++k) {
    const bdem_DOBJ& D = d[k];
    if (D.num_vertex < 0 || D.num_vertex > 8 || D.num_faces < 0 || D.num_faces > 3 || D.num_edges < 0 ||
        D.num_edges > 4)
      return fail(c, BDEM_EINVAL, "dynamic object vertex/face/edge count out of range");
    for (int f = 0; f < D.num_faces; ++f) {
      const bdem_Surface& S = D.faces[f];
      if (S.num_vertex < 1 || S.num_vertex > 8) return fail(c, BDEM_EINVAL, "lifter face vertex count out of range");
      for (int q = 0; q < S.num_vertex; ++q)
        if (S.vertex_Order[q] < 0 || S.vertex_Order[q] >= D.num_vertex)
          return fail(c, BDEM_EINVAL, "lifter face vertex index out of range");
    }
    for (int e = 0; e < D.num_edges; ++e)
      if (D.edge[e].point_vindex[0] >= (unsigned)D.num_vertex || D.edge[e].point_vindex[1] >= (unsigned)D.num_vertex)
        return fail(c, BDEM_EINVAL, "lifter edge vertex index out of range");
  }
  return BDEM_OK;
}

/* WOBJ tables (KSimulationObjects.h:104-152): surface kinds and the index
 * ranges the wall kernels read (is_PointOnSurface walks vertex_Order).
 * Called by every entry point that uploads a world-object table. */
int check_wobj(bdem_ctx* c, const bdem_WOBJ* wobj, int num_wobj) {
  for (int k = 0; k < num_wobj; ++k) {
    const bdem_WOBJ& w = wobj[k];
    if (w.surface_type != BDEM_SURFACE_PLANE && w.surface_type != BDEM_SURFACE_CYLINDER)
      return fail(c, BDEM_EINVAL, "unknown world-object surface type");
    if (w.surface_type != BDEM_SURFACE_PLANE) continue;
    if (w.num_surfaces < 0 || w.num_surfaces > 16 || w.num_vertex < 0 || w.num_vertex > 16)
      return fail(c, BDEM_EINVAL, "world object surface/vertex count out of range");
    for (int s = 0; s < w.num_surfaces; ++s) {
      const bdem_Surface& S = w.surfaces[s];
      if (S.area == -1.0f) continue;  /* infinite plane: no polygon */
      if (S.num_vertex < 1 || S.num_vertex > 8)
        return fail(c, BDEM_EINVAL, "surface vertex count out of range");
      for (int q = 0; q < S.num_vertex; ++q)
        if (S.vertex_Order[q] < 0 || S.vertex_Order[q] >= w.num_vertex)
          return fail(c, BDEM_EINVAL, "surface vertex index out of range");
    }
  }
  return BDEM_OK;
}

/* POBJ polyhedron tables (KSimulationObjects.h:272-300) and per-particle
 * shape tables: the index ranges the polyhedra kernels read. */
template <class T>
int check_poly(bdem_ctx* c, const T* p, int np, int max_vertex, int max_face, int max_edge) {
  for (int t = 0; t < np; ++t) {
    const T& P = p[t];
    if (P.num_vertex < 1 || P.num_vertex > max_vertex || P.num_faces < 1 || P.num_faces > max_face ||
        P.num_edges < 0 || P.num_edges > max_edge)
      return fail(c, BDEM_EINVAL, "polyhedron vertex/face/edge count out of range (row " + std::to_string(t) + ")");
    if (!(P.radius > 0.0f) || !(P.mass > 0.0f))
      return fail(c, BDEM_EINVAL, "polyhedron radius and mass must be positive");
    for (int f = 0; f < P.num_faces; ++f) {
      const bdem_Surface& S = P.face[f];
      if (S.num_vertex < 1 || S.num_vertex > 8) return fail(c, BDEM_EINVAL, "polyhedron face vertex count out of range");
      /* unit normals: the separating-plane test reads face-plane distances as
       * lengths (its vertex-bound shortcut is exact only then) */
      const double nn = std::sqrt(double(S.normal.x) * S.normal.x + double(S.normal.y) * S.normal.y +
                                  double(S.normal.z) * S.normal.z);
      if (!(std::fabs(nn - 1.0) <= 1e-5))
        return fail(c, BDEM_EINVAL, "polyhedron face normal is not unit length (row " + std::to_string(t) +
                                        ", face " + std::to_string(f) + ")");
      for (int q = 0; q < S.num_vertex; ++q)
        if (S.vertex_Order[q] < 0 || S.vertex_Order[q] >= P.num_vertex)
          return fail(c, BDEM_EINVAL, "polyhedron face vertex index out of range");
    }
    for (int e = 0; e < P.num_edges; ++e)
      if (P.edge[e].point_vindex[0] >= (unsigned)P.num_vertex || P.edge[e].point_vindex[1] >= (unsigned)P.num_vertex)
        return fail(c, BDEM_EINVAL, "polyhedron edge vertex index out of range");
  }
  return BDEM_OK;
}

/* Largest |Local_Vertex| x 1.0001 (cm), in double: a bound on the distance
 * of any turned vertex from the centre (unit quaternion). */
float vertex_bound(const bdem_float3* lv, int nv) {
  double vmax = 0.0;
  for (int v = 0; v < nv; ++v) {
    const double x = lv[v].x, y = lv[v].y, z = lv[v].z;
    vmax = std::fmax(vmax, std::sqrt(x * x + y * y + z * z));
  }
  return float(vmax * 1.0001);
}

/* Largest height of a vertex above a face plane of the same body, in
 * double (0 when the faces support the body: a convex hull up to rounding).
 * The edge search uses the face planes as bounds only up to this slack. */
float plane_slack(const bdem_float3* lv, int nv, const bdem_Surface* face, int nf, bdem_float3 com) {
  double worst = 0.0;
  for (int f = 0; f < nf; ++f) {
    const bdem_Surface& S = face[f];
    for (int v = 0; v < nv; ++v) {
      const double h = double(S.normal.x) * (double(lv[v].x) + com.x - S.centroid.x) +
                       double(S.normal.y) * (double(lv[v].y) + com.y - S.centroid.y) +
                       double(S.normal.z) * (double(lv[v].z) + com.z - S.centroid.z);
      worst = std::fmax(worst, h);
    }
  }
  return float(worst * 1.0001 + 1e-6);
}

/* The geometric part of a POBJ row as a shape-table row. */
bdem_poly_shape shape_of_pobj(const bdem_POBJ& po) {
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
  s.vbound = vertex_bound(po.Local_Vertex, po.num_vertex);
  s.plane_slack = plane_slack(po.Local_Vertex, po.num_vertex, po.face, po.num_faces, po.COM);
  return s;
}

/* The shape-table row of a packed P_ID * 8 + type word (the device's
 * shape_index), or -1 when the table has no such row. */
int shape_row(const bdem_ctx* c, int packed) {
  int row = packed & 7;
  if (c->shape_by_pid) row = (packed >> 3) > 0 ? (packed >> 3) : 0;
  return row < (int)c->shapes.size() ? row : -1;
}

/* rot's row layout from the largest vertex and face counts of the table. */
PolyLayout poly_layout(const std::vector<bdem_poly_shape>& t) {
  int pv = 0, pf = 0;
  for (const bdem_poly_shape& s : t) {
    pv = s.num_vertex > pv ? s.num_vertex : pv;
    pf = s.num_faces > pf ? s.num_faces : pf;
  }
  return PolyLayout{pv, pf, pv + 2 * pf};
}

/* Skin masks for the sphere step: allowed when the hash is injective
 * (Nx == Nz, so cells can be recovered from keys); the buffers are
 * allocated on first use. */
int apply_skin(bdem_ctx* c) {
  StepBuffers& b = c->b;
  float skin = c->skin_req > 0.0f ? c->skin_req : 0.0f;   /* off unless asked for */
  if (c->poly || c->sim.num_NNCells.x != c->sim.num_NNCells.z || !(skin > 0.0f)) skin = 0.0f;
  if (skin > 0.0f && !b.vmeta) {
    int rc;
    const size_t n = (size_t)b.n;
    if ((rc = dalloc(c, &b.mask_cur, n)) != BDEM_OK) return rc;
    if ((rc = dalloc(c, &b.mask_srt, n)) != BDEM_OK) return rc;
    if ((rc = dalloc(c, &b.inv_map, n)) != BDEM_OK) return rc;
    if ((rc = dalloc(c, &b.rowflag, n)) != BDEM_OK) return rc;
    if ((rc = dalloc(c, &b.vmeta, VM_WORDS)) != BDEM_OK) return rc;
    launch_fill_u32(reinterpret_cast<uint32_t*>(b.mask_cur), 2 * n, 0u, c->stream);
    launch_fill_u32(reinterpret_cast<uint32_t*>(b.mask_srt), 2 * n, 0u, c->stream);
    launch_fill_u32(b.inv_map, n, 0u, c->stream);
    launch_fill_u32(b.rowflag, n, 0u, c->stream);
    launch_fill_u32(b.vmeta, VM_WORDS, 0u, c->stream);
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  b.skin = skin;
  c->lists_full = true;
  return BDEM_OK;
}

SkinArgs skin_args(bdem_ctx* c) {
  const StepBuffers& b = c->b;
  SkinArgs a{};
  a.on = b.skin > 0.0f ? 1 : 0;
  if (!a.on) return a;
  a.mask_in = b.mask_srt;
  a.mask_out = b.mask_cur;
  a.rowflag = b.rowflag;
  a.vmeta = b.vmeta;
  a.skin = b.skin;
  a.par = c->vpar ^ 1;
  return a;
}

int require(bdem_ctx* c) {
  if (!c) return BDEM_EINVAL;
  if (!c->configured) return fail(c, BDEM_ESTATE, "bdem_set_simulation_data has not been called");
  return BDEM_OK;
}

/* Host AoS -> device float4 streams.  Radius and mass come from the type
 * table (the reference's only source) unless per-particle values exist. */
int upload_state(bdem_ctx* c, const bdem_float3* pos, const bdem_Quaterion* ornt,
                 const bdem_float3* vel, const bdem_float3* avel, const unsigned int* ptype,
                 const int* pid, bool keep_vel) {
  const int n = c->b.n;
  /* every check before the first copy: a rejected upload leaves the state,
   * keys and flags of the previous one untouched */
  const bool hist = need_history(c);
  const bool check_q = c->orient && ornt && !c->sim.sphere_orient;
  for (int i = 0; i < n; ++i) {
    const unsigned t = ptype[i];
    if (t >= c->pobj.size() || t >= (unsigned)MAX_TYPES)
      return fail(c, BDEM_EINVAL, "particle type out of range at index " + std::to_string(i));
    /* P_ID * 8 + type is packed into one int; -1 marks a dead particle */
    if (pid[i] < -1 || pid[i] >= (1 << 28))
      return fail(c, BDEM_EINVAL, "P_ID out of range [-1, 2^28) at index " + std::to_string(i));
    if (c->shape_by_pid && pid[i] >= (int)c->shapes.size())
      return fail(c, BDEM_EINVAL, "P_ID beyond the per-particle shape table at index " + std::to_string(i));
    if (hist && pid[i] >= c->b.hist_cap)
      return fail(c, BDEM_EINVAL, "history laws need P_ID < Num_Particles (index " + std::to_string(i) + ")");
    if (check_q && pid[i] >= 0) {
      const double s2 = double(ornt[i].w) * ornt[i].w + double(ornt[i].x) * ornt[i].x +
                        double(ornt[i].y) * ornt[i].y + double(ornt[i].z) * ornt[i].z;
      if (!std::isfinite(s2))
        return fail(c, BDEM_EINVAL, "orientation quaternion is not finite at index " + std::to_string(i));
    }
  }
  /* The reference turns vertices by q v q* as the state gives q, unit or not
   * (the unit-test start files hold (1, 0.5, 0, 0)); q v q* scales by |q|^2
   * until the first step renormalises q.  The separating-plane shortcut's
   * vertex bound is scaled by the largest |q|^2 (at least 1) of the upload,
   * so it stays a bound. */
  double qscale = 1.0;
  if (check_q)
    for (int i = 0; i < n; ++i)
      if (pid[i] >= 0) {
        const double s2 = double(ornt[i].w) * ornt[i].w + double(ornt[i].x) * ornt[i].x +
                          double(ornt[i].y) * ornt[i].y + double(ornt[i].z) * ornt[i].z;
        qscale = std::max(qscale, s2 * (1.0 + 1e-6));
      }
  std::vector<float4> P(n), V(n), O(n), Q;
  std::vector<float4> oldP, oldV, oldO;
  if (keep_vel || c->hp.per_particle) {
    oldP.resize(n); oldV.resize(n); oldO.resize(n);
    HIPCHK(c, hipMemcpy(oldP.data(), c->b.pos_cur, n * sizeof(float4), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(oldV.data(), c->b.vel_cur, n * sizeof(float4), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(oldO.data(), c->b.omg_cur, n * sizeof(float4), hipMemcpyDeviceToHost));
  }
  for (int i = 0; i < n; ++i) {
    const unsigned t = ptype[i];
    float r = c->hp.per_particle ? oldP[i].w : c->pobj[t].radius;
    float m = c->hp.per_particle ? oldV[i].w : c->pobj[t].mass;
    if (c->shape_by_pid) {   /* per-particle shapes carry their own radius and mass */
      const bdem_poly_shape& sh = c->shapes[pid[i] > 0 ? pid[i] : 0];
      r = sh.radius;
      m = sh.mass;
    }
    P[i] = make_float4(pos[i].x, pos[i].y, pos[i].z, r);
    if (vel) V[i] = make_float4(vel[i].x, vel[i].y, vel[i].z, m);
    else V[i] = make_float4(oldV[i].x, oldV[i].y, oldV[i].z, m);
    if (avel) O[i] = make_float4(avel[i].x, avel[i].y, avel[i].z, bits_f(pack_id_type(pid[i], t)));
    else O[i] = make_float4(oldO[i].x, oldO[i].y, oldO[i].z, bits_f(pack_id_type(pid[i], t)));
  }
  /* from here on the device state changes: whatever happens, the next step
   * re-hashes and fully re-sorts */
  c->keys_valid = false;
  c->flags_valid = false;
  c->lists_full = true;
  HIPCHK(c, hipMemcpy(c->b.pos_cur, P.data(), n * sizeof(float4), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->b.vel_cur, V.data(), n * sizeof(float4), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->b.omg_cur, O.data(), n * sizeof(float4), hipMemcpyHostToDevice));
  if (c->orient) {
    Q.resize(n);
    for (int i = 0; i < n; ++i) {
      if (ornt && !c->sim.sphere_orient) Q[i] = make_float4(ornt[i].w, ornt[i].x, ornt[i].y, ornt[i].z);
      else Q[i] = make_float4(1.f, 0.f, 0.f, 0.f);  /* Set_Pos: identity for sphere_orient */
    }
    HIPCHK(c, hipMemcpy(c->b.quat_cur, Q.data(), n * sizeof(float4), hipMemcpyHostToDevice));
  }
  if (c->b.hwc_cur) {   /* spring counts and lifter contact types restart */
    launch_fill_u32(c->b.hwc_cur, c->b.n, 0u, c->stream);
    launch_fill_u32(c->b.hwc_srt, c->b.n, 0u, c->stream);
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  c->keys_valid = false;
  c->lists_full = true;
  if (c->poly && qscale != c->vbound_scale) {
    c->vbound_scale = qscale;
    int rc = upload_params(c);
    if (rc != BDEM_OK) return rc;
  }
  return BDEM_OK;
}

int download(bdem_ctx* c, std::vector<float4>& P, std::vector<float4>& V, std::vector<float4>& O,
             std::vector<float4>& Q) {
  const int n = c->b.n;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  P.resize(n); V.resize(n); O.resize(n);
  HIPCHK(c, hipMemcpy(P.data(), c->b.pos_cur, n * sizeof(float4), hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(V.data(), c->b.vel_cur, n * sizeof(float4), hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(O.data(), c->b.omg_cur, n * sizeof(float4), hipMemcpyDeviceToHost));
  if (c->orient) {
    Q.resize(n);
    HIPCHK(c, hipMemcpy(Q.data(), c->b.quat_cur, n * sizeof(float4), hipMemcpyDeviceToHost));
  }
  return BDEM_OK;
}

hipEvent_t* timing_slot(bdem_ctx* c) {
  if (c->t_used >= c->t_cap) return nullptr;
  return &c->tev[(size_t)c->t_used * (BDEM_NSTAGE + 1)];
}

/* hash (if needed) + sort + gather + cells: the sorted scratch state.
 * On return b.keys_sorted[c->sorted_par] and b.idx_sorted hold the order. */
int prepare_sorted(bdem_ctx* c, hipEvent_t* ev = nullptr) {
  if (ev) (void)hipEventRecord(ev[0], c->stream);
  const bool incremental = c->keys_valid && c->flags_valid && !c->full_next;
  if (!c->keys_valid) launch_hash(c->dp, c->b, c->stream);
  const int prev = c->sorted_par;
  const int next = incremental ? prev ^ 1 : prev;
  SortArgs a{};
  const StepBuffers& b = c->b;
  a.n = b.n;
  a.C = c->hp.C;
  a.passes = sort_passes(c->hp.C + 1u);
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
  a.quat_cur = c->orient ? b.quat_cur : nullptr;
  a.pos_srt = b.pos_srt; a.vel_srt = b.vel_srt; a.omg_srt = b.omg_srt;
  a.quat_srt = c->orient ? b.quat_srt : nullptr;
  a.hwc_cur = b.hwc_cur;
  a.hwc_srt = b.hwc_srt;
  a.inv_map = b.skin > 0.0f ? b.inv_map : nullptr;
  a.mask_cur = b.skin > 0.0f ? b.mask_cur : nullptr;
  a.mask_srt = b.skin > 0.0f ? b.mask_srt : nullptr;
  hipError_t e = launch_sort(a, incremental, c->append_par, c->stream);
  if (ev) (void)hipEventRecord(ev[1], c->stream);
  if (e != hipSuccess) return fail(c, BDEM_EHIP, std::string("sort launch: ") + hipGetErrorString(e));
  c->sorted_par = next;
  c->append_par ^= 1;
  c->full_next = false;
  if (!incremental) c->lists_full = true;   /* masks were not carried through: test every row */
  /* the step kernel must produce new keys and flags before the next sort */
  c->keys_valid = false;
  c->flags_valid = false;
  HIPCHK(c, hipGetLastError());
  return BDEM_OK;
}

/* Two steps whose launches are the same every time: the incremental path
 * of a sphere context with no discharge check, skin masks, stage timing or
 * recording in them.  Every per-step alternation (sorted keys, append
 * counters, spring buffers) has period two, so such a pair is captured
 * once per parity and replayed. */
bool graph_pair_ok(const bdem_ctx* c) {
  if (!c->graphs_on || c->poly || c->tev.size() || c->dbg_list || c->dbg_nn) return false;
  if (!(c->keys_valid && c->flags_valid && !c->full_next)) return false;
  if (c->b.skin > 0.0f) return false;
  if (c->sim.Kill_Particles) {
    const int interval = (c->has_ext && c->ext.kill_check_interval > 0) ? c->ext.kill_check_interval : 1000;
    if (c->stepc + 2 >= interval) return false;   /* a check falls in the pair */
  }
  return true;
}

int graph_pair(bdem_ctx* c) {
  const int key = c->append_par;
  hipGraphExec_t& g = c->graph[key];
  if (g && (c->graph_hist[key] != c->hist_par || c->graph_sorted[key] != c->sorted_par)) {
    (void)hipGraphExecDestroy(g);
    g = nullptr;
  }
  if (!g) {
    const int hp = c->hist_par, sp = c->sorted_par, ap = c->append_par, st = c->stepc;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = bdem_update_sim(c);
    if (rc == BDEM_OK) rc = bdem_update_sim(c);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &graph);
    if (rc != BDEM_OK) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    if (e != hipSuccess) return fail(c, BDEM_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    if (c->hist_par != hp || c->sorted_par != sp || c->append_par != ap) {
      (void)hipGraphDestroy(graph);
      return fail(c, BDEM_ESTATE, "internal: a step pair did not return to its parity");
    }
    const hipError_t ei = hipGraphInstantiate(&g, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) {
      g = nullptr;
      return fail(c, BDEM_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
    }
    c->graph_hist[key] = hp;
    c->graph_sorted[key] = sp;
    c->stepc = st;   /* captured, not run: the replay below runs the pair */
  }
  HIPCHK(c, hipGraphLaunch(g, c->stream));
  if (c->sim.Kill_Particles) c->stepc += 2;
  if (c->lift_on && !c->dobj.empty()) c->dobj_dirty = true;
  return BDEM_OK;
}

}  // namespace

namespace bdem {
/* The parameter block rebuilt from the context's tables and uploaded (for
 * the adjoint runs, which change the pair constants between runs). */
int ctx_upload_params(bdem_ctx* c) { return upload_params(c); }
}  // namespace bdem

extern "C" {

int bdem_create(bdem_ctx** out, int device) {
  if (!out) return BDEM_EINVAL;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return BDEM_EHIP;
  if (device < 0 || device >= ndev) return BDEM_EINVAL;
  bdem_ctx* c = new bdem_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return BDEM_EHIP;
  }
  *out = c;
  return BDEM_OK;
}

int bdem_destroy(bdem_ctx* c) {
  if (!c) return BDEM_EINVAL;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (hipEvent_t e : c->tev) (void)hipEventDestroy(e);
  free_all(c);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return BDEM_OK;
}

const char* bdem_last_error(const bdem_ctx* c) { return c ? c->err.c_str() : "null context"; }

int bdem_set_simulation_data(bdem_ctx* c, const bdem_WOBJ* wobj, const bdem_POBJ* pobj,
                             const bdem_DOBJ* dobj, const bdem_SimulationInfo* sim, int num_wobj,
                             int num_pobj, int num_dobj, const bdem_InitConfig* init) {
  if (!c || !sim || !pobj || (num_wobj > 0 && !wobj)) return c ? fail(c, BDEM_EINVAL, "null argument") : BDEM_EINVAL;
  if (sim->particle_type != 0 && sim->particle_type != 1)
    return fail(c, BDEM_EINVAL, "particle_type must be 0 (spheres) or 1 (polyhedra)");
  if (sim->Integration_Type != 0) return fail(c, BDEM_ENOTSUP, "only symplectic Euler (Integration_Type 0) is live in the reference");
  if (sim->use_symmetry) return fail(c, BDEM_ENOTSUP, "the symmetric path is disabled in the reference (main.cpp:811)");
  if (num_wobj < 0 || num_wobj > MAX_WOBJ) return fail(c, BDEM_EINVAL, "num_wobj must be in [0, 8]");
  if (num_pobj <= 0 || num_pobj > MAX_TYPES) return fail(c, BDEM_EINVAL, "num_pobj must be in [1, 8]");
  if (sim->Num_Particles <= 0) return fail(c, BDEM_EINVAL, "Num_Particles must be positive");
  if (sim->num_NNCells.x <= 0 || sim->num_NNCells.y <= 0 || sim->num_NNCells.z <= 0)
    return fail(c, BDEM_EINVAL, "num_NNCells must be positive");
  const unsigned long long C = (unsigned long long)sim->num_NNCells.x * sim->num_NNCells.y * sim->num_NNCells.z;
  if (C + 2 > 0xFFFFFFF0ull) return fail(c, BDEM_EINVAL, "grid too large for 32-bit keys");
  if (num_dobj < 0 || num_dobj > MAX_DOBJ) return fail(c, BDEM_EINVAL, "num_dobj must be in [0, 32]");
  if (num_dobj > 0 && !dobj) return fail(c, BDEM_EINVAL, "null dynamic-object table");
  int rc0 = BDEM_OK;
  if (num_dobj > 0 && (rc0 = check_dobj(c, dobj, num_dobj)) != BDEM_OK) return rc0;
  /* use_hist (Kernel_ParticleInteraction_Polyhedra_NonSymmetryHist,
   * DEM_Kernels_Compute.cuh:998-1329) is accepted and changes nothing: the
   * reference's set-up always clears it (main.cpp:830) and its kernel is
   * undefined as written.  The polyhedra step keeps its own contact memory
   * in every mode, the face that separated each listed pair at the previous
   * step, which it tries first; that memory never changes a result. */
  if (sim->particle_type == 1) {
    if ((rc0 = check_poly(c, pobj, num_pobj, 32, 32, 32)) != BDEM_OK) return rc0;
  }
  if ((rc0 = check_wobj(c, wobj, num_wobj)) != BDEM_OK) return rc0;

  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  free_all(c);
  c->sim = *sim;
  c->wobj.assign(wobj, wobj + num_wobj);
  c->pobj.assign(pobj, pobj + num_pobj);
  c->dobj.assign(dobj ? dobj : nullptr, dobj ? dobj + num_dobj : nullptr);
  c->num_wobj_active = num_wobj;
  c->has_ext = false;
  std::memset(&c->ext, 0, sizeof(c->ext));
  c->poly = sim->particle_type == 1;
  c->orient = sim->sphere_orient || c->poly;
  c->vbound_scale = 1.0;
  c->stepc = 0;
  c->hist_par = 0;
  c->lift_on = false;
  c->dobj_dirty = false;
  std::memset(c->tallys, 0, sizeof(c->tallys));

  const int n = sim->Num_Particles;
  StepBuffers& b = c->b;
  b.n = n;
  int rc;
#define ALLOC(p, cnt) if ((rc = dalloc(c, &(p), (cnt))) != BDEM_OK) return rc
  ALLOC(c->dp, 1);
  ALLOC(b.pos_cur, n); ALLOC(b.vel_cur, n); ALLOC(b.omg_cur, n);
  ALLOC(b.pos_srt, n); ALLOC(b.vel_srt, n); ALLOC(b.omg_srt, n);
  if (c->orient) { ALLOC(b.quat_cur, n); ALLOC(b.quat_srt, n); }
  ALLOC(b.keys_new, n); ALLOC(b.keys_sorted[0], n); ALLOC(b.keys_sorted[1], n);
  ALLOC(b.idx_sorted, n);
  ALLOC(b.changed, n);
  ALLOC(b.sbuf, n);
  ALLOC(b.skey, n);
  ALLOC(b.meta, META_WORDS);
  ALLOC(b.rk[0], n); ALLOC(b.rk[1], n); ALLOC(b.rv[0], n); ALLOC(b.rv[1], n);
  ALLOC(b.hist, (size_t)RADIX * sort_blocks(n));
  ALLOC(b.cells, (size_t)C);
  ALLOC(b.status, 1); ALLOC(b.kill_count, 1);
  ALLOC(c->d_tally, 8);
  ALLOC(c->d_stats, 1);
  ALLOC(c->d_profile, PROFILE_MAX_BINS);
  ALLOC(b.wall_force, (size_t)n * 3);
  if (c->poly) {
    PolyBuffers& pb = c->pb;
    for (int t = 0; t < num_pobj; ++t) c->shapes.push_back(shape_of_pobj(pobj[t]));
    pb.L = poly_layout(c->shapes);
    ALLOC(c->dshape, c->shapes.size());
    ALLOC(pb.hdr, c->shapes.size());
    ALLOC(pb.local, c->shapes.size() * (size_t)pb.L.stride);
    ALLOC(pb.rot, (size_t)n * pb.L.stride);
    ALLOC(pb.list, (size_t)n * LIST_STRIDE);
    ALLOC(pb.nn, (size_t)n);
    ALLOC(pb.queue, (size_t)n * LIST_STRIDE);
    ALLOC(pb.equeue, (size_t)n * LIST_STRIDE);
    ALLOC(pb.qcount, 4);
    ALLOC(pb.pf, (size_t)n * LIST_STRIDE);
    ALLOC(pb.pt, (size_t)n * LIST_STRIDE);
    ALLOC(pb.hint, (size_t)n * LIST_STRIDE);
    launch_fill_u32(reinterpret_cast<uint32_t*>(pb.hint), (size_t)n * LIST_STRIDE * 2, 0xFFFFFFFFu, c->stream);
  }
#undef ALLOC
  HIPCHK(c, hipMemset(b.status, 0, sizeof(uint32_t)));
  HIPCHK(c, hipMemset(b.kill_count, 0, sizeof(int)));
  HIPCHK(c, hipMemset(c->d_tally, 0, 8 * sizeof(float)));   /* Set_Tallys with zeros (:1644-1660) */
  HIPCHK(c, hipMemset(b.meta, 0, META_WORDS * sizeof(uint32_t)));
  HIPCHK(c, hipMemset(b.wall_force, 0, (size_t)n * 3 * sizeof(float)));
  HIPCHK(c, hipMemset(b.pos_cur, 0, (size_t)n * sizeof(float4)));
  c->configured = true;
  if ((rc = upload_params(c)) != BDEM_OK) return rc;
  if ((rc = ensure_counts(c)) != BDEM_OK) return rc;
  c->skin_req = -1.0f;
  c->vpar = 0;
  if ((rc = apply_skin(c)) != BDEM_OK) return rc;

  /* Set_Init_Config_Device (GPU_Memory_Transfer.cuh:205-271): velocity =
   * InitVel[0], angular velocity = 0.  InitVel[0] is init->velocity[0] for a
   * grid start, launch_Vel for a mill fill, zero when reading a file
   * (Device/DeviceInterface.cu:1708-1735). */
  bdem_float3 v0{0.f, 0.f, 0.f};
  if (init && !init->use_file) v0 = init->grid_type < 2 ? init->velocity[0] : init->launch_Vel;
  std::vector<float4> P(n), V(n), O(n);
  for (int i = 0; i < n; ++i) {
    P[i] = make_float4(0.f, 0.f, 0.f, pobj[0].radius);
    V[i] = make_float4(v0.x, v0.y, v0.z, pobj[0].mass);
    O[i] = make_float4(0.f, 0.f, 0.f, bits_f(pack_id_type(-1, 0)));
  }
  HIPCHK(c, hipMemcpy(b.pos_cur, P.data(), n * sizeof(float4), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(b.vel_cur, V.data(), n * sizeof(float4), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(b.omg_cur, O.data(), n * sizeof(float4), hipMemcpyHostToDevice));
  if (c->orient) {
    std::vector<float4> Q(n, make_float4(1.f, 0.f, 0.f, 0.f));
    HIPCHK(c, hipMemcpy(b.quat_cur, Q.data(), n * sizeof(float4), hipMemcpyHostToDevice));
  }
  c->keys_valid = false;
  std::memset(c->tallys, 0, sizeof(c->tallys));
  return BDEM_OK;
}

int bdem_set_ext_params(bdem_ctx* c, const bdem_ext_params* ext) {
  int rc = require(c);
  if (rc) return rc;
  if (!ext) return fail(c, BDEM_EINVAL, "null ext");
  if (ext->contact_law < 0 || ext->contact_law > 2) return fail(c, BDEM_EINVAL, "unknown contact law");
  if (ext->num_cones < 0 || ext->num_cones > BDEM_MAX_CONES) return fail(c, BDEM_EINVAL, "num_cones out of range");
  if (c->poly && (ext->contact_law != BDEM_LAW_REFERENCE || ext->per_particle_radius))
    return fail(c, BDEM_ENOTSUP, "polyhedra use the reference pair law and their type's bounding radius");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->ext = *ext;
  c->has_ext = true;
  const int n = c->b.n;
  if (need_history(c) && !c->b.hrec[0]) {
    c->b.hist_cap = n;
    for (int h = 0; h < 2; ++h) {
      if ((rc = dalloc(c, &c->b.hrec[h], (size_t)n * HMAX)) != BDEM_OK) return rc;
      if ((rc = dalloc(c, &c->b.wrec[h], (size_t)n * WMAX)) != BDEM_OK) return rc;
    }
  }
  if ((rc = ensure_counts(c)) != BDEM_OK) return rc;
  return upload_params(c);
}

int bdem_update_simulation_info(bdem_ctx* c, const bdem_SimulationInfo* sim) {
  int rc = require(c);
  if (rc) return rc;
  if (!sim) return fail(c, BDEM_EINVAL, "null argument");
  const bdem_SimulationInfo& o = c->sim;
  if (sim->Num_Particles != o.Num_Particles || sim->num_NNCells.x != o.num_NNCells.x ||
      sim->num_NNCells.y != o.num_NNCells.y || sim->num_NNCells.z != o.num_NNCells.z ||
      sim->particle_type != o.particle_type || sim->sphere_orient != o.sphere_orient)
    return fail(c, BDEM_EINVAL, "particle count, grid size, particle type and orientation are fixed at setup");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const bool same_grid = std::memcmp(&sim->worldOrigin, &o.worldOrigin, sizeof(bdem_float3)) == 0 &&
                         std::memcmp(&sim->cellSize, &o.cellSize, sizeof(bdem_float3)) == 0;
  c->sim = *sim;
  if (!same_grid) c->keys_valid = false;
  int rc2 = apply_skin(c);   /* the default skin follows the cell size; lists rebuilt */
  if (rc2) return rc2;
  return upload_params(c);
}

int bdem_set_pos(bdem_ctx* c, const bdem_float3* pos, const bdem_Quaterion* ornt,
                 const unsigned int* ptype, const int* pid) {
  int rc = require(c);
  if (rc) return rc;
  if (!pos || !ptype || !pid) return fail(c, BDEM_EINVAL, "null argument");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return upload_state(c, pos, ornt, nullptr, nullptr, ptype, pid, true);
}

int bdem_set_pos_velocity(bdem_ctx* c, const bdem_float3* pos, const bdem_Quaterion* ornt,
                          const bdem_float3* vel, const bdem_float3* avel,
                          const unsigned int* ptype, const int* pid) {
  int rc = require(c);
  if (rc) return rc;
  if (!pos || !vel || !avel || !ptype || !pid) return fail(c, BDEM_EINVAL, "null argument");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return upload_state(c, pos, ornt, vel, avel, ptype, pid, false);
}

int bdem_set_system_state(bdem_ctx* c, const bdem_float3* pos, const bdem_Quaterion* ornt,
                          const bdem_float3* vel, const bdem_float3* avel, const bdem_float3* acc,
                          const unsigned int* ptype, const int* pid) {
  (void)acc;
  return bdem_set_pos_velocity(c, pos, ornt, vel, avel, ptype, pid);
}

int bdem_set_radii(bdem_ctx* c, const float* radius, const float* mass) {
  int rc = require(c);
  if (rc) return rc;
  if (!c->hp.per_particle) return fail(c, BDEM_ESTATE, "per_particle_radius is not enabled");
  if (!radius || !mass) return fail(c, BDEM_EINVAL, "null argument");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const int n = c->b.n;
  std::vector<float4> P(n), V(n);
  HIPCHK(c, hipMemcpy(P.data(), c->b.pos_cur, n * sizeof(float4), hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(V.data(), c->b.vel_cur, n * sizeof(float4), hipMemcpyDeviceToHost));
  for (int i = 0; i < n; ++i) {
    if (!(radius[i] > 0.f) || !(mass[i] > 0.f)) return fail(c, BDEM_EINVAL, "radius and mass must be positive");
    P[i].w = radius[i];
    V[i].w = mass[i];
  }
  HIPCHK(c, hipMemcpy(c->b.pos_cur, P.data(), n * sizeof(float4), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->b.vel_cur, V.data(), n * sizeof(float4), hipMemcpyHostToDevice));
  c->lists_full = true;   /* list membership depends on the radii */
  return BDEM_OK;
}

int bdem_set_tallys(bdem_ctx* c, const float* t) {
  int rc = require(c);
  if (rc) return rc;
  if (!t) return fail(c, BDEM_EINVAL, "null argument");
  for (int i = 0; i < 3; ++i) c->tallys[i] = t[i];
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(c->d_tally, c->tallys, 3 * sizeof(float), hipMemcpyHostToDevice));
  return BDEM_OK;
}

int bdem_get_system_state(bdem_ctx* c, bdem_float3* pos, bdem_Quaterion* ornt, bdem_float3* vel,
                          bdem_float3* avel, bdem_float3* acc, unsigned int* ptype, int* pid) {
  int rc = require(c);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  std::vector<float4> P, V, O, Q;
  if ((rc = download(c, P, V, O, Q)) != BDEM_OK) return rc;
  for (int i = 0; i < c->b.n; ++i) {
    if (pos) {
      pos[i] = bdem_float3{P[i].x, P[i].y, P[i].z};
      const int row = shape_row(c, f_bits(O[i].w));
      if (c->poly && row >= 0) {
        /* Get_SystemState / Get_Particle_PosArray return position - COM for
         * polyhedra (GPU_Memory_Transfer.cuh:461, 504) */
        const bdem_float3 cm = c->shapes[row].COM;
        pos[i] = bdem_float3{P[i].x - cm.x, P[i].y - cm.y, P[i].z - cm.z};
      }
    }
    if (vel) vel[i] = bdem_float3{V[i].x, V[i].y, V[i].z};
    if (avel) avel[i] = bdem_float3{O[i].x, O[i].y, O[i].z};
    if (acc) acc[i] = bdem_float3{0.f, 0.f, 0.f};
    const int packed = f_bits(O[i].w);
    if (ptype) ptype[i] = (unsigned)(packed & 7);
    if (pid) pid[i] = packed >> 3;
    if (ornt && c->orient) ornt[i] = bdem_Quaterion{Q[i].x, Q[i].y, Q[i].z, Q[i].w};
  }
  return BDEM_OK;
}

int bdem_get_positions(bdem_ctx* c, bdem_float3* pos, bdem_Quaterion* ornt, unsigned int* ptype, int* pid) {
  int rc = require(c);
  if (rc) return rc;
  if (c->b.n == 0) { if (pid) pid[0] = -2; return BDEM_OK; }  /* Device/DeviceInterface.cu:2599-2602 */
  return bdem_get_system_state(c, pos, ornt, nullptr, nullptr, nullptr, ptype, pid);
}

int bdem_get_velocity(bdem_ctx* c, bdem_float3* vel, bdem_float3* avel, int flag, unsigned int* ptype, int* pid) {
  return bdem_get_system_state(c, nullptr, nullptr, vel, flag == 1 ? avel : nullptr, nullptr, ptype, pid);
}

int bdem_record_wall_forces(bdem_ctx* c, int on) {
  int rc = require(c);
  if (rc) return rc;
  drop_graphs(c);   /* the step's launch arguments change */
  c->record_wf = on != 0;
  return BDEM_OK;
}

int bdem_get_wall_forces(bdem_ctx* c, bdem_float3* wf, bdem_float3* cp) {
  int rc = require(c);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const int n = c->b.n;
  std::vector<float> F((size_t)n * 3);
  HIPCHK(c, hipMemcpy(F.data(), c->b.wall_force, F.size() * sizeof(float), hipMemcpyDeviceToHost));
  for (int i = 0; i < n; ++i) {
    /* Get_Wall_Forces returns the negated wall force, GPU_Memory_Transfer.cuh:552-566 */
    if (wf) wf[i] = bdem_float3{-1.0f * F[3 * i], -1.0f * F[3 * i + 1], -1.0f * F[3 * i + 2]};
    if (cp) cp[i] = bdem_float3{0.f, 0.f, 0.f};
  }
  return BDEM_OK;
}

int bdem_get_energy(bdem_ctx* c, float* e) {
  int rc = require(c);
  if (rc) return rc;
  if (!e) return fail(c, BDEM_EINVAL, "null argument");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(c->tallys, c->d_tally, 3 * sizeof(float), hipMemcpyDeviceToHost));
  for (int i = 0; i < 3; ++i) e[i] = float(double(c->tallys[i]) * 1E-4);
  return BDEM_OK;
}

int bdem_get_particle_stats(bdem_ctx* c, float min_y, const bdem_float3* box_lo, const bdem_float3* box_hi,
                            bdem_particle_stats* out) {
  int rc = require(c);
  if (rc) return rc;
  if (!out) return fail(c, BDEM_EINVAL, "null argument");
  if ((box_lo == nullptr) != (box_hi == nullptr)) return fail(c, BDEM_EINVAL, "give both box corners or neither");
  HIPCHK(c, hipSetDevice(c->device));
  /* the current state is in the *_cur streams, behind the last step on the stream */
  float3 lo = make_float3(1.0f, 1.0f, 1.0f), hi = make_float3(0.0f, 0.0f, 0.0f);   /* empty box */
  if (box_lo) {
    lo = make_float3(box_lo->x, box_lo->y, box_lo->z);
    hi = make_float3(box_hi->x, box_hi->y, box_hi->z);
  }
  const StatsWords zero{};
  HIPCHK(c, hipMemcpyAsync(c->d_stats, &zero, sizeof(zero), hipMemcpyHostToDevice, c->stream));
  launch_stats(c->b, min_y, lo, hi, !c->poly, c->d_stats, c->stream);
  HIPCHK(c, hipGetLastError());
  StatsWords w{};
  HIPCHK(c, hipMemcpyAsync(&w, c->d_stats, sizeof(w), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  out->count = (int)w.count;
  out->mean_speed = w.count ? w.sum_speed / (double)w.count : 0.0;
  if (w.best) {
    const uint32_t bits = (uint32_t)(w.best >> 32);
    std::memcpy(&out->max_speed, &bits, 4);
    out->max_pid = (int)(uint32_t)(w.best & 0xFFFFFFFFull) - 1;
  } else {
    out->max_speed = 0.0f;
    out->max_pid = -1;
  }
  out->box_count = (int)w.box_count;
  const double bv = box_lo ? (double(hi.x) - lo.x) * (double(hi.y) - lo.y) * (double(hi.z) - lo.z) : 0.0;
  out->box_solid_fraction = c->poly ? -1.0 : (bv > 0.0 ? w.box_volume / bv : 0.0);
  return BDEM_OK;
}

int bdem_set_poly_shapes(bdem_ctx* c, const bdem_poly_shape* shapes, int count) {
  int rc = require(c);
  if (rc) return rc;
  if (!c->poly) return fail(c, BDEM_EINVAL, "not a polyhedra context (particle_type 1)");
  if (!shapes || count < 1 || count > (1 << 28))
    return fail(c, BDEM_EINVAL, "shape count must be in [1, 2^28] with a non-null table");
  if ((rc = check_poly(c, shapes, count, BDEM_SHAPE_MAX_VERTEX, BDEM_SHAPE_MAX_FACE, BDEM_SHAPE_MAX_EDGE)) !=
      BDEM_OK)
    return rc;
  for (int k = 0; k < count; ++k)
    if (!(shapes[k].radius > 0.0f) || !(shapes[k].mass > 0.0f))
      return fail(c, BDEM_EINVAL, "shape radius and mass must be positive (row " + std::to_string(k) + ")");
  HIPCHK(c, hipSetDevice(c->device));
  std::vector<float4> P, V, O, Q;
  if ((rc = download(c, P, V, O, Q)) != BDEM_OK) return rc;
  const int n = c->b.n;
  for (int i = 0; i < n; ++i)
    if ((f_bits(O[i].w) >> 3) >= count)
      return fail(c, BDEM_EINVAL, "a particle's P_ID has no row in the shape table (index " + std::to_string(i) + ")");
  std::vector<bdem_poly_shape> t(shapes, shapes + count);
  for (bdem_poly_shape& s : t) {
    s.vbound = vertex_bound(s.Local_Vertex, s.num_vertex);
    s.plane_slack = plane_slack(s.Local_Vertex, s.num_vertex, s.face, s.num_faces, s.COM);
  }
  const PolyLayout L = poly_layout(t);
  drop_graphs(c);
  bdem_poly_shape* d = nullptr;
  ShapeHdr* dh = nullptr;
  float4* dl = nullptr;
  if ((rc = dalloc(c, &d, t.size())) != BDEM_OK) return rc;
  if ((rc = dalloc(c, &dh, t.size())) != BDEM_OK || (rc = dalloc(c, &dl, t.size() * (size_t)L.stride)) != BDEM_OK) {
    dfree(c, d);
    dfree(c, dh);
    return rc;
  }
  if (L.stride != c->pb.L.stride) {
    float4* nr = nullptr;
    if ((rc = dalloc(c, &nr, (size_t)n * L.stride)) != BDEM_OK) {
      dfree(c, d);
      dfree(c, dh);
      dfree(c, dl);
      return rc;
    }
    dfree(c, c->pb.rot);
    c->pb.rot = nr;
  }
  dfree(c, c->dshape);
  dfree(c, c->pb.hdr);
  dfree(c, c->pb.local);
  c->dshape = d;
  c->pb.hdr = dh;
  c->pb.local = dl;
  c->pb.L = L;
  c->shapes.swap(t);
  c->shape_by_pid = true;
  /* the current state's radius and mass from its rows */
  for (int i = 0; i < n; ++i) {
    const int row = shape_row(c, f_bits(O[i].w));
    P[i].w = c->shapes[row].radius;
    V[i].w = c->shapes[row].mass;
  }
  HIPCHK(c, hipMemcpy(c->b.pos_cur, P.data(), n * sizeof(float4), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->b.vel_cur, V.data(), n * sizeof(float4), hipMemcpyHostToDevice));
  c->keys_valid = false;
  c->flags_valid = false;
  c->lists_full = true;
  return upload_params(c);
}

int bdem_get_num_poly_shapes(bdem_ctx* c, int* count, int* by_pid) {
  int rc = require(c);
  if (rc) return rc;
  if (count) *count = (int)c->shapes.size();
  if (by_pid) *by_pid = c->shape_by_pid ? 1 : 0;
  return BDEM_OK;
}

int bdem_get_poly_vertices(bdem_ctx* c, int max_vertex, bdem_float3* verts, int* num_vertex, int* pid) {
  int rc = require(c);
  if (rc) return rc;
  if (!verts) return fail(c, BDEM_EINVAL, "null argument");
  if (!c->poly) return fail(c, BDEM_EINVAL, "not a polyhedra context (particle_type 1)");
  const PolyLayout L = c->pb.L;
  if (max_vertex < L.pv) return fail(c, BDEM_EINVAL, "max_vertex is below the largest vertex count of the types");
  HIPCHK(c, hipSetDevice(c->device));
  const int n = c->b.n;
  launch_poly_world_shapes(c->dp, c->dshape, c->b, c->pb, c->stream);
  HIPCHK(c, hipGetLastError());
  std::vector<float4> R((size_t)n * L.stride), O((size_t)n);
  HIPCHK(c, hipMemcpyAsync(R.data(), c->pb.rot, R.size() * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(O.data(), c->b.omg_cur, O.size() * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int i = 0; i < n; ++i) {
    const int packed = f_bits(O[i].w);
    const int row = shape_row(c, packed);
    const int nv = row >= 0 ? c->shapes[row].num_vertex : 0;
    if (num_vertex) num_vertex[i] = nv;
    if (pid) pid[i] = packed >> 3;
    for (int v = 0; v < max_vertex; ++v) {
      const float4 r = v < nv ? R[(size_t)i * L.stride + v] : make_float4(0.f, 0.f, 0.f, 0.f);
      verts[(size_t)i * max_vertex + v] = bdem_float3{r.x, r.y, r.z};
    }
  }
  return BDEM_OK;
}

int bdem_get_surface_profile(bdem_ctx* c, int axis, float lo, float hi, int bins, float* top) {
  int rc = require(c);
  if (rc) return rc;
  if (!top) return fail(c, BDEM_EINVAL, "null argument");
  if (axis != 0 && axis != 2) return fail(c, BDEM_EINVAL, "profile axis must be 0 (x) or 2 (z)");
  if (bins < 1 || bins > PROFILE_MAX_BINS) return fail(c, BDEM_EINVAL, "bins must be 1..4096");
  if (!(hi > lo)) return fail(c, BDEM_EINVAL, "profile range must have hi > lo");
  HIPCHK(c, hipSetDevice(c->device));
  const float inv_w = (float)bins / (hi - lo);
  HIPCHK(c, hipMemsetAsync(c->d_profile, 0, sizeof(uint32_t) * (size_t)bins, c->stream));
  launch_profile(c->b, axis, lo, inv_w, bins, c->d_profile, c->stream);
  HIPCHK(c, hipGetLastError());
  std::vector<uint32_t> keys((size_t)bins);
  HIPCHK(c, hipMemcpyAsync(keys.data(), c->d_profile, sizeof(uint32_t) * (size_t)bins, hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int k = 0; k < bins; ++k) {
    if (keys[k] == 0u) {
      top[k] = std::numeric_limits<float>::quiet_NaN();
      continue;
    }
    const uint32_t u = (keys[k] & 0x80000000u) ? (keys[k] & 0x7FFFFFFFu) : ~keys[k];
    std::memcpy(&top[k], &u, 4);
  }
  return BDEM_OK;
}

int bdem_get_surface_profile_noalloc(bdem_ctx* c, int axis, float lo, float hi, int bins, float* top) {
  int rc = require(c);
  if (rc) return rc;
  if (!top) return fail(c, BDEM_EINVAL, "null argument");
  if (axis != 0 && axis != 2) return fail(c, BDEM_EINVAL, "profile axis must be 0 (x) or 2 (z)");
  if (bins < 1 || bins > PROFILE_MAX_BINS) return fail(c, BDEM_EINVAL, "bins must be 1..4096");
  if (!(hi > lo)) return fail(c, BDEM_EINVAL, "profile range must have hi > lo");
  HIPCHK(c, hipSetDevice(c->device));
  uint32_t* scr = nullptr;
  HIPCHK(c, noalloc_scratch(c->device, &scr));
  const float inv_w = (float)bins / (hi - lo);
  launch_profile_noalloc(c->b, axis, lo, inv_w, bins, scr, c->stream);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(top, scr + PROFILE_MAX_BINS, sizeof(float) * (size_t)bins, hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return BDEM_OK;
}

int bdem_surface_angle(const float* top, int bins, float lo, float hi, double* slope, double* intercept,
                       double* angle_deg) {
  if (!top || bins < 2 || !(hi > lo)) return BDEM_EINVAL;
  /* least squares h = a + b u over the non-empty bins, u at the bin centres */
  const double w = ((double)hi - (double)lo) / bins;
  double su = 0.0, sh = 0.0, m = 0.0;
  for (int k = 0; k < bins; ++k)
    if (top[k] == top[k]) { su += lo + (k + 0.5) * w; sh += top[k]; m += 1.0; }
  if (m < 2.0) return BDEM_EINVAL;
  const double ub = su / m, hb = sh / m;
  double suu = 0.0, suh = 0.0;
  for (int k = 0; k < bins; ++k)
    if (top[k] == top[k]) {
      const double du = lo + (k + 0.5) * w - ub;
      suu += du * du;
      suh += du * ((double)top[k] - hb);
    }
  const double b = suh / suu;
  if (slope) *slope = b;
  if (intercept) *intercept = hb - b * ub;
  if (angle_deg) *angle_deg = std::atan(std::fabs(b)) * (180.0 / 3.14159265358979323846);
  return BDEM_OK;
}

int bdem_get_status(bdem_ctx* c, unsigned int* w) {
  int rc = require(c);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  uint32_t s = 0;
  HIPCHK(c, hipMemcpy(&s, c->b.status, sizeof(s), hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemset(c->b.status, 0, sizeof(uint32_t)));
  if (w) *w = s;
  return BDEM_OK;
}

int bdem_get_kill_count(bdem_ctx* c, int* k) {
  int rc = require(c);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(k, c->b.kill_count, sizeof(int), hipMemcpyDeviceToHost));
  return BDEM_OK;
}

int bdem_sort_changed(bdem_ctx* c, unsigned int* changed) {
  int rc = require(c);
  if (rc) return rc;
  if (!changed) return fail(c, BDEM_EINVAL, "null output");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(changed, c->b.meta + META_COUNT, sizeof(uint32_t), hipMemcpyDeviceToHost));
  return BDEM_OK;
}

int bdem_calibration_copy(bdem_ctx* c, unsigned long long bytes, int repeats) {
  int rc = require(c);
  if (rc) return rc;
  if (bytes == 0 || bytes % 16 != 0 || repeats < 1) return fail(c, BDEM_EINVAL, "bytes must be a positive multiple of 16");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  void *src = nullptr, *dst = nullptr;
  if (hipMalloc(&src, bytes) != hipSuccess) return fail(c, BDEM_ENOMEM, "calibration buffer");
  if (hipMalloc(&dst, bytes) != hipSuccess) { (void)hipFree(src); return fail(c, BDEM_ENOMEM, "calibration buffer"); }
  launch_fill_u32(static_cast<uint32_t*>(src), bytes / 4, 0x3F800000u, c->stream);
  for (int r = 0; r < repeats; ++r)
    launch_copy_u4(static_cast<uint4*>(dst), static_cast<const uint4*>(src), bytes / 16, c->stream);
  const hipError_t e = hipStreamSynchronize(c->stream);
  (void)hipFree(src);
  (void)hipFree(dst);
  if (e != hipSuccess) return fail(c, BDEM_EHIP, std::string("calibration copy: ") + hipGetErrorString(e));
  return BDEM_OK;
}

int bdem_num_particles(const bdem_ctx* c) { return (c && c->configured) ? c->b.n : 0; }
long long bdem_num_cells(const bdem_ctx* c) { return (c && c->configured) ? (long long)c->hp.C : 0; }

int bdem_update_dobject_positions(bdem_ctx* c, const bdem_DOBJ* d) {
  int rc = require(c);
  if (rc) return rc;
  if (c->dobj.empty()) return BDEM_OK;
  if (!d) return fail(c, BDEM_EINVAL, "null dynamic-object table");
  if ((rc = check_dobj(c, d, int(c->dobj.size()))) != BDEM_OK) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->dobj_dirty = false;   /* the caller's table replaces the device's turned one */
  c->dobj.assign(d, d + c->dobj.size());
  return upload_params(c);
}

int bdem_update_world_object_positions(bdem_ctx* c, const bdem_WOBJ* w) {
  int rc = require(c);
  if (rc) return rc;
  if (!w) return fail(c, BDEM_EINVAL, "null argument");
  if ((rc = check_wobj(c, w, int(c->wobj.size()))) != BDEM_OK) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (size_t k = 0; k < c->wobj.size(); ++k) c->wobj[k] = w[k];
  return upload_params(c);
}

int bdem_open_hatch(bdem_ctx* c, const bdem_WOBJ* w, int nw) {
  int rc = require(c);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (w) {
    if (nw < 0 || nw > MAX_WOBJ) return fail(c, BDEM_EINVAL, "num_wobj out of range");
    if ((rc = check_wobj(c, w, nw)) != BDEM_OK) return rc;
    c->wobj.assign(w, w + nw);
  }
  if (c->num_wobj_active > 0) c->num_wobj_active -= 1;
  return upload_params(c);
}

int bdem_reset_energy(bdem_ctx* c, int num) {
  int rc = require(c);
  if (rc) return rc;
  if (num <= 3) {
    c->tallys[0] = c->tallys[1] = c->tallys[2] = 0.f;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(c->d_tally, c->tallys, 3 * sizeof(float), hipMemcpyHostToDevice));
  } else {
    c->tallys[3] = c->tallys[4] = 0.f;
  }
  return BDEM_OK;
}

int bdem_drum_rotation(bdem_ctx* c, int on) {
  int rc = require(c);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->cyl_rotation = on != 0;
  return upload_params(c);
}

int bdem_set_lifter_rotation(bdem_ctx* c, double rad_per_step, float ax, float ay, float az) {
  int rc = require(c);
  if (rc) return rc;
  if (!std::isfinite(rad_per_step) || !std::isfinite(ax) || !std::isfinite(ay) || !std::isfinite(az))
    return fail(c, BDEM_EINVAL, "rotation angle and axis point must be finite");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  drop_graphs(c);
  c->lift_on = rad_per_step != 0.0;
  c->lift_c = std::cos(rad_per_step);
  c->lift_s = std::sin(rad_per_step);
  c->lift_about[0] = ax; c->lift_about[1] = ay; c->lift_about[2] = az;
  return BDEM_OK;
}

int bdem_get_dobjects(bdem_ctx* c, bdem_DOBJ* out, int* num) {
  int rc = require(c);
  if (rc) return rc;
  if (!num) return fail(c, BDEM_EINVAL, "null argument");
  HIPCHK(c, hipSetDevice(c->device));
  if ((rc = pull_dobj(c)) != BDEM_OK) return rc;
  *num = (int)c->dobj.size();
  if (out) for (size_t k = 0; k < c->dobj.size(); ++k) out[k] = c->dobj[k];
  return BDEM_OK;
}

int bdem_start_bin(bdem_ctx* c) {
  int rc = require(c);
  if (rc) return rc;
  c->start_bin = true;
  return BDEM_OK;
}

int bdem_add_force(bdem_ctx* c) { return require(c); }
int bdem_sub_force(bdem_ctx* c) { return require(c); }
int bdem_zero_force(bdem_ctx* c) { return require(c); }

int bdem_update_sim(bdem_ctx* c) {
  int rc = require(c);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  hipEvent_t* ev = timing_slot(c);
  if (c->lift_on && !c->dobj.empty()) {
    /* UpdateSim turns the lifters before the step (main.cpp:114-121) */
    launch_turn_lifters(c->dp, c->lift_c, c->lift_s, c->lift_about[0], c->lift_about[1], c->lift_about[2],
                        c->stream);
    c->dobj_dirty = true;
  }
  if ((rc = prepare_sorted(c, ev)) != BDEM_OK) return rc;
  bool do_kill = false;
  if (c->sim.Kill_Particles) {
    const int interval = (c->has_ext && c->ext.kill_check_interval > 0) ? c->ext.kill_check_interval : 1000;
    if (c->stepc + 1 >= interval) { do_kill = true; c->stepc = 0; }
    else c->stepc += 1;
  }
  const SkinArgs sa = skin_args(c);
  if (sa.on) {
    launch_skin(c->dp, c->b, c->b.keys_sorted[c->sorted_par ^ 1], sa.par, c->lists_full, c->stream);
    c->lists_full = false;
  }
  /* the lists boundary only when there was a lists stage (an event between
   * two dependent launches costs a few microseconds) */
  if (ev && sa.on) (void)hipEventRecord(ev[2], c->stream);
  if (ev) c->t_lists[(size_t)c->t_used] = sa.on ? 1 : 0;
  if (c->poly)
    launch_poly_step(c->dp, c->dshape, c->b, c->pb, c->b.keys_sorted[c->sorted_par], c->append_par, do_kill,
                     c->record_wf, !c->dobj.empty(), c->sim.EnergyCalc, c->stream);
  else
    launch_step(c->dp, c->b, c->b.keys_sorted[c->sorted_par], c->hp.law, c->orient, c->hist_par,
                c->append_par, do_kill, c->record_wf, c->sim.EnergyCalc, sa, c->dbg_list, c->dbg_nn, c->stream);
  if (sa.on) c->vpar = sa.par;
  if (do_kill) c->full_next = true;
  if (ev) {
    (void)hipEventRecord(ev[3], c->stream);
    c->t_used += 1;
  }
  HIPCHK(c, hipGetLastError());
  if (need_history(c)) c->hist_par ^= 1;
  c->keys_valid = true;
  c->flags_valid = true;
  return BDEM_OK;
}

int bdem_step(bdem_ctx* c, int nsteps) {
  int rc = require(c);
  if (rc) return rc;
  int s = 0;
  while (s < nsteps) {
    if (nsteps - s >= 2 && graph_pair_ok(c)) {
      HIPCHK(c, hipSetDevice(c->device));
      if ((rc = graph_pair(c)) != BDEM_OK) return rc;
      s += 2;
    } else {
      if ((rc = bdem_update_sim(c)) != BDEM_OK) return rc;
      s += 1;
    }
  }
  return BDEM_OK;
}

int bdem_set_graphs(bdem_ctx* c, int on) {
  int rc = require(c);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->graphs_on = on != 0;
  if (!c->graphs_on) drop_graphs(c);
  return BDEM_OK;
}

int bdem_sync(bdem_ctx* c) {
  int rc = require(c);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return BDEM_OK;
}

int bdem_timing_begin(bdem_ctx* c, int max_steps) {
  int rc = require(c);
  if (rc) return rc;
  if (max_steps < 0 || max_steps > 1000000) return fail(c, BDEM_EINVAL, "max_steps out of range");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (hipEvent_t e : c->tev) (void)hipEventDestroy(e);
  c->tev.assign((size_t)max_steps * (BDEM_NSTAGE + 1), nullptr);
  c->t_lists.assign((size_t)max_steps, 0);
  for (auto& e : c->tev) HIPCHK(c, hipEventCreate(&e));
  c->t_cap = max_steps;
  c->t_used = 0;
  return BDEM_OK;
}

int bdem_timing_end(bdem_ctx* c, double* stage_ms, int* nsteps) {
  int rc = require(c);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  double acc[BDEM_NSTAGE] = {};
  for (int s = 0; s < c->t_used; ++s) {
    hipEvent_t* ev = &c->tev[(size_t)s * (BDEM_NSTAGE + 1)];
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, ev[0], ev[1]));
    acc[0] += ms;
    if (c->t_lists[(size_t)s]) {
      HIPCHK(c, hipEventElapsedTime(&ms, ev[1], ev[2]));
      acc[1] += ms;
      HIPCHK(c, hipEventElapsedTime(&ms, ev[2], ev[3]));
    } else {
      HIPCHK(c, hipEventElapsedTime(&ms, ev[1], ev[3]));
    }
    acc[2] += ms;
  }
  if (stage_ms) for (int k = 0; k < BDEM_NSTAGE; ++k) stage_ms[k] = acc[k];
  if (nsteps) *nsteps = c->t_used;
  for (hipEvent_t e : c->tev) (void)hipEventDestroy(e);
  c->tev.clear();
  c->t_cap = c->t_used = 0;
  return BDEM_OK;
}

int bdem_debug_hash(bdem_ctx* c, unsigned int* keys, int* idx) {
  int rc = require(c);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  launch_hash(c->dp, c->b, c->stream);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const int n = c->b.n;
  HIPCHK(c, hipMemcpy(keys, c->b.keys_new, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
  for (int i = 0; i < n; ++i) idx[i] = i;
  c->keys_valid = true;
  c->flags_valid = false;
  return BDEM_OK;
}

int bdem_debug_sorted(bdem_ctx* c, unsigned int* keys, int* idx) {
  int rc = require(c);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  if ((rc = prepare_sorted(c)) != BDEM_OK) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const int n = c->b.n;
  HIPCHK(c, hipMemcpy(keys, c->b.keys_sorted[c->sorted_par], n * sizeof(uint32_t), hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(idx, c->b.idx_sorted, n * sizeof(int), hipMemcpyDeviceToHost));
  c->keys_valid = false;
  return BDEM_OK;
}

int bdem_debug_cells(bdem_ctx* c, unsigned int* cs, unsigned int* ce) {
  int rc = require(c);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  if ((rc = prepare_sorted(c)) != BDEM_OK) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  std::vector<uint2> cells(c->hp.C);
  HIPCHK(c, hipMemcpy(cells.data(), c->b.cells, cells.size() * sizeof(uint2), hipMemcpyDeviceToHost));
  for (size_t k = 0; k < cells.size(); ++k) { cs[k] = cells[k].x; ce[k] = cells[k].y; }
  c->keys_valid = false;
  return BDEM_OK;
}

int bdem_debug_contacts(bdem_ctx* c, unsigned int* list32, unsigned int* nn) {
  int rc = require(c);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  if ((rc = prepare_sorted(c)) != BDEM_OK) return rc;
  const int n = c->b.n;
  uint32_t *dl = nullptr, *dn = nullptr;
  HIPCHK(c, hipMalloc((void**)&dl, (size_t)n * LIST_STRIDE * sizeof(uint32_t)));
  HIPCHK(c, hipMalloc((void**)&dn, (size_t)n * sizeof(uint32_t)));
  launch_fill_u32(dl, (size_t)n * LIST_STRIDE, 0u, c->stream);
  launch_contact_list(c->dp, c->b, dl, dn, c->stream);
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess) e = hipMemcpy(list32, dl, (size_t)n * LIST_STRIDE * sizeof(uint32_t), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(nn, dn, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToHost);
  (void)hipFree(dl);
  (void)hipFree(dn);
  c->keys_valid = false;
  if (e != hipSuccess) return fail(c, BDEM_EHIP, std::string("debug_contacts: ") + hipGetErrorString(e));
  return BDEM_OK;
}

int bdem_debug_forces(bdem_ctx* c, float* fpp, float* tpp, float* fw, float* tw, int* sorted_pid) {
  int rc = require(c);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  if ((rc = prepare_sorted(c)) != BDEM_OK) return rc;
  const int n = c->b.n;
  float* d = nullptr;
  HIPCHK(c, hipMalloc((void**)&d, (size_t)n * 12 * sizeof(float)));
  if (c->poly)
    launch_poly_forces(c->dp, c->dshape, c->b, c->pb, d, d + 3 * (size_t)n, d + 6 * (size_t)n, d + 9 * (size_t)n,
                       c->stream);
  else
    launch_forces(c->dp, c->b, c->hp.law, c->hist_par, d, d + 3 * (size_t)n, d + 6 * (size_t)n,
                  d + 9 * (size_t)n, c->stream);
  hipError_t e = hipStreamSynchronize(c->stream);
  std::vector<float> h((size_t)n * 12);
  if (e == hipSuccess) e = hipMemcpy(h.data(), d, h.size() * sizeof(float), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(c, BDEM_EHIP, std::string("debug_forces: ") + hipGetErrorString(e));
  std::memcpy(fpp, h.data(), (size_t)n * 3 * sizeof(float));
  std::memcpy(tpp, h.data() + 3 * (size_t)n, (size_t)n * 3 * sizeof(float));
  std::memcpy(fw, h.data() + 6 * (size_t)n, (size_t)n * 3 * sizeof(float));
  std::memcpy(tw, h.data() + 9 * (size_t)n, (size_t)n * 3 * sizeof(float));
  if (sorted_pid) {
    std::vector<float4> O(n);
    HIPCHK(c, hipMemcpy(O.data(), c->b.omg_srt, n * sizeof(float4), hipMemcpyDeviceToHost));
    for (int i = 0; i < n; ++i) sorted_pid[i] = f_bits(O[i].w) >> 3;
  }
  c->keys_valid = false;
  return BDEM_OK;
}

int bdem_set_skin(bdem_ctx* c, float skin) {
  int rc = require(c);
  if (rc) return rc;
  if (!(skin == skin)) return fail(c, BDEM_EINVAL, "skin must be a number");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  drop_graphs(c);
  c->skin_req = skin;
  return apply_skin(c);
}

float bdem_get_skin(const bdem_ctx* c) { return (c && c->configured) ? c->b.skin : 0.0f; }

int bdem_debug_poly_queues(bdem_ctx* c, unsigned int* counts) {
  int rc = require(c);
  if (rc) return rc;
  if (!counts) return fail(c, BDEM_EINVAL, "null output");
  if (!c->poly) return fail(c, BDEM_EINVAL, "not a polyhedra context (particle_type 1)");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(counts, c->pb.qcount, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost));
  return BDEM_OK;
}

int bdem_debug_step_contacts(bdem_ctx* c, unsigned int* list32, unsigned int* nn) {
  int rc = require(c);
  if (rc) return rc;
  if (!list32 || !nn) return fail(c, BDEM_EINVAL, "null output");
  if (c->poly) return fail(c, BDEM_ENOTSUP, "the polyhedra step keeps its list in its own buffers (bdem_debug_contacts)");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const int n = c->b.n;
  uint32_t *dl = nullptr, *dn = nullptr;
  HIPCHK(c, hipMalloc((void**)&dl, (size_t)n * LIST_STRIDE * sizeof(uint32_t)));
  if (hipMalloc((void**)&dn, (size_t)n * sizeof(uint32_t)) != hipSuccess) {
    (void)hipFree(dl);
    return fail(c, BDEM_ENOMEM, "debug buffer");
  }
  launch_fill_u32(dl, (size_t)n * LIST_STRIDE, 0u, c->stream);
  launch_fill_u32(dn, (size_t)n, 0u, c->stream);
  c->dbg_list = dl;
  c->dbg_nn = dn;
  rc = bdem_update_sim(c);
  c->dbg_list = nullptr;
  c->dbg_nn = nullptr;
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess) e = hipMemcpy(list32, dl, (size_t)n * LIST_STRIDE * sizeof(uint32_t), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(nn, dn, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToHost);
  (void)hipFree(dl);
  (void)hipFree(dn);
  if (rc) return rc;
  if (e != hipSuccess) return fail(c, BDEM_EHIP, std::string("debug_step_contacts: ") + hipGetErrorString(e));
  return BDEM_OK;
}

}  // extern "C"
