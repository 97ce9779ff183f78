/*
 * bdem_poly.hip - polyhedra (particle_type 1) step for gfx950.
 *
 * Kernel_ParticleInteraction_Polyhedra_NonSymmetry (DEM_Kernels_Compute.cuh:
 * 736-992) rotates both shapes for every listed pair and tests one pair per
 * particle thread in a serial loop.  Here a step is five launches:
 *
 *   k_poly_rotate   one thread per (particle, vertex / face-normal /
 *                   face-centroid slot): world-frame shapes into `rot`,
 *                   written coalesced (RotatePolyhedra, Utilities.cuh:420-479)
 *   k_poly_list     bounding-sphere contact list in the reference order (the
 *                   sphere broad phase), plus a queue of (particle, slot)
 *                   work items appended one atomic per wave
 *   k_poly_pairs    one thread per queued pair, grid-stride over the queue:
 *                   the separating-plane test of the two facing faces; the
 *                   pairs it does not separate go to a second queue
 *   k_poly_narrow   one wave per pair of that queue, both bodies staged in
 *                   LDS: the full separating-plane test (a lane per face),
 *                   vertex-face both ways (a lane per vertex), edge-edge
 *                   (lanes over the edge pairs a distance bound keeps),
 *                   depth clamps, pair law.  Results go to the pair's own
 *                   slot, so the queue orders do not matter
 *   k_poly_finish   one thread per particle: gravity, then the slots in list
 *                   order (the reference's summation order), walls, the
 *                   tensor integrator and quaternion, discharge, next key,
 *                   changed flag and changed-list append
 *
 * Every float expression follows the restatement's operation order, so the
 * results equal the oracle's bit for bit away from sin/cos in the quaternion.
 */
#include <atomic>

#include "bdem_kernels.hpp"

namespace bdem {

namespace {

__device__ __forceinline__ f3 ld3(const float4* __restrict__ p) {
  const float4 v = *p;
  return mk3(v.x, v.y, v.z);
}

/* q v q* with (0, v), the reference's make_quaterionD / make_vectorD. */
__device__ __forceinline__ f3 qrot(QuatF Q, QuatF QT, f3 v) {
  const QuatF p{0.0f, v.x, v.y, v.z};
  const QuatF r = qmul(qmul(Q, p), QT);
  return mk3(r.x, r.y, r.z);
}

/* The shape-table row of a particle: its type (the POBJ table of the deck)
 * or, with per-particle shapes, its P_ID.  A dead particle (P_ID -1) takes
 * row 0; nothing reads its shape. */
__device__ __forceinline__ int shape_index(const DevParams& P, float w) {
  if (!P.shape_by_pid) return packed_type(w);
  const int pid = packed_pid(w);
  return pid > 0 ? pid : 0;
}

/* A rotated shape in `rot`: vertices, then face normals, then face centroids.
 * The wave-per-pair kernel adds the face polygons from its LDS stage (vertex
 * orders packed four to a word, areas, vertex counts). */
struct Shape {
  const float4* V;
  const float4* Fn;
  const float4* Fc;
  int nv, nf;
  const uint2* vo = nullptr;
  const float* area = nullptr;
  const int* fnv = nullptr;
};
__device__ __forceinline__ Shape shape_of(const float4* __restrict__ rot, const PolyLayout& L, int i,
                                          const bdem_poly_shape& po) {
  const float4* base = rot + (size_t)i * L.stride;
  return Shape{base, base + L.pv, base + L.pv + L.pf, po.num_vertex, po.num_faces};
}

__device__ __forceinline__ Shape shape_of(const float4* __restrict__ rot, const PolyLayout& L, int i,
                                          const ShapeHdr& h) {
  const float4* base = rot + (size_t)i * L.stride;
  return Shape{base, base + L.pv, base + L.pv + L.pf, h.nv, h.nf};
}

/* Face f; f = -1 (no face of B had A's centre in front) is a zero plane. */
__device__ __forceinline__ f3 fnorm(const Shape& B, int f) { return f >= 0 ? ld3(B.Fn + f) : mk3(0.f, 0.f, 0.f); }
__device__ __forceinline__ f3 fcent(const Shape& B, int f) { return f >= 0 ? ld3(B.Fc + f) : mk3(0.f, 0.f, 0.f); }

enum { C_UNDEF = 0, C_VF = 1, C_EF = 2, C_FF = 3, C_EE = 4, C_SP = 5 };
struct PColl { bool hit; int cls; float dis; f3 n; f3 cp; };

/* Vertex q of a face's packed vertex order. */
__device__ __forceinline__ int vo_at(uint2 vo, int q) {
  const uint32_t w = q < 4 ? vo.x : vo.y;
  return (int)((w >> (8 * (q & 3))) & 0xFFu);
}

/* is_PointOnSurface over a polyhedron's rotated vertices (Utilities.cuh:20-53):
 * face of nv vertices in order vo, area farea. */
__device__ __forceinline__ bool point_on_face(int nv, uint2 vo, float farea, const float4* __restrict__ V, f3 point,
                                              float tol) {
  float area = 0.000000f;
  for (int q = 0; q < nv - 1; ++q) {
    const f3 V1 = point - ld3(V + vo_at(vo, q));
    const f3 V2 = point - ld3(V + vo_at(vo, q + 1));
    area = area + 0.500000f * length(cross(V1, V2));
  }
  const f3 V1 = point - ld3(V + vo_at(vo, nv - 1));
  const f3 V2 = point - ld3(V + vo_at(vo, 0));
  area = area + 0.500000f * length(cross(V1, V2));
  return (fabsf(area - farea) / farea) * 100.0000f <= tol;
}

/* B's centre farther beyond a face plane of A than B's vertex bound vb (any
 * turned vertex lies within vb of the centre): then every vertex of B is in
 * front of that plane, and the vertex test of the face must find them so.
 * The margin (10 um plus 4e-6 of the coordinates' size) exceeds the float
 * error of the turned vertices and of the plane distances by a wide factor,
 * and vb carries 1e-4 for the unit-length tolerance of the quaternion. */
__device__ __forceinline__ bool beyond_by_sphere(float dist, f3 c, f3 PosB, float vb) {
  const float size = fabsf(PosB.x) + fabsf(PosB.y) + fabsf(PosB.z) + fabsf(c.x) + fabsf(c.y) + fabsf(c.z);
  return dist > vb + 0.001f + 4.0e-6f * size;
}

/* Wave reductions by DPP moves inside each row of 16 lanes (swap
 * neighbours, swap pairs, rotate the row by 4, then by 8: every lane then
 * holds its row's result) and readlane across the four rows: a few cycles
 * a step, where a shuffle goes through the LDS crossbar.  Every lane of the
 * wave must be active; the results are wave-uniform. */
constexpr int DPP_XOR1 = 0xB1, DPP_XOR2 = 0x4E, DPP_ROR4 = 0x124, DPP_ROR8 = 0x128;
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ float dpp_f32(float v) { return __uint_as_float(dpp_u32<CTRL>(__float_as_uint(v))); }
__device__ __forceinline__ float lane_f32(float v, int l) {
  return __uint_as_float((uint32_t)__builtin_amdgcn_readlane((int)__float_as_uint(v), l));
}

__device__ __forceinline__ float wave_min(float v) {
  v = fminf(v, dpp_f32<DPP_XOR1>(v));
  v = fminf(v, dpp_f32<DPP_XOR2>(v));
  v = fminf(v, dpp_f32<DPP_ROR4>(v));
  v = fminf(v, dpp_f32<DPP_ROR8>(v));
  return fminf(fminf(lane_f32(v, 0), lane_f32(v, 16)), fminf(lane_f32(v, 32), lane_f32(v, 48)));
}

__device__ __forceinline__ int wave_min_i32(int v) {
  v = min(v, (int)dpp_u32<DPP_XOR1>((uint32_t)v));
  v = min(v, (int)dpp_u32<DPP_XOR2>((uint32_t)v));
  v = min(v, (int)dpp_u32<DPP_ROR4>((uint32_t)v));
  v = min(v, (int)dpp_u32<DPP_ROR8>((uint32_t)v));
  return min(min(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
             min(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}

/* The minimum of a 64-bit key over each half of the wave (lanes 0-31,
 * 32-63), returned to the lanes of that half. */
template <int CTRL>
__device__ __forceinline__ uint64_t dpp_min_u64(uint64_t k) {
  const uint64_t o = ((uint64_t)dpp_u32<CTRL>((uint32_t)(k >> 32)) << 32) | dpp_u32<CTRL>((uint32_t)k);
  return o < k ? o : k;
}
__device__ __forceinline__ uint64_t half_min_u64(uint64_t k) {
  k = dpp_min_u64<DPP_XOR1>(k);
  k = dpp_min_u64<DPP_XOR2>(k);
  k = dpp_min_u64<DPP_ROR4>(k);
  k = dpp_min_u64<DPP_ROR8>(k);
  auto row = [&](int l) {
    return ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(k >> 32), l) << 32) |
           (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)k, l);
  };
  const uint64_t r0 = row(0), r1 = row(16), r2 = row(32), r3 = row(48);
  const uint64_t h0 = r0 < r1 ? r0 : r1, h1 = r2 < r3 ? r2 : r3;
  return (threadIdx.x & 32) ? h1 : h0;
}

/* The face and the vertex of each body of a pair turned most towards the
 * other: largest n . (other centre - own centre), resp.
 * (v - own centre) . (other centre - own centre), the lowest index on ties;
 * lanes 0-31 work on A, 32-63 on B.  They only choose which face or vertex
 * a test tries first; no result depends on them. */
struct PairDirs { int fA, fB, vA, vB; };
__device__ __forceinline__ uint64_t argmax_key(bool valid, float v, int idx) {
  if (!valid) return ~0ull;
  uint32_t u = __float_as_uint(v);
  u = (u & 0x80000000u) ? ~u : (u | 0x80000000u);   /* ordered as the floats */
  return ((uint64_t)(~u) << 32) | (uint32_t)idx;    /* smallest key: largest value, then lowest index */
}
__device__ __forceinline__ PairDirs pair_dirs(const Shape& A, const Shape& B, f3 PosA, f3 PosB) {
  const int lane = threadIdx.x & 63;
  const bool hb = lane >= 32;
  const int k = lane & 31;
  const Shape& own = hb ? B : A;
  const f3 Pown = hb ? PosB : PosA;
  const f3 d = hb ? PosA - PosB : PosB - PosA;
  const bool fk = k < own.nf, vk = k < own.nv;
  const uint64_t kf = half_min_u64(argmax_key(fk, dot(ld3(own.Fn + (fk ? k : 0)), d), k));
  const uint64_t kv = half_min_u64(argmax_key(vk, dot(ld3(own.V + (vk ? k : 0)) - Pown, d), k));
  auto pick = [](uint64_t key, int l) {
    const int i = __builtin_amdgcn_readlane((int)(uint32_t)key, l);
    return (i >= 0 && i < 32) ? i : 0;
  };
  return PairDirs{pick(kf, 0), pick(kf, 32), pick(kv, 0), pick(kv, 32)};
}

struct InPoly { int f0, f1, f2; float d0, d1, d2; };

/* The inside part of Collision_Detection_Vertex_InPoly_3Faces
 * (Collision_Detection_Polyhedra.cuh:287-375) for every vertex of P against
 * Q: vertex v is inside when every face of Q has it behind its plane,
 * vd = n . (v - c) < -1e-5 (the serial scan stops at the first face that
 * does not).  Each vertex is first tried against face fq of Q (the face
 * turned towards P, which most of P's vertices are in front of), a lane per
 * vertex; the rest take every face, a lane per (vertex, face): lanes 0-31
 * face f = lane with one vertex, lanes 32-63 the same face with the next.
 * Bit v of the result is set for an inside vertex. */
__device__ __forceinline__ uint32_t vertices_inside(const Shape& P, const Shape& Q, int fq) {
  const int lane = threadIdx.x & 63;
  uint32_t rest;
  {
    const int v = lane < P.nv ? lane : 0;
    const float vd = dot(ld3(Q.Fn + fq), ld3(P.V + v) - ld3(Q.Fc + fq));
    rest = (uint32_t)__ballot(lane < P.nv && vd < -0.0000100f);
  }
  const int f = lane & 31, h = lane >> 5;
  const bool fv = f < Q.nf;
  const int fr = fv ? f : 0;
  const f3 n = ld3(Q.Fn + fr), c = ld3(Q.Fc + fr);
  uint32_t inside = 0u;
  while (rest) {
    const int v0 = __ffs(rest) - 1;
    rest &= rest - 1u;
    const int v1 = rest ? __ffs(rest) - 1 : -1;
    if (v1 >= 0) rest &= rest - 1u;
    const float vd = dot(n, ld3(P.V + (h && v1 >= 0 ? v1 : v0)) - c);
    const uint64_t out = __ballot(fv && !(vd < -0.0000100f));
    if ((uint32_t)out == 0u) inside |= 1u << v0;
    if (v1 >= 0 && (uint32_t)(out >> 32) == 0u) inside |= 1u << v1;
  }
  return inside;
}

/* The selection part of the same function for an inside vertex: the three
 * faces of Q with PosP in front (n . (PosP - c) > 0) that have the smallest
 * |vd|, the earlier face first on ties, as the serial scan inserts them
 * (strictly below the third distance so far, which starts at 10); -1 and 10
 * where fewer qualify.  Lanes 0-31 work for vertex v, lanes 32-63 for vertex
 * w, one face each; every lane returns its half's vertex's result. */
__device__ __forceinline__ InPoly closest_faces(f3 PosP, const Shape& P, int v, int w, const Shape& Q) {
  const int lane = threadIdx.x & 63;
  const int f = lane & 31;
  uint64_t key = ~0ull;   /* (bits of |vd|, face): the order of the insertion */
  if (f < Q.nf) {
    const f3 n = ld3(Q.Fn + f), c = ld3(Q.Fc + f);
    const float a = fabsf(dot(n, ld3(P.V + (lane < 32 ? v : w)) - c));
    if (dot(n, PosP - c) > 0.0f && a < 10.0f) key = ((uint64_t)__float_as_uint(a) << 32) | (uint32_t)f;
  }
  InPoly r{-1, -1, -1, 10.0f, 10.0f, 10.0f};
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const uint64_t m = half_min_u64(key);
    if (m != ~0ull) {   /* the same in every lane of the half */
      const int fm = (int)(uint32_t)m;
      const float dm = __uint_as_float((uint32_t)(m >> 32));
      if (q == 0) { r.f0 = fm; r.d0 = dm; }
      else if (q == 1) { r.f1 = fm; r.d1 = dm; }
      else { r.f2 = fm; r.d2 = dm; }
      if (f == fm) key = ~0ull;
    }
  }
  return r;
}

/* The ray from PosA through `through` meets face f of B inside its polygon. */
__device__ __forceinline__ bool ray_hits_face(f3 PosA, f3 through, const Shape& B, int f, float tol) {
  if (f < 0) return false;
  f3 Ray = through - PosA;
  Ray = Ray / length(Ray);
  const f3 n = ld3(B.Fn + f);
  const float dir = dot(Ray, n);
  if (!(fabsf(dir) > 0.0f)) return false;
  const float d = dot((ld3(B.Fc + f) - PosA), n) / dir;
  return point_on_face(B.fnv[f], B.vo[f], B.area[f], B.V, PosA + d * Ray, tol);
}

__device__ __forceinline__ int face_at(const InPoly& s, int q) { return q == 0 ? s.f0 : (q == 1 ? s.f1 : s.f2); }

/* Collision_Detection_Polyhedara_Polyhedra_VertexFaces
 * (Collision_Detection_Polyhedra.cuh:1154-1630): vertices of A inside B; one
 * -> vertex-face, two -> edge-face, more -> face-face (the first eight
 * inside vertices: the reference's lists hold eight), with a ray test
 * choosing among the three closest faces.  A whole wave works on one pair:
 * lane v < A.nv tests vertex v
 * (Collision_Detection_Vertex_InPoly_3Faces); the inside vertices are then
 * taken in ascending order, as the serial loop appends them, so the first
 * two vertices, their closest faces and the ordered sum of the first eight
 * are the serial ones, and the selection runs on wave-uniform values. */
/* The selection among the inside vertices of A (bit v of mask: vertex v):
 * the closest faces of the first (and the second's second-closest face) are
 * found by closest_faces. */
__device__ __forceinline__ PColl vertex_faces_select(f3 PosA, const Shape& A, const Shape& B, uint32_t mask) {
  PColl res{false, C_UNDEF, 0.0f, mk3(0.f, 0.f, 0.f), mk3(0.f, 0.f, 0.f)};
  const int np_all = __popc(mask);
  if (np_all == 0) return res;
  const int v0 = __ffs(mask) - 1;
  const uint32_t rest = mask & (mask - 1u);
  const int v1 = rest ? __ffs(rest) - 1 : v0;
  const InPoly t = closest_faces(PosA, A, v0, v1, B);
  InPoly s0;   /* t is uniform over each half: lane 0 holds v0's faces, lane 32 v1's */
  s0.f0 = __builtin_amdgcn_readlane(t.f0, 0);
  s0.f1 = __builtin_amdgcn_readlane(t.f1, 0);
  s0.f2 = __builtin_amdgcn_readlane(t.f2, 0);
  s0.d0 = lane_f32(t.d0, 0);
  s0.d1 = lane_f32(t.d1, 0);
  s0.d2 = lane_f32(t.d2, 0);
  const int s1f1 = __builtin_amdgcn_readlane(t.f1, 32);
  const int np = np_all < 8 ? np_all : 8;
  res.hit = true;
  if (np == 1) {
    res.cls = C_VF;
    res.dis = s0.d0;
    res.n = fnorm(B, s0.f0);
    res.cp = ld3(A.V + v0);
    if (s0.f1 == -1) return res;
    if (ray_hits_face(PosA, res.cp, B, s0.f0, 0.10f)) return res;
    if (s0.f1 > -1 && ray_hits_face(PosA, res.cp, B, s0.f1, 0.10f)) {
      res.n = fnorm(B, s0.f1);
      res.dis = s0.d1;
      return res;
    }
    if (s0.f2 > -1 && ray_hits_face(PosA, res.cp, B, s0.f2, 0.10f)) {
      res.n = fnorm(B, s0.f2);
      res.dis = s0.d2;
      return res;
    }
    return res;
  }
  if (np == 2) {
    res.cls = C_EF;
    res.cp = (ld3(A.V + v0) + ld3(A.V + v1)) * 0.50f;
    res.dis = fabsf(dot(fnorm(B, s0.f0), (res.cp - fcent(B, s0.f0))));
    res.n = fnorm(B, s0.f0);
    if (s0.f1 == -1 && s1f1 == -1) return res;
    for (int q = 0; q < 3; ++q) {
      const int f = face_at(s0, q);
      if (q > 0 && !(f > -1)) continue;
      if (ray_hits_face(PosA, res.cp, B, f, 0.10f)) {
        res.dis = fabsf(dot(fnorm(B, f), (res.cp - fcent(B, f))));
        res.n = fnorm(B, f);
        return res;
      }
    }
    return res;
  }
  res.cls = C_FF;
  f3 sum = mk3(0.0f, 0.0f, 0.0f);
  uint32_t m = mask;
  for (int k = 0; k < np; ++k) {
    const int v = __ffs(m) - 1;
    m &= m - 1u;
    sum = sum + ld3(A.V + v);
  }
  res.cp = sum / (float)np;
  for (int q = 0; q < 3; ++q) {
    const int f = face_at(s0, q);
    if (q > 0 && !(f > -1)) continue;
    if (ray_hits_face(PosA, res.cp, B, f, q == 1 ? 0.010f : 0.10f)) {
      res.dis = fabsf(dot(fnorm(B, f), (res.cp - fcent(B, f))));
      res.n = fnorm(B, f);
      return res;
    }
  }
  return res;   /* face logic fall-through: depth 0, zero normal */
}

/* Vertices of A in B, and if none, vertices of B in A (normal reversed). */
__device__ __forceinline__ PColl vertex_faces_pair(f3 PosA, const Shape& A, f3 PosB, const Shape& B,
                                                   const PairDirs& D) {
  PColl c = vertex_faces_select(PosA, A, B, vertices_inside(A, B, D.fB));
  if (!c.hit) {
    c = vertex_faces_select(PosB, B, A, vertices_inside(B, A, D.fA));
    c.n = c.n * -1.0f;   /* B into A: normal reversed */
  }
  return c;
}

/* One edge pair of Collision_Detection_Polyhedra_Polyhedra_Edges
 * (Collision_Detection_Polyhedra.cuh:537-707): true, with the distance of
 * the closest points, when the closest points of the two lines fall inside
 * both edges (1 % slack) and apart. */
__device__ __forceinline__ bool edge_pair_dis(const Shape& A, uint2 ea, const Shape& B, uint2 eb, float& dis) {
  const f3 PA = ld3(A.V + ea.x);
  const f3 DA = ld3(A.V + ea.y) - PA;
  const float DEAD = dot(DA, DA);
  const f3 PB = ld3(B.V + eb.x);
  const f3 DB = ld3(B.V + eb.y) - PB;
  const f3 dir = cross(DA, DB);
  const float mag = dot(dir, dir);
  if (!(mag > 0.00000f)) return false;
  const f3 DV = PA - PB;
  const float DEBD = dot(DB, DB);
  const float DEABD = dot(DA, DB);
  const float DEADAB = dot(DA, DV);
  const float DEBDAB = dot(DB, DV);
  const float detJinv = 1.00000f / (DEABD * DEABD - DEAD * DEBD);
  const float dA = detJinv * (DEBD * DEADAB - DEABD * DEBDAB);
  if (!(dA >= -0.010f && dA <= 1.0100f)) return false;
  const float dB = detJinv * (DEABD * DEADAB - DEAD * DEBDAB);
  if (!(dB >= -0.0100f && dB <= 1.010f)) return false;
  const f3 pA = PA + DA * dA;
  const f3 pB = PB + DB * dB;
  const f3 D = (pA - pB);
  dis = dot(D, D);
  if (!(dis > 0.00000f)) return false;
  dis = __builtin_sqrtf(dis);
  return true;
}

/* A lower bound of |pA - pB| over every pA of the edge [p0, p1] extended by
 * 1 % at both ends (the range the search accepts) and every pB the search
 * can pick on body Q: centre c, vertex bound vb, face plane (n, fc) with
 * slack sl (no vertex of Q lies more than sl in front of it).  A point of
 * Q's extended edges is within 1.02 vb of c and at most sl + 0.02 vb in
 * front of the plane; n may be longer than 1 (a non-unit orientation
 * quaternion turns it so) and is divided out. */
__device__ __forceinline__ float edge_bound(f3 p0, f3 p1, f3 c, float vb, f3 n, f3 fc, float sl) {
  const f3 d = p1 - p0;
  const f3 e0 = p0 - 0.01f * d, e1 = p1 + 0.01f * d;
  const float ln = length(n);
  const float hp = ln > 0.0f ? fminf(dot(n, e0 - fc), dot(n, e1 - fc)) / ln - sl - 0.02f * vb : -3.0e38f;
  const f3 de = e1 - e0;
  const float dd = dot(de, de);
  float u = dd > 0.0f ? dot(c - e0, de) / dd : 0.0f;
  u = fminf(fmaxf(u, 0.0f), 1.0f);
  const float hs = length(c - (e0 + de * u)) - 1.02f * vb;
  return fmaxf(hp, hs);
}

/* Collision_Detection_Polyhedra_Polyhedra_Edges (Collision_Detection_Polyhedra.cuh:537-707),
 * with a whole wave on one pair.  The serial loop walks t = i * E_B + j and
 * keeps the first strictly smaller distance below 5: the smallest distance
 * and among equal distances the smallest t.  Here every edge of A gets a
 * lower bound la_i of the distance any of its pairs can have (edge_bound
 * against B's bounding sphere and the face of B turned most towards A), and
 * every edge of B one against A, lb_j.  Pass 1 evaluates the pairs whose
 * edges both come within w1 of the other body (the closest edges); its best
 * distance U then bounds the answer, and pass 2 evaluates every remaining
 * pair with la_i and lb_j below U plus a rounding margin.  A pair left out
 * has a distance above U, so it can be neither the minimum nor tie with it;
 * the lanes keep the smallest (distance, t) of what they evaluate and a wave
 * minimum selects the serial loop's pair, whose points every lane then
 * recomputes with the same arithmetic.  The result is wave-uniform. */
__device__ __forceinline__ PColl edges_wave(const Shape& A, const uint2* EA, int ea, const Shape& B, const uint2* EB,
                                            int eb, f3 RelPos, f3 PosA, f3 PosB, float vbA, float vbB, float slA,
                                            float slB, unsigned char* ia, unsigned char* ib, int fA, int fB) {
  const int lane = threadIdx.x & 63;
  float la = 3.0e38f, lb = 3.0e38f;
  if (lane < ea)
    la = edge_bound(ld3(A.V + EA[lane].x), ld3(A.V + EA[lane].y), PosB, vbB, ld3(B.Fn + fB), ld3(B.Fc + fB), slB);
  if (lane < eb)
    lb = edge_bound(ld3(B.V + EB[lane].x), ld3(B.V + EB[lane].y), PosA, vbA, ld3(A.Fn + fA), ld3(A.Fc + fA), slA);
  const float size = fabsf(PosA.x) + fabsf(PosA.y) + fabsf(PosA.z) + fabsf(PosB.x) + fabsf(PosB.y) + fabsf(PosB.z);
  const float margin = 0.001f + 4.0e-6f * size;
  float best = 5.0f;
  int bt = 0x7FFFFFFF;
  const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  /* the pairs of SA x SB, from compacted index lists; lane u takes pairs
   * u, u + 64, ... in row-major order (64 = qa nb + rb) */
  auto sweep = [&](uint64_t SA, uint64_t SB) {
    __builtin_amdgcn_wave_barrier();
    if ((SA >> lane) & 1ull) ia[__popcll(SA & lt)] = (unsigned char)lane;
    if ((SB >> lane) & 1ull) ib[__popcll(SB & lt)] = (unsigned char)lane;
    __builtin_amdgcn_wave_barrier();
    const int na = __popcll(SA), nb = __popcll(SB);
    const int total = na * nb;
    if (total == 0) return;
    const int qa = 64 / nb, rb = 64 - qa * nb;
    int a = lane / nb, b = lane - a * nb;
    for (int u = lane; u < total; u += 64) {
      const int i = ia[a], j = ib[b];
      a += qa;
      b += rb;
      if (b >= nb) { b -= nb; ++a; }
      float dis;
      if (edge_pair_dis(A, EA[i], B, EB[j], dis) && dis < 5.0f) {
        const int t = i * eb + j;
        if (dis < best || (dis == best && t < bt)) { best = dis; bt = t; }
      }
    }
  };
  const float t0 = fmaxf(wave_min(la), wave_min(lb));
  const float w1 = fmaxf(t0, 0.0f) + 0.05f * fminf(vbA, vbB) + margin;
  const uint64_t SA1 = __ballot(lane < ea && la <= w1), SB1 = __ballot(lane < eb && lb <= w1);
  sweep(SA1, SB1);
  /* pass 2: the pairs of SA2 x SB2 that pass 1 left, as (SA2 \ SA1) x SB2
   * and (SA2 n SA1) x (SB2 \ SB1) */
  const float w2 = wave_min(best) + margin;
  const uint64_t SA2 = __ballot(lane < ea && la <= w2), SB2 = __ballot(lane < eb && lb <= w2);
  if ((SA2 & ~SA1) && SB2) sweep(SA2 & ~SA1, SB2);
  if ((SA2 & SA1) && (SB2 & ~SB1)) sweep(SA2 & SA1, SB2 & ~SB1);
  /* wave minimum of (distance, t) */
  const float m = wave_min(best);
  const int tm = wave_min_i32((best == m) ? bt : 0x7FFFFFFF);
  PColl res{false, C_SP, 0.0f, mk3(0.f, 0.f, 0.f), mk3(0.0f, 0.0f, 0.0f)};
  if (!(m < 5.0f) || tm == 0x7FFFFFFF) return res;
  /* the selected pair, recomputed by every lane */
  const int si = tm / eb, sj = tm - (tm / eb) * eb;
  const f3 PA = ld3(A.V + EA[si].x);
  const f3 DA = ld3(A.V + EA[si].y) - PA;
  const float DEAD = dot(DA, DA);
  const f3 PB = ld3(B.V + EB[sj].x);
  const f3 DB = ld3(B.V + EB[sj].y) - PB;
  const f3 DV = PA - PB;
  const float DEBD = dot(DB, DB);
  const float DEABD = dot(DA, DB);
  const float DEADAB = dot(DA, DV);
  const float DEBDAB = dot(DB, DV);
  const float detJinv = 1.00000f / (DEABD * DEABD - DEAD * DEBD);
  const float dA = detJinv * (DEBD * DEADAB - DEABD * DEBDAB);
  const float dB = detJinv * (DEABD * DEADAB - DEAD * DEBDAB);
  const f3 selA = PA + DA * dA;
  const f3 selB = PB + DB * dB;
  res.cp = selA;
  res.n = cross(DA, DB);
  res.n = res.n * (1.0f / length(res.n));
  if (length(PosB - selB) < length(PosB - selA)) {
    /* isPointinPoly (Utilities.cuh:63-98), a lane per face of B */
    const int f = lane < B.nf ? lane : 0;
    if (__ballot(lane < B.nf && !(dot(ld3(B.Fn + f), selA - ld3(B.Fc + f)) <= 0.000100f))) return res;
  }
  res.hit = true;
  res.cls = C_EE;
  const float dn = dot(RelPos, res.n);
  if (fabsf(dn) > 0.0f) {
    res.n = (dot(RelPos, res.n) / fabsf(dot(RelPos, res.n))) * res.n;
    res.dis = fabsf(dot(res.n, (selB - selA))) + 0.000010f;
  } else {
    res.n = RelPos / length(RelPos);
    res.cp = PosB;
    res.dis = fabsf(dot(res.n, (selB - selA))) + 0.000010f;
  }
  return res;
}

/* Face f of A separates B: B's centre and every vertex of B strictly in
 * front (the per-face test of is_SP_Poly_Poly, Utilities.cuh:105-156; vb is
 * B's vertex bound: the sphere shortcut changes which face is found first,
 * never whether one is). */
__device__ __forceinline__ bool face_separates(const Shape& A, int f, f3 PosB, const Shape& B, float vb) {
  const f3 n = ld3(A.Fn + f), c = ld3(A.Fc + f);
  const float dist = dot(n, PosB - c);
  if (!(dist > 0.0f)) return false;
  if (beyond_by_sphere(dist, c, PosB, vb)) return true;
  /* eight vertices per round, their loads in flight together (indices past
   * the count re-read the last vertex and are not tested) */
  const int last = B.nv - 1;
  for (int v = 0; v < B.nv; v += 8) {
    f3 p[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) p[u] = ld3(B.V + min(v + u, last));
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (v + u < B.nv && dot(n, p[u] - c) <= 0.00000f) return false;
  }
  return true;
}

/* The face of A whose normal points most towards B's centre. */
__device__ __forceinline__ int facing_face(const Shape& A, f3 dirAB) {
  int best = 0;
  float bv = -3.0e38f;
  const int last = A.nf - 1;
  for (int f = 0; f < A.nf; f += 8) {   /* eight normals per round, as above */
    float d[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) d[u] = dot(ld3(A.Fn + min(f + u, last)), dirAB);
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (f + u < A.nf && d[u] > bv) { bv = d[u]; best = f + u; }
  }
  return best;
}

/* The contact of A with listed partner B (DEM_Kernels_Compute.cuh:820-905)
 * starts with the separating-plane test both ways (is_SP_Poly_Poly): true
 * when some face of one body separates them.  Which face is found does not
 * matter, only whether one exists, so each body's face turned most towards
 * the other is tried first (quick_separated, one thread per pair): it
 * separates most separated pairs.  The rest go to the wave-per-pair kernel,
 * where separated_wave tests every face at once. */
/* A separating face is named by a code: f for face f of A, 0x100 | f for
 * face f of B; HINT_NONE for none. */
constexpr uint32_t HINT_NONE = 0xFFFFFFFFu;

__device__ __forceinline__ bool quick_separated(const Shape& A, const Shape& B, f3 PosA, f3 PosB, float vbA,
                                                float vbB, uint32_t& code) {
  const f3 dAB = PosB - PosA;
  const int fa = facing_face(A, dAB);
  if (face_separates(A, fa, PosB, B, vbB)) { code = (uint32_t)fa; return true; }
  const int fb = facing_face(B, -1.0f * dAB);
  if (face_separates(B, fb, PosA, A, vbA)) { code = 0x100u | (uint32_t)fb; return true; }
  return false;
}

/* The separating-face hint of a listed pair: the partner's P_ID and the code
 * of the face that separated the two at the previous step.  Whether some
 * face separates a pair does not depend on which face shows it, so a face
 * that separates them now proves it, and a stale or foreign hint costs one
 * test and changes nothing. */
__device__ __forceinline__ bool hint_separates(uint2 h, uint32_t pid_b, const Shape& A, const Shape& B, f3 PosA,
                                               f3 PosB, float vbA, float vbB) {
  if (h.x != pid_b || h.y == HINT_NONE) return false;
  const int f = (int)(h.y & 0xFFu);
  if (h.y & 0x100u) return f < B.nf && face_separates(B, f, PosA, A, vbA);
  return f < A.nf && face_separates(A, f, PosB, B, vbB);
}

/* is_SP_Poly_Poly both ways (Utilities.cuh:105-156): some face of A has
 * B's centre and every vertex of B strictly in front, or some face of B so
 * for A.  A face is a candidate when the centre and the other body's vertex
 * turned towards it (D) are in front, a lane per face (0-31 A's, 32-63
 * B's); the candidates then take all vertices, two faces per round and a
 * lane per vertex, in ascending order, so the first separating face found
 * is the lowest-numbered one.  Returns its code, or HINT_NONE. */
__device__ __forceinline__ uint32_t separated_wave(const Shape& A, const Shape& B, f3 PosA, f3 PosB,
                                                   const PairDirs& D) {
  const int lane = threadIdx.x & 63;
  uint64_t cand;
  {
    const bool hb = lane >= 32;
    const int f = lane & 31;
    const Shape& own = hb ? B : A;
    const Shape& other = hb ? A : B;
    const bool fv = f < own.nf;
    const int fr = fv ? f : 0;
    const f3 n = ld3(own.Fn + fr), c = ld3(own.Fc + fr);
    cand = __ballot(fv && dot(n, (hb ? PosA : PosB) - c) > 0.0f &&
                    !(dot(n, ld3(other.V + (hb ? D.vA : D.vB)) - c) <= 0.00000f));
  }
  const int h = lane >> 5, v = lane & 31;
  while (cand) {
    const int c0 = __ffsll((unsigned long long)cand) - 1;
    cand &= cand - 1ull;
    const int c1 = cand ? __ffsll((unsigned long long)cand) - 1 : -1;
    if (c1 >= 0) cand &= cand - 1ull;
    const int cf = (h && c1 >= 0) ? c1 : c0;   /* this half's face: code cf (B's faces from 32) */
    const bool fb = cf >= 32;
    const Shape& own = fb ? B : A;
    const Shape& other = fb ? A : B;
    const int f = cf & 31;
    const f3 n = ld3(own.Fn + f), c = ld3(own.Fc + f);
    const bool vv = v < other.nv;
    const bool behind = vv && dot(n, ld3(other.V + (vv ? v : 0)) - c) <= 0.00000f;
    const uint64_t m = __ballot(behind);
    if ((uint32_t)m == 0u) return c0 < 32 ? (uint32_t)c0 : (0x100u | (uint32_t)(c0 - 32));
    if (c1 >= 0 && (uint32_t)(m >> 32) == 0u) return c1 < 32 ? (uint32_t)c1 : (0x100u | (uint32_t)(c1 - 32));
  }
  return HINT_NONE;
}

/* The depth clamps (DEM_Kernels_Compute.cuh:918-950). */
__device__ __forceinline__ void depth_clamps(PColl& res, float rA, f3 PosA, f3 PosB, f3 RVel) {
  const f3 RelPos = (PosA - PosB) / length(PosA - PosB);
  if (res.dis > 0.1500f * rA) {   /* contact error: fixed small overlap along the centres */
    res.n = RelPos / length(RelPos);
    res.dis = 0.025f * rA;
    res.cp = PosA;
    res.cls |= 8;
  }
  if (res.dis > 0.0500f * rA) {   /* large overlap: limited by speed, first branch wins as written */
    res.n = RelPos / length(RelPos);
    const float sp_rel = length(RVel);
    if (sp_rel > 50.0f) res.dis = 0.010f * rA;
    else if (sp_rel > 150.0f) res.dis = 0.015f * rA;
    else if (sp_rel > 250.0f) res.dis = 0.020f * rA;
    else res.dis = 0.0030f * rA;
    res.cp = PosA;
    res.cls |= 16;
  }
}

/* A queued pair's operands. */
struct PairOps {
  int a, b, ta, tb, sa, sb;   /* types (materials) and shape-table rows */
  f3 PosA, PosB, va, vb, wa, wb;
  float ma, mb;
};
__device__ __forceinline__ PairOps pair_ops(const DevParams& P, uint32_t slot, const uint32_t* __restrict__ list32,
                                            const float4* __restrict__ pos, const float4* __restrict__ vel,
                                            const float4* __restrict__ omg) {
  PairOps o;
  o.a = (int)(slot / LIST_STRIDE);
  o.b = (int)list32[slot];
  const float4 oa = omg[o.a], ob = omg[o.b];
  o.ta = packed_type(oa.w);
  o.tb = packed_type(ob.w);
  o.sa = shape_index(P, oa.w);
  o.sb = shape_index(P, ob.w);
  const float4 va4 = vel[o.a], vb4 = vel[o.b];
  o.PosA = xyz(pos[o.a]);
  o.PosB = xyz(pos[o.b]);
  o.va = xyz(va4); o.vb = xyz(vb4);
  o.ma = va4.w; o.mb = vb4.w;
  o.wa = xyz(oa); o.wb = xyz(ob);
  return o;
}

/* The operands of a queued pair whose partner is known. */
__device__ __forceinline__ PairOps pair_ops_ab(const DevParams& P, uint32_t slot, uint32_t b,
                                               const float4* __restrict__ pos, const float4* __restrict__ vel,
                                               const float4* __restrict__ omg) {
  PairOps o;
  o.a = (int)(slot / LIST_STRIDE);
  o.b = (int)b;
  const float4 oa = omg[o.a], ob = omg[o.b];
  o.ta = packed_type(oa.w);
  o.tb = packed_type(ob.w);
  o.sa = shape_index(P, oa.w);
  o.sb = shape_index(P, ob.w);
  const float4 va4 = vel[o.a], vb4 = vel[o.b];
  o.PosA = xyz(pos[o.a]);
  o.PosB = xyz(pos[o.b]);
  o.va = xyz(va4); o.vb = xyz(vb4);
  o.ma = va4.w; o.mb = vb4.w;
  o.wa = xyz(oa); o.wb = xyz(ob);
  return o;
}

/* Clamps, then the reference pair law into the pair's slot. */
__device__ __forceinline__ void pair_store(const DevParams& P, const bdem_poly_shape& pa, const PairOps& o, PColl c,
                                           uint32_t slot, float4* __restrict__ pf, float4* __restrict__ pt,
                                           float* tally = nullptr) {
  float4 F4 = make_float4(0.f, 0.f, 0.f, 0.f), T4 = F4;
  if (c.hit) {
    depth_clamps(c, pa.radius, o.PosA, o.PosB, o.va - o.vb);
    f3 F, T;
    pair_resp(P, o.va, o.ma, o.wa, o.ta, o.vb, o.mb, o.wb, o.tb, c.n, c.dis, c.cp - o.PosA, c.cp - o.PosB, F, T,
              tally);
    F4 = make_float4(F.x, F.y, F.z, 1.0f);
    T4 = make_float4(T.x, T.y, T.z, 0.0f);
  }
  pf[slot] = F4;
  pt[slot] = T4;
}

/* ---------------------------------------------------------------- walls */
/* Collision_Detection_Vertcies_Surface (Collision_Detection_Polyhedra.cuh:85-185) */
__device__ __forceinline__ PColl vertices_surface(const Shape& A, const bdem_Surface& S, const bdem_float3* svtx) {
  PColl res{false, C_UNDEF, 0.0f, mk3(0.f, 0.f, 0.f), mk3(0.0f, 0.0f, 0.0f)};
  float dis_max = 0.0f;
  int cnt = 0;
  const f3 n = cvt(S.normal), c = cvt(S.centroid);
  for (int v = 0; v < A.nv; ++v) {
    const f3 pv = ld3(A.V + v);
    const float vd = dot(n, (pv - c));
    if (vd < 0.000000f) {
      const f3 hit = pv + fabsf(vd) * n;
      if (point_on_surface(S, svtx, hit)) {
        ++cnt;
        res.cp = res.cp + hit;
        if (fabsf(vd) > dis_max) { res.hit = true; dis_max = fabsf(vd); }
        if (cnt > 2) {
          res.cls = C_FF; res.n = n; res.dis = dis_max;
          res.cp = mk3(0.0f, 0.0f, 0.0f);
          return res;
        }
      }
    }
  }
  if (!res.hit) { res.cls = C_SP; return res; }
  res.cls = cnt == 1 ? C_VF : C_EF;
  res.n = n;
  res.dis = dis_max;
  res.cp = res.cp / (float)cnt;
  return res;
}

/* Collision_Detection_Polyhedra_Surface (Collision_Detection_Polyhedra.cuh:489-529) */
__device__ __forceinline__ PColl poly_surface(f3 p, float r, const Shape& A, const bdem_Surface& S,
                                              const bdem_float3* svtx) {
  PColl res{false, C_UNDEF, 0.0f, mk3(0.f, 0.f, 0.f), mk3(0.f, 0.f, 0.f)};
  const f3 n = cvt(S.normal), c = cvt(S.centroid);
  if (dot(n, (p - c)) > 0.000000f) {
    if (dot(n, (p - c)) <= r) {
      res = vertices_surface(A, S, svtx);
      if (res.cls == C_FF) res.cp = p;   /* no moment for face-face */
    }
  }
  return res;
}

/* Collision_Detection_Vertcies_Plane (Collision_Detection_Polyhedra.cuh:6-84) */
__device__ __forceinline__ PColl vertices_plane(const Shape& A, f3 n, f3 c) {
  PColl res{false, C_SP, 0.0f, mk3(0.f, 0.f, 0.f), mk3(0.0f, 0.0f, 0.0f)};
  float dis_max = 0.0f;
  int cnt = 0;
  for (int v = 0; v < A.nv; ++v) {
    const f3 pv = ld3(A.V + v);
    const float vd = dot(n, (pv - c));
    if (vd < 0.000000f) {
      ++cnt;
      res.cp = res.cp + pv;
      if (fabsf(vd) > dis_max) dis_max = fabsf(vd);
    }
  }
  if (cnt == 0) return res;
  res.hit = true;
  res.cls = cnt == 2 ? C_EF : C_FF;
  res.cp = res.cp / (float)cnt;
  res.n = n;
  res.dis = dis_max;
  return res;
}

/* Collision_Detection_Polyhedra_Cylinder (Collision_Detection_Polyhedra.cuh:2010-2132);
 * the running depth is averaged in double, as written. */
__device__ __forceinline__ PColl poly_cylinder(f3 p, float r, const Shape& A, const bdem_Macro_Cylinder& cyl) {
  PColl res{false, C_UNDEF, 0.0f, mk3(0.f, 0.f, 0.f), mk3(0.f, 0.f, 0.f)};
  const f3 AX = cvt(cyl.AXIS), cb = cvt(cyl.center_bot_cap);
  const f3 Pvec = p * AX - cb;
  const float Pvec_dis = length(Pvec) + r;
  if (!(Pvec_dis > cyl.radius)) return res;
  float dis = 0.0f;
  int cnt = 0;
  f3 s_point = mk3(0.0f, 0.0f, 0.0f);
  for (int v = 0; v < A.nv; ++v) {
    const f3 pv = ld3(A.V + v);
    const f3 vp = pv * AX - cb;
    const float rc = length(vp);
    if (rc > cyl.radius) {
      const float vd = rc - cyl.radius;
      ++cnt;
      s_point = s_point + pv;
      dis = dis + vd;
      if (cnt > 2) {
        dis = (float)((double)dis * (1.0 / (double)(float)cnt));
        res.hit = true;
        res.cls = C_FF;
        res.n = -1.0f * Pvec;
        res.n = res.n * (1.0f / length(res.n));
        res.dis = fminf(0.00100f, fabsf(dis));
        res.cp = p;
        return res;
      }
    }
  }
  if (cnt == 0) return res;
  dis = (float)((double)dis * (1.0 / (double)(float)cnt));
  res.n = -1.0f * Pvec;
  res.n = res.n * (1.0f / length(res.n));
  res.hit = true;
  if (cnt == 2) {
    res.cls = C_EF;
    res.cp = s_point / (float)cnt;
    res.dis = fabsf(dis);
    return res;
  }
  res.cp = s_point;
  res.dis = dis;
  res.cls = C_VF;
  return res;
}

/* WorldObject_InteractionParticle, polyhedra branches (DEM_Kernels_Compute.cuh:
 * 254-610).  Infinite planes have no polyhedra method: the previous result
 * stands, as written. */
/* No vertex of a body with centre p and vertex bound vb (every turned
 * vertex within vb of p) lies behind the plane (n, c) when p is farther in
 * front of it than |n| vb plus a margin (10 um and 4e-6 of the coordinates'
 * size) beyond the float error of the vertex heights: the vertex test of
 * the plane would find nothing. */
__device__ __forceinline__ bool clear_of_plane(f3 n, f3 c, f3 p, float vb) {
  const float size = fabsf(p.x) + fabsf(p.y) + fabsf(p.z) + fabsf(c.x) + fabsf(c.y) + fabsf(c.z);
  return dot(n, p - c) > length(n) * vb + 0.001f + 4.0e-6f * size;
}

__device__ __forceinline__ void wall_forces_poly(const DevParams& P, int t, f3 p, float r, f3 v, f3 w, float m,
                                                 const Shape& A, float vb, f3& F, f3& T, float* tally) {
  F = mk3(0.f, 0.f, 0.f);
  T = mk3(0.f, 0.f, 0.f);
  const bool moving_planes = P.rotating_wobj && P.cyl_rotation;
  PColl res{false, C_UNDEF, 0.0f, mk3(0.f, 0.f, 0.f), mk3(0.f, 0.f, 0.f)};
  for (int k = 0; k < P.num_wobj; ++k) {
    const bdem_WOBJ& W = P.wobj[k];
    if (W.surface_type == BDEM_SURFACE_PLANE) {
      for (int s = 0; s < W.num_surfaces; ++s) {
        const bdem_Surface& S = W.surfaces[s];
        if (S.area != -1.0f) res = poly_surface(p, r, A, S, W.vertex);
        if (!res.hit) continue;
        const Coll c{true, res.dis, res.cp, res.n};
        f3 ft = mk3(0.f, 0.f, 0.f), tq = ft;
        if (moving_planes) ft = resp_moving(P, t, p, v, w, m, c, drum_velocity(P, res.cp), tq, tally);
        else resp_static(P, t, p, v, w, m, c, ft, tq, tally);
        F = F + ft;
        if (P.rotation && res.cls != C_FF) T = T + tq;
      }
    } else if (W.surface_type == BDEM_SURFACE_CYLINDER) {
      const bdem_Macro_Cylinder& cyl = W.cyl_geo;
      res = poly_cylinder(p, r, A, cyl);
      if (res.hit) {
        const Coll c{true, res.dis, res.cp, res.n};
        f3 ft = mk3(0.f, 0.f, 0.f), tq = ft;
        if (P.cyl_rotation) ft = resp_moving(P, t, p, v, w, m, c, drum_velocity(P, res.cp), tq, tally);
        else resp_static(P, t, p, v, w, m, c, ft, tq, tally);
        F = ft;
        if (P.rotation && res.cls != C_FF) T = T + tq;
      }
      if (cyl.has_top_cap && !clear_of_plane(cvt(cyl.normal_top_cap), cvt(cyl.center_top_cap), p, vb)) {
        res = vertices_plane(A, cvt(cyl.normal_top_cap), cvt(cyl.center_top_cap));
        if (res.hit) F = resp_endplate(P, t, v, Coll{true, res.dis, res.cp, res.n});
      }
      if (cyl.has_bot_cap && !clear_of_plane(cvt(cyl.normal_bot_cap), cvt(cyl.center_bot_cap), p, vb)) {
        res = vertices_plane(A, cvt(cyl.normal_bot_cap), cvt(cyl.center_bot_cap));
        if (res.hit) F = resp_endplate(P, t, v, Coll{true, res.dis, res.cp, res.n});
      }
    }
  }
}

/* ---- polyhedra against lifters (VolumeObject_InteractionParticle,
 * DEM_Kernels_Compute.cuh:100-224), one thread per particle: lifter contacts
 * are few, so the serial form of the restatement is kept. */

/* is_SP_VolumeObject_Poly (Utilities.cuh:272-309) */
__device__ __forceinline__ bool sp_dobj_poly(const bdem_DOBJ& D, f3 PosP, const Shape& A) {
  for (int f = 0; f < D.num_faces; ++f) {
    const f3 n = cvt(D.faces[f].normal), c = cvt(D.faces[f].centroid);
    if (dot(n, PosP - c) > 0.0f) {
      bool contact = false;
      for (int v = 0; v < A.nv; ++v)
        if (dot(n, ld3(A.V + v) - c) <= 0.00000f) { contact = true; break; }
      if (!contact) return true;
    }
  }
  return false;
}

/* is_SP_Poly_VolumeObject (Utilities.cuh:212-247) */
__device__ __forceinline__ bool sp_poly_dobj(const Shape& A, f3 PosD, const bdem_DOBJ& D) {
  for (int f = 0; f < A.nf; ++f) {
    const f3 n = ld3(A.Fn + f), c = ld3(A.Fc + f);
    if (dot(n, PosD - c) > 0.0f) {
      bool contact = false;
      for (int v = 0; v < D.num_vertex; ++v)
        if (dot(n, cvt(D.vertex[v]) - c) <= 0.00000f) { contact = true; break; }
      if (!contact) return true;
    }
  }
  return false;
}

/* Collision_Detection_Vertex_InDObject_3Faces (Collision_Detection_Polyhedra.cuh:383-476) */
__device__ __forceinline__ bool vertex_in_dobj(f3 PosP, f3 v, const bdem_DOBJ& D, InPoly& r) {
  r = InPoly{-1, -1, -1, 10.0f, 10.0f, 10.0f};
  for (int f = 0; f < D.num_faces; ++f) {
    const f3 n = cvt(D.faces[f].normal), c = cvt(D.faces[f].centroid);
    const float vd = dot(n, v - c);
    if (!(vd <= 0.000010f)) return false;
    if (dot(n, PosP - c) > 0.0f) {
      const float a = fabsf(vd);
      if (a < r.d0) {
        if (r.f0 > -1) {
          if (r.f1 > -1) { r.d2 = r.d1; r.f2 = r.f1; }
          r.d1 = r.d0; r.f1 = r.f0;
        }
        r.d0 = a; r.f0 = f;
      } else if (a < r.d1) {
        if (r.f1 > -1) { r.d2 = r.d1; r.f2 = r.f1; }
        r.d1 = a; r.f1 = f;
      } else if (a < r.d2) {
        r.d2 = a; r.f2 = f;
      }
    }
  }
  return true;
}

__device__ __forceinline__ f3 dnorm(const bdem_DOBJ& D, int f) {
  return f >= 0 ? cvt(D.faces[f].normal) : mk3(0.f, 0.f, 0.f);
}
__device__ __forceinline__ f3 dcent(const bdem_DOBJ& D, int f) {
  return f >= 0 ? cvt(D.faces[f].centroid) : mk3(0.f, 0.f, 0.f);
}
__device__ __forceinline__ float dist_at(const InPoly& s, int q) { return q == 0 ? s.d0 : (q == 1 ? s.d1 : s.d2); }

/* is_PointOnSurface with an explicit tolerance over a lifter's vertices */
__device__ __forceinline__ bool point_on_dface(const bdem_Surface& f, const bdem_float3* V, f3 point, float tol) {
  float area = 0.000000f;
  const int nv = f.num_vertex;
  for (int q = 0; q < nv - 1; ++q) {
    const f3 V1 = point - cvt(V[f.vertex_Order[q]]);
    const f3 V2 = point - cvt(V[f.vertex_Order[q + 1]]);
    area = area + 0.500000f * length(cross(V1, V2));
  }
  const f3 V1 = point - cvt(V[f.vertex_Order[nv - 1]]);
  const f3 V2 = point - cvt(V[f.vertex_Order[0]]);
  area = area + 0.500000f * length(cross(V1, V2));
  return (fabsf(area - f.area) / f.area) * 100.0000f <= tol;
}

__device__ __forceinline__ bool ray_hits_dface(f3 PosP, f3 through, const bdem_DOBJ& D, int f, float tol) {
  if (f < 0) return false;
  f3 Ray = through - PosP;
  Ray = Ray / length(Ray);
  const float dir = dot(Ray, dnorm(D, f));
  if (!(fabsf(dir) > 0.0f)) return false;
  const float d = dot((dcent(D, f) - PosP), dnorm(D, f)) / dir;
  return point_on_dface(D.faces[f], D.vertex, PosP + d * Ray, tol);
}

/* Collision_Detection_Polyhedara_VolumeObject_VertexFace (Collision_Detection_Polyhedra.cuh:1638-1990) */
__device__ __forceinline__ PColl vertex_faces_dobj(f3 PosP, const Shape& A, const bdem_DOBJ& D) {
  PColl res{false, C_UNDEF, 0.0f, mk3(0.f, 0.f, 0.f), mk3(0.f, 0.f, 0.f)};
  int np = 0, valid[8];
  InPoly sel[8];
  for (int v = 0; v < A.nv; ++v) {
    InPoly t;
    if (vertex_in_dobj(PosP, ld3(A.V + v), D, t)) {
      if (np < 8) { sel[np] = t; valid[np] = v; }
      ++np;
    }
  }
  if (np == 0) return res;
  np = min(np, 8);
  res.hit = true;
  const InPoly s0 = sel[0];
  if (np == 1) {
    res.cls = C_VF;
    res.dis = s0.d0;
    res.n = dnorm(D, s0.f0);
    res.cp = ld3(A.V + valid[0]);
    if (s0.f1 == -1) return res;
    if (ray_hits_dface(PosP, res.cp, D, s0.f0, 0.1f)) return res;
    for (int q = 1; q < 3; ++q) {
      const int f = face_at(s0, q);
      if (f > -1 && ray_hits_dface(PosP, res.cp, D, f, 0.10f)) {
        res.n = dnorm(D, f);
        res.dis = dist_at(s0, q);
        return res;
      }
    }
    return res;
  }
  if (np == 2) {
    res.cls = C_EF;
    res.cp = (ld3(A.V + valid[0]) + ld3(A.V + valid[1])) * 0.50f;
    res.n = dnorm(D, s0.f0);
    res.dis = fabsf(dot(dnorm(D, s0.f0), (res.cp - dcent(D, s0.f0))));
    if (s0.f1 == -1 && sel[1].f1 == -1) return res;
    for (int q = 0; q < 3; ++q) {
      const int f = face_at(s0, q);
      if (q > 0 && !(f > -1)) continue;
      if (ray_hits_dface(PosP, res.cp, D, f, 0.10f)) {
        res.dis = fabsf(dot(dnorm(D, f), (res.cp - dcent(D, f))));
        res.n = dnorm(D, f);
        return res;
      }
    }
    return res;
  }
  res.cls = C_FF;
  f3 cp = mk3(0.0f, 0.0f, 0.0f);
  for (int q = 0; q < np; ++q) cp = cp + ld3(A.V + valid[q]);
  res.cp = cp / (float)np;
  int common[3] = {0, 0, 0};
  float cdis[3] = {10.0f, 10.0f, 10.0f};
  for (int q = 0; q < np; ++q)
    for (int j = 0; j < 3; ++j) {
      const int f = face_at(sel[q], j);
      if (f > -1 && f < 3) {
        common[f]++;
        const float dj = dist_at(sel[q], j);
        if (dj < cdis[f]) cdis[f] = dj;
      }
    }
  for (int f = 0; f < D.num_faces && f < 3; ++f)
    if (common[f] == np) {
      res.dis = cdis[f];
      res.n = dnorm(D, f);
      return res;
    }
  return res;
}

/* Collision_Detection_DObject_Polyhedra_Edges (Collision_Detection_Polyhedra.cuh:716-890);
 * no crossing within 0.1 cm of the other point (the reference then reads an
 * unset normal) is no contact. */
__device__ __forceinline__ PColl edges_dobj(const bdem_DOBJ& D, const Shape& B, const bdem_poly_shape& pb, f3 PosB,
                                            f3 RelPos) {
  PColl res{false, C_SP, 0.0f, mk3(0.f, 0.f, 0.f), mk3(0.0f, 0.0f, 0.0f)};
  float min_edgedis = 0.10f;
  f3 selA = mk3(0.f, 0.f, 0.f), selB = selA, nsel = selA, cpsum = selA;
  int count = 0;
  bool found = false, have_min = false;
  for (int i = 0; i < D.num_edges; ++i) {
    const f3 PA = cvt(D.vertex[D.edge[i].point_vindex[0]]);
    const f3 DA = cvt(D.vertex[D.edge[i].point_vindex[1]]) - PA;
    const float DEAD = dot(DA, DA);
    for (int j = 0; j < pb.num_edges; ++j) {
      const f3 PB = ld3(B.V + pb.edge[j].point_vindex[0]);
      const f3 DB = ld3(B.V + pb.edge[j].point_vindex[1]) - PB;
      const f3 dir = cross(DA / length(DA), DB / length(DB));
      const float mag = dot(dir, dir);
      if (!(mag > 0.00000f)) continue;
      f3 DV = PA - PB;
      const float DEBD = dot(DB, DB);
      const float DEABD = dot(DA, DB);
      const float DEADAB = dot(DA, DV);
      const float DEBDAB = dot(DB, DV);
      const float detJinv = 1.00000f / (DEABD * DEABD - DEAD * DEBD);
      const float dA = detJinv * (DEBD * DEADAB - DEABD * DEBDAB);
      if (!(dA >= -0.010f && dA <= 1.0100f)) continue;
      const float dB = detJinv * (DEABD * DEADAB - DEAD * DEBDAB);
      if (!(dB >= -0.0100f && dB <= 1.010f)) continue;
      const f3 pA = PA + DA * dA;
      const f3 pB = PB + DB * dB;
      DV = (pA - pB);
      float dis = dot(DV, DV);
      if (!(dis > 0.00000f)) continue;
      dis = __builtin_sqrtf(dis);
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
  const f3 n = nsel * (1.0f / length(nsel));
  res.hit = true;
  res.cls = C_EE;
  const float dn = dot(RelPos, n);
  if (fabsf(dn) > 0.0f) {
    res.n = -1.0f * ((dot(RelPos, n) / fabsf(dot(RelPos, n))) * n);
    res.dis = fabsf(dot(res.n, (selB - selA)));
    res.cp = cpsum / (float)count;
  } else {
    res.n = -1.0f * RelPos / length(RelPos);
    res.cp = PosB;
    res.dis = fabsf(dot(res.n, (selB - selA))) + 0.000010f;
  }
  return res;
}

/* Lifter force and torque on a polyhedron: the first lifter in contact acts. */
__device__ __forceinline__ void lifter_forces_poly(const DevParams& P, int t, f3 p, f3 v, f3 w, float m,
                                                   const Shape& A, const bdem_poly_shape& pa, f3& F, f3& T,
                                                   float* tally) {
  F = mk3(0.f, 0.f, 0.f);
  T = F;
  const float r = pa.radius;
  const bool check = P.dobj[0].is_translating || !P.mill_sim || P.rotating_wobj ||
                     length(mk3(p.x, p.y, 0.0f) - cvt(P.mill_cent)) > (P.dobj_bound_r - 2.0f * r);
  if (!check) return;
  for (int d = 0; d < P.num_dobj; ++d) {
    const bdem_DOBJ& D = P.dobj[d];
    const f3 BV = (p - cvt(D.COM)) * cvt(D.Axis);
    if (!(length(BV) < (D.boundR + r))) continue;
    PColl res{false, C_SP, 0.0f, mk3(0.f, 0.f, 0.f), mk3(0.f, 0.f, 0.f)};
    bool sp = sp_dobj_poly(D, p, A);
    if (!sp) sp = sp_poly_dobj(A, cvt(D.COM), D);
    if (!sp) {
      f3 PosD = cvt(D.COM);
      PosD.z = p.z;
      res = vertex_faces_dobj(p, A, D);
      if (!res.hit) {
        const f3 RelPos = (PosD - p) / length(PosD - p);
        res = edges_dobj(D, A, pa, p, RelPos);
      }
    }
    if (!res.hit) continue;
    f3 Vl = mk3(0.f, 0.f, 0.f);
    if (D.is_attached_DWorld) Vl = drum_velocity(P, res.cp);
    else if (D.is_translating) Vl = cvt(D.velocity);
    const DColl c{true, 0, res.dis, res.n, res.cp};
    f3 Ftr, Tq;
    resp_dobject(P, t, p, v, w, m, c, Vl, Ftr, Tq, tally);
    F = F + Ftr;
    if (P.rotation) T = T + Tq;
    break;
  }
}

}  // namespace

/* ------------------------------------------------------------------ kernels */
__global__ __launch_bounds__(256) void k_poly_rotate(const DevParams* __restrict__ Pp,
                                                     const float4* __restrict__ local, int n,
                                                     PolyLayout L, const float4* __restrict__ pos,
                                                     const float4* __restrict__ omg,
                                                     const float4* __restrict__ quat,
                                                     float4* __restrict__ rot, uint32_t* __restrict__ qcount) {
  const DevParams& P = *Pp;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t == 0) { qcount[0] = 0u; qcount[1] = 0u; qcount[2] = 0u; }   /* this step's queues start empty */
  const size_t i = t / (size_t)L.stride;
  const int e = (int)(t - i * (size_t)L.stride);
  if (i >= (size_t)n) return;
  /* the shape row's slot e in the same layout, local frame (w = 1: used) */
  const float4 l4 = local[(size_t)shape_index(P, omg[i].w) * L.stride + e];
  const f3 l = xyz(l4);
  const f3 c = xyz(pos[i]);
  const float4 q4 = quat[i];
  const QuatF Q{q4.x, q4.y, q4.z, q4.w};
  const QuatF QT{Q.w, -Q.x, -Q.y, -Q.z};
  f3 out = mk3(0.f, 0.f, 0.f);
  if (l4.w != 0.0f) {
    out = P.rotation ? qrot(Q, QT, l) : l;
    if (e < L.pv || e >= L.pv + L.pf) out = out + c;   /* vertices and centroids; normals turn only */
  }
  rot[t] = make_float4(out.x, out.y, out.z, 0.0f);
}

constexpr int LIST_LANES = 32;   /* particles per wave in the polyhedra list kernel (a power of two) */

/* Bounding-sphere contact list (the sphere broad phase: Kernel_BroadCollision-
 * Detection_NonSymmetry, 32 slots, full count) and the pair work queue. */
__global__ __launch_bounds__(256) void k_poly_list(const DevParams* __restrict__ Pp, int n,
                                                   const uint2* __restrict__ cells,
                                                   const float4* __restrict__ pos_r,
                                                   const float4* __restrict__ omg,
                                                   uint32_t* __restrict__ list32, uint32_t* __restrict__ nn,
                                                   uint32_t* __restrict__ queue, uint32_t* __restrict__ qcount,
                                                   uint32_t* status) {
  const DevParams& P = *Pp;
  __shared__ __align__(16) unsigned char wlds[4][sizeof(WaveStage)];
  /* 32 consecutive particles per wave (lanes 32-63 idle): twice the waves
   * of one particle per lane, to hide the scan's memory latency at the
   * particle counts of a polyhedra run */
  const int lane = threadIdx.x & 63;
  const bool own = lane < LIST_LANES;
  const int i = (blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6)) * LIST_LANES + (lane & (LIST_LANES - 1));
  const bool live = own && i < n && packed_pid(omg[i].w) >= 0;
  const float4 pr = live ? pos_r[i] : make_float4(0.f, 0.f, 0.f, 0.f);
  /* the staged stencil scan of a wave of consecutive particles */
  int c = for_each_contact_wave(P, cells, pos_r, (uint32_t)n, i, xyz(pr), pr.w, live,
                                *reinterpret_cast<WaveStage*>(wlds[threadIdx.x >> 6]), [&](int j, float4, int k) {
                                  if (k < LIST_STRIDE) list32[(size_t)i * LIST_STRIDE + k] = (uint32_t)j;
                                });
  if (!live) c = 0;
  if (c > LIST_STRIDE) atomicOr(status, ST_LIST_OVERFLOW);
  if (own && i < n) nn[i] = (uint32_t)c;
  const int m = c < LIST_STRIDE ? c : LIST_STRIDE;
  /* wave prefix of the per-lane pair counts, one atomic per wave */
  int incl = m;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int o = __shfl_up(incl, d, 64);
    if (lane >= d) incl += o;
  }
  const int total = __shfl(incl, 63, 64);
  uint32_t base = 0u;
  if (lane == 63 && total > 0) base = atomicAdd(qcount, (uint32_t)total);
  base = __shfl(base, 63, 64);
  const uint32_t at = base + (uint32_t)(incl - m);
  const uint32_t cap = (uint32_t)n * LIST_STRIDE;
  for (int k = 0; k < m; ++k)
    if (at + (uint32_t)k < cap) queue[at + k] = (uint32_t)i * LIST_STRIDE + (uint32_t)k;
}

/* One queued pair per thread, grid-stride so that every wave exits once the
 * queue (at most n * 32 entries) is drained: the separating-plane test of
 * each body's face turned most towards the other.  The pairs it does not
 * separate (about 1 in 5 in a dense pack of hulls) go to a second queue,
 * where a wave per pair runs the full separating-plane test, then the
 * vertex-face and edge-edge searches. */
__global__ __launch_bounds__(256) void k_poly_pairs(const DevParams* __restrict__ Pp,
                                                    const bdem_poly_shape* __restrict__ shapes,
                                                    const ShapeHdr* __restrict__ hdr, int n, PolyLayout L,
                                                    const float4* __restrict__ rot,
                                                    const float4* __restrict__ pos, const float4* __restrict__ vel,
                                                    const float4* __restrict__ omg,
                                                    const uint32_t* __restrict__ list32,
                                                    const uint32_t* __restrict__ queue,
                                                    uint32_t* __restrict__ qcount, uint2* __restrict__ equeue,
                                                    float4* __restrict__ pf, float4* __restrict__ pt,
                                                    uint2* __restrict__ hint) {
  const DevParams& P = *Pp;
  const uint32_t total = qcount[0];
  const uint32_t cap = (uint32_t)n * LIST_STRIDE;
  const uint32_t cnt = total < cap ? total : cap;
  const int lane = threadIdx.x & 63;
  /* Each XCD (workgroup w runs on XCD w mod 8) takes one contiguous eighth
   * of the queue, and its blocks stride through it together, so the turned
   * shapes of neighbouring particles are read through that XCD's L2.  Every
   * lane of a wave runs the same number of iterations (the bounds are whole
   * waves) so that the append below sees the whole wave. */
  const uint32_t xcd = blockIdx.x % NUM_XCD, per = gridDim.x / NUM_XCD;
  const uint32_t cnt_w = (cnt + 63u) & ~63u;
  const uint32_t chunk = ((cnt_w / 64u + NUM_XCD - 1u) / NUM_XCD) * 64u;
  const uint32_t lo = xcd * chunk, hi = lo + chunk < cnt_w ? lo + chunk : cnt_w;
  const uint32_t stride = per * blockDim.x;
  for (uint32_t q = lo + (blockIdx.x / NUM_XCD) * blockDim.x + threadIdx.x; q < hi; q += stride) {
    bool need = false;
    uint32_t slot = 0u, partner = 0u;
    if (q < cnt) {
      slot = queue[q];
      const PairOps o = pair_ops(P, slot, list32, pos, vel, omg);
      partner = (uint32_t)o.b;
      const ShapeHdr ha = hdr[o.sa], hb = hdr[o.sb];
      const Shape A = shape_of(rot, L, o.a, ha);
      const Shape B = shape_of(rot, L, o.b, hb);
      const float vbA = ha.vbound * P.vbound_scale, vbB = hb.vbound * P.vbound_scale;
      /* the pair's hint, kept by P_ID and list slot (rows past n: no hint) */
      const int ida = packed_pid(omg[o.a].w);
      const uint32_t idb = (uint32_t)packed_pid(omg[o.b].w);
      uint2* hp = (ida >= 0 && ida < n) ? hint + (size_t)ida * LIST_STRIDE + (slot % LIST_STRIDE) : nullptr;
      uint32_t code = HINT_NONE;
      bool sep = false;
      if (hp) {
        const uint2 h = *hp;
        if (hint_separates(h, idb, A, B, o.PosA, o.PosB, vbA, vbB)) { sep = true; code = h.y; }
      }
      if (!sep) sep = quick_separated(A, B, o.PosA, o.PosB, vbA, vbB, code);
      if (hp) *hp = make_uint2(idb, sep ? code : HINT_NONE);
      need = !sep;
      if (!need)
        pair_store(P, shapes[o.sa], o, PColl{false, C_SP, 0.0f, mk3(0.f, 0.f, 0.f), mk3(0.f, 0.f, 0.f)}, slot, pf, pt);
    }
    const uint64_t mask = __ballot(need);
    if (mask) {
      const int leader = __ffsll((unsigned long long)mask) - 1;
      uint32_t base = 0u;
      if (lane == leader) base = atomicAdd(&qcount[1], (uint32_t)__popcll(mask));
      base = __shfl(base, leader, 64);
      const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
      const uint32_t at = base + (uint32_t)__popcll(mask & lt);
      if (need && at < cap) equeue[at] = make_uint2(slot, partner);
    }
  }
}

/* One pair's bodies staged in LDS by its wave: turned vertices, face
 * normals and centroids, and edge lists of A (row 0) and B (row 1). */
/* A pair's contact, kept until the chunk's pair law runs. */
struct PairHit { uint32_t hit; float dis, nx, ny, nz, cx, cy, cz; };
constexpr uint32_t NARROW_CHUNK = 32;   /* largest chunk of the narrow kernel's queue */

struct PairStage {
  float4 v[2][BDEM_SHAPE_MAX_VERTEX];
  float4 fn[2][BDEM_SHAPE_MAX_FACE];
  float4 fc[2][BDEM_SHAPE_MAX_FACE];
  uint2 e[2][BDEM_SHAPE_MAX_EDGE];
  uint2 vo[2][BDEM_SHAPE_MAX_FACE];      /* face vertex orders, four to a word */
  float area[2][BDEM_SHAPE_MAX_FACE];
  int fnv[2][BDEM_SHAPE_MAX_FACE];
  unsigned char ia[BDEM_SHAPE_MAX_EDGE], ib[BDEM_SHAPE_MAX_EDGE];   /* edge lists of the pruned search */
};

/* The pairs the first pass could not separate, one wave per pair,
 * grid-stride over the queue by waves: vertices of A in B, vertices of B in
 * A (normal reversed), then edge-edge with its depth limited to 0.025 rA
 * (DEM_Kernels_Compute.cuh:852-904).  The wave first copies both bodies
 * into its LDS stage, so the searches read LDS instead of chains of
 * dependent global loads.
 *
 * A wave takes its pairs from the queue in chunks by an atomic counter
 * (qcount[2]), guided: a chunk is about 1 / (2 x waves) of what is left, 4
 * to 32 pairs, so waves that drew expensive pairs (the edge-edge search)
 * do not hold up the launch.  The searches leave each pair's contact in
 * LDS, and the pair law then runs for the whole chunk, a lane per pair.
 *
 * Staging is one round trip to memory per pair.  Each lane fetches one of
 * the chunk's queue entries and the shape rows of its two bodies ahead of
 * time; a pair's values are then read from that lane, so they are
 * wave-uniform (scalar loads for the pair's state and shape counts), and
 * every row of the stage is loaded without a condition on the shape's
 * counts (lanes 0-31 body A, 32-63 body B; rows past the counts are inside
 * the record and never read) before the first LDS write.
 *
 * Five waves per SIMD: at most 96 VGPRs without spills, and five blocks of
 * 25 KiB of LDS (four stages and their chunk's pair records) per CU. */
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void k_poly_narrow(const DevParams* __restrict__ Pp,
                                                     const bdem_poly_shape* __restrict__ shapes,
                                                     const ShapeHdr* __restrict__ hdr, int n, PolyLayout L,
                                                     const float4* __restrict__ rot,
                                                     const float4* __restrict__ pos, const float4* __restrict__ vel,
                                                     const float4* __restrict__ omg,
                                                     uint32_t* __restrict__ qcount,
                                                     const uint2* __restrict__ equeue,
                                                     float4* __restrict__ pf, float4* __restrict__ pt, int energy,
                                                     uint2* __restrict__ hint) {
  const DevParams& P = *Pp;
  float* const tally = energy ? P.tally : nullptr;   /* dissipation tallies (EnergyCalc) */
  __shared__ PairStage stage[4];
  __shared__ PairHit hits[4][NARROW_CHUNK];
  PairStage& S = stage[threadIdx.x >> 6];
  PairHit* const R = hits[threadIdx.x >> 6];
  const int lane = threadIdx.x & 63;
  const int h = lane >> 5, k = lane & 31;
  const int kv = k < L.pv ? k : L.pv - 1, kf = k < L.pf ? k : L.pf - 1;
  const uint32_t cap = (uint32_t)n * LIST_STRIDE;
  const uint32_t cnt = qcount[1] < cap ? qcount[1] : cap;
  const uint32_t waves = gridDim.x * (blockDim.x / 64u);
  uint32_t seen = 0u;   /* the counter's value at this wave's last grab */
  for (;;) {
    uint32_t size = (cnt > seen ? cnt - seen : 0u) / (2u * waves);
    size = size < 4u ? 4u : (size > NARROW_CHUNK ? NARROW_CHUNK : size);
    uint32_t base = 0u;
    if (lane == 0) base = atomicAdd(&qcount[2], size);
    base = __builtin_amdgcn_readfirstlane(base);
    if (base >= cnt) break;
    seen = base + size;
    const uint32_t m = cnt - base < size ? cnt - base : size;
    uint32_t lslot = cap, lpart = 0u, lsa = 0u, lsb = 0u;
    if ((uint32_t)lane < m) {
      const uint2 item = equeue[base + (uint32_t)lane];   /* (slot, partner) */
      if (item.x < cap && item.y < (uint32_t)n) {
        lslot = item.x;
        lpart = item.y;
        lsa = (uint32_t)shape_index(P, omg[item.x / LIST_STRIDE].w);
        lsb = (uint32_t)shape_index(P, omg[item.y].w);
      }
    }
    for (int j = 0; j < (int)m; ++j) {
      const uint32_t slot = __builtin_amdgcn_readlane(lslot, j);
      if (slot >= cap) continue;
      const uint32_t partner = __builtin_amdgcn_readlane(lpart, j);
      const int sa = __builtin_amdgcn_readlane(lsa, j), sb = __builtin_amdgcn_readlane(lsb, j);
      const bdem_poly_shape& pa = shapes[sa];
      const bdem_poly_shape& pb = shapes[sb];
      const ShapeHdr ha = hdr[sa], hb = hdr[sb];
      /* turned rows at the table's largest counts, edges and face polygons */
      const float4* rh = rot + (size_t)(h ? partner : slot / LIST_STRIDE) * L.stride;
      const float4 v4 = rh[kv], n4 = rh[L.pv + kf], c4 = rh[L.pv + L.pf + kf];
      const bdem_Edge_Def eA = pa.edge[lane], eB = pb.edge[lane];
      const bdem_Surface& F = (h ? pb : pa).face[k];
      uint32_t w0 = 0u, w1 = 0u;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        w0 |= ((uint32_t)F.vertex_Order[q] & 0xFFu) << (8 * q);
        w1 |= ((uint32_t)F.vertex_Order[q + 4] & 0xFFu) << (8 * q);
      }
      const float farea = F.area;
      const int fnv = F.num_vertex;
      const PairOps o = pair_ops_ab(P, slot, partner, pos, vel, omg);
      const float vbA = ha.vbound * P.vbound_scale, vbB = hb.vbound * P.vbound_scale;
      /* the previous pair's reads are done before the stage is overwritten */
      __builtin_amdgcn_wave_barrier();
      S.v[h][k] = v4;
      S.fn[h][k] = n4;
      S.fc[h][k] = c4;
      S.e[0][lane] = make_uint2(eA.point_vindex[0], eA.point_vindex[1]);
      S.e[1][lane] = make_uint2(eB.point_vindex[0], eB.point_vindex[1]);
      S.vo[h][k] = make_uint2(w0, w1);
      S.area[h][k] = farea;
      S.fnv[h][k] = fnv;
      __builtin_amdgcn_wave_barrier();
      const Shape A{S.v[0], S.fn[0], S.fc[0], ha.nv, ha.nf, S.vo[0], S.area[0], S.fnv[0]};
      const Shape B{S.v[1], S.fn[1], S.fc[1], hb.nv, hb.nf, S.vo[1], S.area[1], S.fnv[1]};
      const PairDirs D = pair_dirs(A, B, o.PosA, o.PosB);
      const uint32_t code = separated_wave(A, B, o.PosA, o.PosB, D);
      PColl c{false, C_SP, 0.0f, mk3(0.f, 0.f, 0.f), mk3(0.f, 0.f, 0.f)};
      if (code != HINT_NONE) {
        if (lane == 0) {
          const int ida = packed_pid(omg[o.a].w);
          if (ida >= 0 && ida < n)   /* the next step tries this face first */
            hint[(size_t)ida * LIST_STRIDE + (slot % LIST_STRIDE)] = make_uint2((uint32_t)packed_pid(omg[o.b].w), code);
        }
      } else {
        c = vertex_faces_pair(o.PosA, A, o.PosB, B, D);
        if (!c.hit) {
          const f3 RelPos = (o.PosA - o.PosB) / length(o.PosA - o.PosB);
          c = edges_wave(A, S.e[0], ha.ne, B, S.e[1], hb.ne, RelPos, o.PosA, o.PosB, vbA, vbB,
                         ha.plane_slack * P.vbound_scale, hb.plane_slack * P.vbound_scale, S.ia, S.ib, D.fA, D.fB);
          if (c.hit) c.dis = fminf(0.025f * ha.radius, c.dis);
        }
      }
      /* the pair law runs for the whole chunk at once below, a lane per pair */
      if (lane == 0) R[j] = PairHit{c.hit ? 1u : 0u, c.dis, c.n.x, c.n.y, c.n.z, c.cp.x, c.cp.y, c.cp.z};
    }
    __builtin_amdgcn_wave_barrier();
    if ((uint32_t)lane < m && lslot < cap) {
      const PairHit r = R[lane];
      const PairOps o = pair_ops_ab(P, lslot, lpart, pos, vel, omg);
      pair_store(P, shapes[lsa], o, PColl{r.hit != 0u, C_UNDEF, r.dis, mk3(r.nx, r.ny, r.nz), mk3(r.cx, r.cy, r.cz)},
                 lslot, pf, pt, tally);
    }
    __builtin_amdgcn_wave_barrier();   /* the records are read before the next chunk writes them */
  }
}

/* Per particle: gravity, the pair slots in list order, walls, integration,
 * discharge, next key, changed flag and changed-list append (as k_step).
 * FORCES_ONLY stops after the forces and writes them (debug stage). */
template <bool FORCES_ONLY, bool LIFTERS>
__global__ __launch_bounds__(256) void k_poly_finish(
    const DevParams* __restrict__ Pp, const bdem_poly_shape* __restrict__ shapes, int n, PolyLayout L,
    const float4* __restrict__ rot, const uint32_t* __restrict__ nn, const float4* __restrict__ pf,
    const float4* __restrict__ pt, const float4* __restrict__ pos_in, const float4* __restrict__ vel_in,
    const float4* __restrict__ omg_in, const float4* __restrict__ quat_in, float4* __restrict__ pos_out,
    float4* __restrict__ vel_out, float4* __restrict__ omg_out, float4* __restrict__ quat_out,
    uint32_t* __restrict__ keys_out, const uint32_t* __restrict__ sorted_keys, uint8_t* __restrict__ changed,
    unsigned long long* __restrict__ sbuf, uint32_t* __restrict__ append, const uint32_t* __restrict__ hwc_in,
    uint32_t* __restrict__ hwc_out, float* __restrict__ wall_force, int do_kill, uint32_t* status,
    int* kill_count, float* __restrict__ fpp, float* __restrict__ tpp, float* __restrict__ fw,
    float* __restrict__ tw, int energy) {
  const DevParams& P = *Pp;
  float* const tally = energy ? P.tally : nullptr;   /* dissipation tallies (EnergyCalc) */
  const int i = xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  const bool in = i < n;
  bool chg = false;
  uint32_t new_key = 0u;
  if (in) {
    const float4 pr = pos_in[i], vm = vel_in[i], ot = omg_in[i], q4 = quat_in[i];
    int pid = packed_pid(ot.w);
    const int ta = packed_type(ot.w);
    f3 p = xyz(pr), v = xyz(vm), w = xyz(ot);
    const float r = pr.w, m = vm.w;
    QuatF q{q4.x, q4.y, q4.z, q4.w};
    if (pid >= 0) {
      const bdem_poly_shape& po = shapes[shape_index(P, ot.w)];
      f3 Fpp = cvt(P.gravity) * m, Tpp = mk3(0.f, 0.f, 0.f);
      const uint32_t c = nn[i];
      const int cnt = c < (uint32_t)LIST_STRIDE ? (int)c : LIST_STRIDE;
      for (int k = 0; k < cnt; ++k) {
        const float4 f = pf[(size_t)i * LIST_STRIDE + k];
        if (f.w != 0.0f) {
          Fpp = Fpp + xyz(f);
          Tpp = Tpp + xyz(pt[(size_t)i * LIST_STRIDE + k]);
        }
      }
      if (!P.rotation) Tpp = mk3(0.f, 0.f, 0.f);
      const Shape A = shape_of(rot, L, i, po);
      f3 Fw, Tw;
      wall_forces_poly(P, ta, p, r, v, w, m, A, po.vbound * P.vbound_scale, Fw, Tw, tally);
      f3 Fl = mk3(0.f, 0.f, 0.f), Tl = Fl;
      if (LIFTERS && P.num_dobj > 0) lifter_forces_poly(P, ta, p, v, w, m, A, po, Fl, Tl, tally);
      if (FORCES_ONLY) {
        fpp[3 * i] = Fpp.x; fpp[3 * i + 1] = Fpp.y; fpp[3 * i + 2] = Fpp.z;
        tpp[3 * i] = Tpp.x; tpp[3 * i + 1] = Tpp.y; tpp[3 * i + 2] = Tpp.z;
        fw[3 * i] = Fw.x; fw[3 * i + 1] = Fw.y; fw[3 * i + 2] = Fw.z;
        tw[3 * i] = Tw.x; tw[3 * i + 1] = Tw.y; tw[3 * i + 2] = Tw.z;
        return;
      }
      if (wall_force) {
        wall_force[3 * i] = Fw.x; wall_force[3 * i + 1] = Fw.y; wall_force[3 * i + 2] = Fw.z;
      }
      integrate(P, 0.0f, m, Fpp, Tpp, Fw, Tw, Fl, Tl, p, v, w, &q, po.InertiaT);
      /* Remove_Silo_Discharged_Particles, GPU_Memory_Transfer.cuh:650-664 */
      if (do_kill && p.y < P.kill_height) {
        p = cvt(P.max_size);
        pid = -1;
        atomicAdd(kill_count, 1);
      }
      pos_out[i] = make_float4(p.x, p.y, p.z, r);
      vel_out[i] = make_float4(v.x, v.y, v.z, m);
      omg_out[i] = make_float4(w.x, w.y, w.z, __int_as_float(pack_id_type(pid, ta)));
      quat_out[i] = make_float4(q.w, q.x, q.y, q.z);
      if (hwc_out) hwc_out[i] = 0u;
    } else {
      if (FORCES_ONLY) {
        for (int k = 0; k < 3; ++k) { fpp[3 * i + k] = 0.f; tpp[3 * i + k] = 0.f; fw[3 * i + k] = 0.f; tw[3 * i + k] = 0.f; }
        return;
      }
      pos_out[i] = pr;
      vel_out[i] = vm;
      omg_out[i] = ot;
      quat_out[i] = q4;
      if (hwc_out) hwc_out[i] = hwc_in ? hwc_in[i] : 0u;
      if (wall_force) { wall_force[3 * i] = 0.f; wall_force[3 * i + 1] = 0.f; wall_force[3 * i + 2] = 0.f; }
    }
    new_key = key_of(P, p, pid, status);
    keys_out[i] = new_key;
    chg = new_key != sorted_keys[i];
    changed[i] = chg ? 1 : 0;
  }
  if (FORCES_ONLY) return;
  const uint64_t mask = __ballot(chg);
  if (mask) {
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((unsigned long long)mask) - 1;
    uint32_t base = 0u;
    if (lane == leader) base = atomicAdd(append, (uint32_t)__popcll(mask));
    base = __shfl(base, leader, 64);
    const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    const uint32_t at = base + (uint32_t)__popcll(mask & lt);
    if (chg) {
      if (at < (uint32_t)n) sbuf[at] = ((unsigned long long)new_key << 32) | (uint32_t)i;
      else atomicOr(status, ST_SORT_ERROR);
    }
  }
}

/* ------------------------------------------------------------------ launchers */
/* Blocks of 256 threads of `kernel` that the device holds at once (every
 * CU full), a multiple of the 8 XCDs: the grid of the two pair kernels, so
 * that no partial second round of blocks starts late.  Queried once per
 * device (slot: which kernel). */
template <class K>
static unsigned resident_blocks(K kernel, int slot) {
  static std::atomic<unsigned> cache[2][64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return (unsigned)POLY_PAIR_BLOCKS;
  unsigned nb = cache[slot][dev].load(std::memory_order_relaxed);
  if (nb == 0u) {
    int per = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, 256, 0) != hipSuccess || per <= 0) per = 8;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
    nb = (unsigned)(per * cus) / NUM_XCD * NUM_XCD;
    if (nb < (unsigned)NUM_XCD) nb = NUM_XCD;
    cache[slot][dev].store(nb, std::memory_order_relaxed);
  }
  return nb;
}

static void launch_poly_forces_common(const DevParams* dP, const bdem_poly_shape* dshape, const StepBuffers& b,
                                      const PolyBuffers& pb, bool energy, hipStream_t s) {
  const size_t rot_threads = (size_t)b.n * pb.L.stride;
  hipLaunchKernelGGL(k_poly_rotate, dim3((unsigned)((rot_threads + 255) / 256)), dim3(256), 0, s, dP, pb.local,
                     b.n, pb.L, b.pos_srt, b.omg_srt, b.quat_srt, pb.rot, pb.qcount);
  hipLaunchKernelGGL(k_poly_list, dim3((b.n + 4 * LIST_LANES - 1) / (4 * LIST_LANES)), dim3(256), 0, s, dP, b.n,
                     b.cells, b.pos_srt,
                     b.omg_srt, pb.list, pb.nn, pb.queue, pb.qcount, b.status);
  /* both pair kernels loop over their queues with the resident grid; the
   * first splits its queue by XCD (a grid of whole XCD multiples) */
  static_assert(POLY_PAIR_BLOCKS % NUM_XCD == 0, "the pair kernel splits its queue by XCD");
  hipLaunchKernelGGL(k_poly_pairs, dim3(resident_blocks(k_poly_pairs, 0)), dim3(256), 0, s, dP, dshape, pb.hdr, b.n,
                     pb.L, pb.rot, b.pos_srt, b.vel_srt, b.omg_srt, pb.list, pb.queue, pb.qcount, pb.equeue, pb.pf,
                     pb.pt, pb.hint);
  hipLaunchKernelGGL(k_poly_narrow, dim3(resident_blocks(k_poly_narrow, 1)), dim3(256), 0, s, dP, dshape, pb.hdr, b.n,
                     pb.L, pb.rot,
                     b.pos_srt, b.vel_srt, b.omg_srt, pb.qcount, pb.equeue, pb.pf, pb.pt, energy ? 1 : 0, pb.hint);
}

/* World-frame shapes of the current (unsorted) state into pb.rot: the
 * point-cloud output, between steps (the next step rewrites pb.rot). */
void launch_poly_world_shapes(const DevParams* dP, const bdem_poly_shape* dshape, const StepBuffers& b,
                              const PolyBuffers& pb, hipStream_t s) {
  const size_t rot_threads = (size_t)b.n * pb.L.stride;
  if (rot_threads == 0) return;
  hipLaunchKernelGGL(k_poly_rotate, dim3((unsigned)((rot_threads + 255) / 256)), dim3(256), 0, s, dP, pb.local,
                     b.n, pb.L, b.pos_cur, b.omg_cur, b.quat_cur, pb.rot, pb.qcount);
}

void launch_poly_step(const DevParams* dP, const bdem_poly_shape* dshape, const StepBuffers& b, const PolyBuffers& pb,
                      const uint32_t* sk, int ap, bool do_kill, bool want_wf, bool lifters, bool energy,
                      hipStream_t s) {
  launch_poly_forces_common(dP, dshape, b, pb, energy, s);
  /* the lifter search needs private arrays (scratch): only its own instantiation carries it */
  auto k = lifters ? k_poly_finish<false, true> : k_poly_finish<false, false>;
  hipLaunchKernelGGL(k, dim3((b.n + 255) / 256), dim3(256), 0, s, dP, dshape, b.n, pb.L,
                     pb.rot, pb.nn, pb.pf, pb.pt, b.pos_srt, b.vel_srt, b.omg_srt, b.quat_srt, b.pos_cur,
                     b.vel_cur, b.omg_cur, b.quat_cur, b.keys_new, sk, b.changed, b.sbuf,
                     b.meta + META_APPEND + ap, b.hwc_srt, b.hwc_cur, want_wf ? b.wall_force : nullptr,
                     do_kill ? 1 : 0, b.status, b.kill_count, nullptr, nullptr, nullptr, nullptr, energy ? 1 : 0);
}

void launch_poly_forces(const DevParams* dP, const bdem_poly_shape* dshape, const StepBuffers& b, const PolyBuffers& pb,
                        float* fpp, float* tpp, float* fw, float* tw, hipStream_t s) {
  launch_poly_forces_common(dP, dshape, b, pb, false, s);
  hipLaunchKernelGGL((k_poly_finish<true, false>), dim3((b.n + 255) / 256), dim3(256), 0, s, dP, dshape, b.n, pb.L,
                     pb.rot, pb.nn, pb.pf, pb.pt, b.pos_srt, b.vel_srt, b.omg_srt, b.quat_srt, nullptr, nullptr,
                     nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 0,
                     b.status, b.kill_count, fpp, tpp, fw, tw, 0);
}

}  // namespace bdem
