/*
 * bdem_shapes.cpp - convex polyhedron shapes from point sets (host only).
 *
 * The hull of a small point set (at most BDEM_SHAPE_MAX_VERTEX points) by
 * enumerating supporting planes: three points span a face plane when every
 * other point lies on one side of it.  Points on that plane form one face
 * (a polygon when more than three are coplanar), emitted once (the first
 * time its point set is met), ordered
 * counter-clockwise seen from outside.  O(n^4) for n <= 32 points: a few
 * microseconds per shape.
 *
 * The derived data follow the deck reader's polyhedron set-up
 * (ReadPObject_Poly + Make_Particle_LocalInfo, Input/KSimulationData.cpp:
 * 1180-1530), in float: COM = mean of the vertices, Local_Vertex =
 * vertex - COM, bounding radius = largest |Local_Vertex|, face centroid =
 * mean of its vertices, face normal = normalise(cross(v1 - v2, v2 - v0)) for
 * triangles and normalise(cross(v2 - v3, v3 - v0)) for polygons, area = the
 * fan about the centroid.  Mass and inertia come from the solid (the
 * reference reads a normalised tensor from its deck): volume and inertia
 * about COM by tetrahedra (COM, face centroid, consecutive face vertices),
 * in double; InertiaT is the inverse of the inertia tensor.
 */
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>
#include "bdem_deck.h"

namespace {

struct D3 { double x, y, z; };
inline D3 sub(D3 a, D3 b) { return D3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D3 crs(D3 a, D3 b) { return D3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double dt(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

struct F3 { float x, y, z; };
inline F3 fsub(F3 a, F3 b) { return F3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline F3 fcrs(F3 a, F3 b) { return F3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float fdot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

/* Inverse of a symmetric 3x3 matrix (row-major); false when singular. */
bool inv3(const double m[9], double out[9]) {
  const double a = m[0], b = m[1], c = m[2], d = m[3], e = m[4], f = m[5], g = m[6], h = m[7], i = m[8];
  const double A = e * i - f * h, B = -(d * i - f * g), C = d * h - e * g;
  const double det = a * A + b * B + c * C;
  if (!(std::fabs(det) > 0.0)) return false;
  out[0] = A / det; out[1] = -(b * i - c * h) / det; out[2] = (b * f - c * e) / det;
  out[3] = B / det; out[4] = (a * i - c * g) / det; out[5] = -(a * f - c * d) / det;
  out[6] = C / det; out[7] = -(a * h - b * g) / det; out[8] = (a * e - b * d) / det;
  return true;
}

int build_one(const bdem_float3* pts, int n, float density, float tol, bdem_poly_shape& S) {
  std::memset(&S, 0, sizeof(S));
  if (n < 4 || n > BDEM_SHAPE_MAX_VERTEX) return BDEM_EINVAL;
  std::vector<D3> P(n);
  double scale = 0.0;
  for (int k = 0; k < n; ++k) {
    P[k] = D3{pts[k].x, pts[k].y, pts[k].z};
    scale = std::max(scale, std::sqrt(dt(sub(P[k], P[0]), sub(P[k], P[0]))));
  }
  if (!(scale > 0.0)) return BDEM_EINVAL;
  const double eps = 1e-9 * scale;
  /* faces as vertex-index lists, counter-clockwise from outside */
  std::vector<std::vector<int>> faces;
  std::vector<std::vector<int>> planes;   /* each emitted face's point set, ascending */
  std::vector<char> used(n, 0);
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j)
      for (int k = j + 1; k < n; ++k) {
        D3 nrm = crs(sub(P[j], P[i]), sub(P[k], P[i]));
        const double ln = std::sqrt(dt(nrm, nrm));
        if (!(ln > eps * scale)) continue;   /* collinear */
        nrm = D3{nrm.x / ln, nrm.y / ln, nrm.z / ln};
        int pos = 0, neg = 0;
        std::vector<int> on;
        for (int l = 0; l < n; ++l) {
          const double s = dt(nrm, sub(P[l], P[i]));
          if (s > eps) ++pos;
          else if (s < -eps) ++neg;
          else on.push_back(l);
        }
        if (pos && neg) continue;            /* not a supporting plane */
        if (!pos && !neg) return BDEM_EINVAL; /* all points coplanar */
        if (std::find(planes.begin(), planes.end(), on) != planes.end()) continue;   /* face already emitted */
        planes.push_back(on);
        if ((int)on.size() > 8) return BDEM_EINVAL;             /* Surface::vertex_Order[8] */
        if (pos) nrm = D3{-nrm.x, -nrm.y, -nrm.z};              /* outward */
        /* order the face's points counter-clockwise about the outward normal */
        D3 c{0, 0, 0};
        for (int l : on) { c.x += P[l].x; c.y += P[l].y; c.z += P[l].z; }
        c = D3{c.x / on.size(), c.y / on.size(), c.z / on.size()};
        D3 u = sub(P[on[0]], c);
        const double lu = std::sqrt(dt(u, u));
        u = D3{u.x / lu, u.y / lu, u.z / lu};
        const D3 w = crs(nrm, u);
        std::vector<std::pair<double, int>> ang;
        for (int l : on) {
          const D3 d = sub(P[l], c);
          ang.push_back({std::atan2(dt(d, w), dt(d, u)), l});
        }
        std::sort(ang.begin(), ang.end());
        std::vector<int> f;
        for (auto& a : ang) f.push_back(a.second);
        for (int l : f) used[l] = 1;
        faces.push_back(f);
      }
  if (faces.size() < 4 || (int)faces.size() > BDEM_SHAPE_MAX_FACE) return BDEM_EINVAL;
  /* only hull vertices are kept, renumbered in input order */
  std::vector<int> remap(n, -1);
  std::vector<F3> V;
  for (int l = 0; l < n; ++l)
    if (used[l]) { remap[l] = (int)V.size(); V.push_back(F3{pts[l].x, pts[l].y, pts[l].z}); }
  for (auto& f : faces)
    for (int& l : f) l = remap[l];
  const int nv = (int)V.size();
  S.num_vertex = nv;
  F3 tot{0.f, 0.f, 0.f};
  for (int l = 0; l < nv; ++l) tot = F3{tot.x + V[l].x, tot.y + V[l].y, tot.z + V[l].z};
  const float inv_n = 1.0f / (float)nv;
  const F3 com{tot.x * inv_n, tot.y * inv_n, tot.z * inv_n};
  S.COM = bdem_float3{com.x, com.y, com.z};
  float rad = 0.0f;
  double vmax = 0.0;
  for (int l = 0; l < nv; ++l) {
    const F3 lv = fsub(V[l], com);
    S.Local_Vertex[l] = bdem_float3{lv.x, lv.y, lv.z};
    rad = std::max(rad, std::sqrt(fdot(lv, lv)));
    vmax = std::max(vmax, std::sqrt(double(lv.x) * lv.x + double(lv.y) * lv.y + double(lv.z) * lv.z));
  }
  S.radius = rad;
  S.vbound = float(vmax * 1.0001);
  S.num_faces = (int)faces.size();
  double vol = 0.0, I[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  std::vector<std::pair<int, int>> edges;
  for (size_t fi = 0; fi < faces.size(); ++fi) {
    const std::vector<int>& f = faces[fi];
    bdem_Surface& sf = S.face[fi];
    sf.num_vertex = (int)f.size();
    sf.Tol = tol;
    F3 ct{0.f, 0.f, 0.f};
    for (size_t q = 0; q < f.size(); ++q) {
      sf.vertex_Order[q] = f[q];
      ct = F3{ct.x + V[f[q]].x, ct.y + V[f[q]].y, ct.z + V[f[q]].z};
    }
    const float kq = 1.0f / (float)f.size();
    ct = F3{kq * ct.x, kq * ct.y, kq * ct.z};
    sf.centroid = bdem_float3{ct.x, ct.y, ct.z};
    F3 nr = f.size() == 3 ? fcrs(fsub(V[f[1]], V[f[2]]), fsub(V[f[2]], V[f[0]]))
                          : fcrs(fsub(V[f[2]], V[f[3]]), fsub(V[f[3]], V[f[0]]));
    const float ln = std::sqrt(fdot(nr, nr));
    sf.normal = bdem_float3{nr.x / ln, nr.y / ln, nr.z / ln};
    float area = 0.0f;
    for (size_t q = 0; q < f.size(); ++q) {
      const F3 a = fsub(V[f[q]], ct), b = fsub(V[f[(q + 1) % f.size()]], ct);
      const F3 x = fcrs(a, b);
      area = area + 0.5f * std::sqrt(fdot(x, x));
    }
    sf.area = area;
    const D3 cc{double(ct.x) - com.x, double(ct.y) - com.y, double(ct.z) - com.z};
    for (size_t q = 0; q < f.size(); ++q) {
      const int a_ = f[q], b_ = f[(q + 1) % f.size()];
      const D3 a{double(V[a_].x) - com.x, double(V[a_].y) - com.y, double(V[a_].z) - com.z};
      const D3 b{double(V[b_].x) - com.x, double(V[b_].y) - com.y, double(V[b_].z) - com.z};
      /* tetrahedron (COM, a, b, centroid): volume and inertia from the
       * covariance of a simplex */
      const double tv = dt(a, crs(b, cc)) / 6.0;
      vol += tv;
      const D3 sm{a.x + b.x + cc.x, a.y + b.y + cc.y, a.z + b.z + cc.z};
      const D3 pp[4] = {a, b, cc, sm};
      double cov[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      for (const D3& x : pp) {
        const double xs[3] = {x.x, x.y, x.z};
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) cov[3 * r + c] += xs[r] * xs[c];
      }
      const double tr = (cov[0] + cov[4] + cov[8]) * tv / 20.0;
      for (int g = 0; g < 9; ++g) I[g] += (g % 4 == 0 ? tr : 0.0) - cov[g] * tv / 20.0;
      const std::pair<int, int> key{std::min(a_, b_), std::max(a_, b_)};
      if (std::find_if(edges.begin(), edges.end(), [&](const std::pair<int, int>& e) {
            return std::min(e.first, e.second) == key.first && std::max(e.first, e.second) == key.second;
          }) == edges.end())
        edges.push_back({a_, b_});
    }
  }
  if ((int)edges.size() > BDEM_SHAPE_MAX_EDGE || !(vol > 0.0)) return BDEM_EINVAL;
  S.num_edges = (int)edges.size();
  for (size_t e = 0; e < edges.size(); ++e) {
    S.edge[e].point_vindex[0] = (unsigned)edges[e].first;
    S.edge[e].point_vindex[1] = (unsigned)edges[e].second;
  }
  S.mass = float(double(density) * vol * 1e-6);
  double Id[9], Iinv[9];
  for (int g = 0; g < 9; ++g) Id[g] = I[g] * double(density) * 1e-6;
  if (!inv3(Id, Iinv)) return BDEM_EINVAL;
  for (int g = 0; g < 9; ++g) S.InertiaT[g] = float(Iinv[g]);
  return BDEM_OK;
}

}  // namespace

extern "C" int bdem_poly_shapes_from_points(int count, const int* npts, const bdem_float3* pts, int max_pts,
                                            float density, float face_tol, bdem_poly_shape* out, int* failed) {
  if (failed) *failed = -1;
  if (count < 0 || (count > 0 && (!npts || !pts || !out)) || max_pts < 4 || max_pts > BDEM_SHAPE_MAX_VERTEX ||
      !(density > 0.0f))
    return BDEM_EINVAL;
  for (int s = 0; s < count; ++s) {
    if (npts[s] > max_pts) { if (failed) *failed = s; return BDEM_EINVAL; }
    const int rc = build_one(pts + (size_t)s * max_pts, npts[s], density, face_tol, out[s]);
    if (rc != BDEM_OK) {
      if (failed) *failed = s;
      return rc;
    }
  }
  return BDEM_OK;
}
