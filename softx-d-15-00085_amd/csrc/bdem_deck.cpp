/*
 * bdem_deck.cpp - project decks and state files (host C++, no device code).
 *
 * Reads the reference's text formats token by token with the semantics of
 * istream extraction (a word is a run of non-blank characters; a number is
 * the longest numeric prefix, so "15," gives 15 and leaves ","), in the
 * order of KSimulationData (Input/KSimulationData.cpp:27-2028), and derives
 * the simulation parameters as Set_SimInfo does (main.cpp:469-833).  Float
 * geometry (normals, areas, centroids, volumes) is computed in float in the
 * reference's operation order; quantities the reference computes in double
 * (sphere mass, spring constants, mill rotation) are computed in double.
 */
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "bdem_deck.h"

namespace bdem {
int derive_springs(bdem_POBJ* pobj, int num_pobj, int mode, const double* cor_pp, const double* cor_surface,
                   const double* cor_dsurface, double tc_pp, const double* tc_ps, const double* tc_pd);
}

struct bdem_deck {
  std::string dir;
  bdem_deck_info info;
  bdem_SimulationInfo sim;
  bdem_InitConfig init;
  int per_object[8];
  std::vector<bdem_WOBJ> wobj;
  std::vector<bdem_POBJ> pobj;
  std::vector<bdem_DOBJ> dobj;
  std::vector<bdem_DOBJ> dobj_initial;   /* as read and repeated: Rotate_VolObject_angle starts from it */
};

namespace {

constexpr int MAX_WOBJ = 8;       /* WorldObject[8], GPU_Memory_Transfer.cuh:80 */
constexpr int MAX_POBJ = 6;       /* POBJ::PP[6], KSimulationObjects.h:290 */
constexpr int MAX_DOBJ = 32;      /* DynamicObject[32], GPU_Memory_Transfer.cuh:82 */
constexpr double PI_MAIN = 3.14159265;   /* main.h:58 (main.cpp's PI) */

/* ---- float3 arithmetic as the host helpers do it (Utilities/Host) ---- */
inline bdem_float3 mk(float x, float y, float z) { return bdem_float3{x, y, z}; }
inline bdem_float3 add(bdem_float3 a, bdem_float3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
inline bdem_float3 sub(bdem_float3 a, bdem_float3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
inline bdem_float3 scale(float s, bdem_float3 a) { return mk(s * a.x, s * a.y, s * a.z); }
inline bdem_float3 divs(bdem_float3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
inline float dot(bdem_float3 a, bdem_float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float length(bdem_float3 v) { return std::sqrt(dot(v, v)); }
inline bdem_float3 cross(bdem_float3 a, bdem_float3 b) {
  return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

/* ---- istream-like token reader over a whole file ---- */
class Reader {
 public:
  bool open(const std::string& path) {
    path_ = path;
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return fail("cannot open " + path);
    char buf[65536];
    size_t k;
    while ((k = std::fread(buf, 1, sizeof(buf), f)) > 0) text_.append(buf, k);
    std::fclose(f);
    return true;
  }
  const std::string& error() const { return err_; }
  bool ok() const { return err_.empty(); }

  std::string word() {
    skip_ws();
    size_t b = pos_;
    while (pos_ < text_.size() && !blank(text_[pos_])) ++pos_;
    if (b == pos_) fail("unexpected end of file");
    return text_.substr(b, pos_ - b);
  }
  std::string peek() {
    size_t save = pos_;
    std::string w = word_quiet();
    pos_ = save;
    return w;
  }
  void skip(int n) { for (int i = 0; i < n && ok(); ++i) word(); }
  /* the rest of the current line (getline) */
  std::string line() {
    size_t b = pos_;
    while (pos_ < text_.size() && text_[pos_] != '\n') ++pos_;
    std::string s = text_.substr(b, pos_ - b);
    if (pos_ < text_.size()) ++pos_;
    return s;
  }
  bool i32(int& v) {
    skip_ws();
    size_t b = pos_;
    if (pos_ < text_.size() && (text_[pos_] == '+' || text_[pos_] == '-')) ++pos_;
    size_t d = pos_;
    while (pos_ < text_.size() && text_[pos_] >= '0' && text_[pos_] <= '9') ++pos_;
    if (pos_ == d) { pos_ = b; return fail("expected an integer"); }
    v = (int)std::strtol(text_.substr(b, pos_ - b).c_str(), nullptr, 10);
    return true;
  }
  bool f64(double& v) {
    std::string s;
    if (!numeric(s)) return false;
    v = std::strtod(s.c_str(), nullptr);
    return true;
  }
  bool f32(float& v) {
    std::string s;
    if (!numeric(s)) return false;
    v = std::strtof(s.c_str(), nullptr);
    return true;
  }
  bool v3(bdem_float3& v) { return f32(v.x) && f32(v.y) && f32(v.z); }
  bool label_v3(bdem_float3& v) { word(); return ok() && v3(v); }
  bool label_i(int& v) { word(); return ok() && i32(v); }
  bool label_f(float& v) { word(); return ok() && f32(v); }
  bool label_d(double& v) { word(); return ok() && f64(v); }
  bool fail(const std::string& what) {
    if (err_.empty()) {
      int line_no = 1;
      for (size_t i = 0; i < pos_ && i < text_.size(); ++i) line_no += text_[i] == '\n';
      err_ = path_ + ":" + std::to_string(line_no) + ": " + what;
    }
    return false;
  }

 private:
  static bool blank(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f'; }
  void skip_ws() { while (pos_ < text_.size() && blank(text_[pos_])) ++pos_; }
  std::string word_quiet() {
    skip_ws();
    size_t b = pos_;
    while (pos_ < text_.size() && !blank(text_[pos_])) ++pos_;
    return text_.substr(b, pos_ - b);
  }
  /* the characters a floating-point extraction accepts: sign, digits, one
   * point, an exponent with optional sign and digits */
  bool numeric(std::string& s) {
    skip_ws();
    size_t b = pos_, n = text_.size();
    size_t p = pos_;
    if (p < n && (text_[p] == '+' || text_[p] == '-')) ++p;
    size_t digits = 0;
    while (p < n && std::isdigit((unsigned char)text_[p])) { ++p; ++digits; }
    if (p < n && text_[p] == '.') {
      ++p;
      while (p < n && std::isdigit((unsigned char)text_[p])) { ++p; ++digits; }
    }
    if (digits == 0) return fail("expected a number");
    if (p < n && (text_[p] == 'e' || text_[p] == 'E')) {
      size_t q = p + 1;
      if (q < n && (text_[q] == '+' || text_[q] == '-')) ++q;
      size_t e0 = q;
      while (q < n && std::isdigit((unsigned char)text_[q])) ++q;
      if (q > e0) p = q;
    }
    s = text_.substr(b, p - b);
    pos_ = p;
    return true;
  }
  std::string path_, text_, err_;
  size_t pos_ = 0;
};

struct Err {
  std::string msg;
  bool set(const std::string& m) { if (msg.empty()) msg = m; return false; }
};

bool starts_with(const std::string& s, const char* p) { return s.compare(0, std::strlen(p), p) == 0; }

void copy_str(char* dst, size_t cap, const std::string& s) {
  if (!cap) return;
  size_t k = s.size() < cap - 1 ? s.size() : cap - 1;
  std::memcpy(dst, s.data(), k);
  dst[k] = 0;
}

/* ---- .WOBJ (ReadWObject, Input/KSimulationData.cpp:709-879) ---- */
bool read_wobj(Reader& r, bdem_WOBJ& w, Err& e) {
  r.word(); r.word();                       /* Title: name */
  r.word(); r.word();                       /* Descp: desc */
  r.word();                                 /* Geometry-Specfications */
  if (!r.label_i(w.num_vertex)) return false;
  if (w.num_vertex < 0 || w.num_vertex > 16) return e.set("a world object holds at most 16 vertices");
  for (int j = 0; j < w.num_vertex; ++j) {
    int idx;
    if (!r.i32(idx) || !r.v3(w.vertex[j])) return false;
  }
  if (!r.label_i(w.num_surfaces)) return false;
  if (w.num_surfaces < 0 || w.num_surfaces > 16) return e.set("a world object holds at most 16 surfaces");
  for (int j = 0; j < w.num_surfaces; ++j) {
    int idx;
    if (!r.i32(idx)) return false;
    r.word();                               /* surface name */
    w.surface_type = idx == -1 ? BDEM_SURFACE_CYLINDER : BDEM_SURFACE_PLANE;
    bdem_Surface& s = w.surfaces[j];
    if (w.surface_type == BDEM_SURFACE_PLANE) {
      if (!r.label_f(s.Tol) || !r.label_i(s.num_vertex)) return false;
      if (s.num_vertex < 1 || s.num_vertex > 8) return e.set("a surface holds 1 to 8 vertices");
      bdem_float3 sum = mk(0, 0, 0);
      for (int f = 0; f < s.num_vertex; ++f) {
        int o;
        if (!r.i32(o)) return false;
        if (o < 1 || o > w.num_vertex) return e.set("surface vertex index out of range");
        s.vertex_Order[f] = o - 1;
        sum = add(sum, w.vertex[o - 1]);
      }
      const bdem_float3* V = w.vertex;
      const int* O = s.vertex_Order;
      if (s.num_vertex == 3) {
        bdem_float3 V1 = sub(V[O[2]], V[O[0]]), V2 = sub(V[O[1]], V[O[2]]);
        bdem_float3 n = cross(V2, V1);
        s.normal = divs(n, length(n));
        s.area = 0.5f * length(n);
      } else if (s.num_vertex == 4) {
        bdem_float3 V1 = sub(V[O[1]], V[O[0]]), V2 = sub(V[O[2]], V[O[1]]);
        bdem_float3 V3 = sub(V[O[3]], V[O[2]]), V4 = sub(V[O[0]], V[O[3]]);
        bdem_float3 n = cross(V2, V1);
        s.normal = divs(n, length(n));
        s.area = 0.5f * length(n) + 0.5f * length(cross(V3, V4));
      }
      s.centroid = scale(1.0f / s.num_vertex, sum);
    } else {
      bdem_Macro_Cylinder& c = w.cyl_geo;
      if (!r.label_f(c.radius)) return false;
      c.radius *= 0.50f;
      if (!r.label_f(c.height)) return false;
      r.word(); c.has_top_cap = r.word() == "true";
      r.word(); c.has_bot_cap = r.word() == "true";
      bdem_float3 offset;
      if (!r.label_v3(c.normal_top_cap) || !r.label_v3(c.normal_bot_cap) || !r.label_v3(offset)) return false;
      c.AXIS = sub(mk(1.0f, 1.0f, 1.0f), c.normal_bot_cap);
      c.center_bot_cap = add(offset, scale(c.radius, sub(mk(1.0f, 1.0f, 1.0f), c.normal_bot_cap)));
      c.center_top_cap = add(c.center_bot_cap, mk(c.normal_bot_cap.x * c.height, c.normal_bot_cap.y * c.height,
                                                   c.normal_bot_cap.z * c.height));
    }
  }
  return r.ok();
}

/* ---- material blocks shared by both POBJ kinds (:1102-1168, :1548-1617) ---- */
struct Material {
  double cor = 0, cor_s = 0, cor_d = 0, tc_pp = 0, tc_ps = 0, tc_pd = 0;
};

/* Contact times are kept as read; the caller scales them by the time step
 * (tc *= Delta_t, :1119-1121), which a World file may give after the
 * objects. */
bool read_material(Reader& r, int force_spec, int i, bdem_POBJ& p, Material& m) {
  if (force_spec == 0) {
    r.word();
    if (!r.label_d(m.cor) || !r.label_d(m.tc_pp) || !r.label_f(p.PP[0].Fric_static) ||
        !r.label_f(p.PP[0].Fric_kinetic))
      return false;
    r.word();
    if (!r.label_d(m.cor_s) || !r.label_d(m.tc_ps) || !r.label_f(p.surface_Fric_static) ||
        !r.label_f(p.surface_Fric_kinetic))
      return false;
    r.word();
    if (!r.label_d(m.cor_d) || !r.label_d(m.tc_pd) || !r.label_f(p.Dsurface_Fric_static) ||
        !r.label_f(p.Dsurface_Fric_kinetic))
      return false;
  } else {
    r.word();
    if (!r.label_d(m.cor) || !r.label_f(p.PP[i].Kn) || !r.label_f(p.PP[i].Fric_static) ||
        !r.label_f(p.PP[i].Fric_kinetic))
      return false;
    r.word();
    if (!r.label_d(m.cor_s) || !r.label_f(p.surface_Kn) || !r.label_f(p.surface_Fric_static) ||
        !r.label_f(p.surface_Fric_kinetic))
      return false;
    r.word();
    if (!r.label_d(m.cor_d) || !r.label_f(p.Dsurface_Kn) || !r.label_f(p.Dsurface_Fric_static) ||
        !r.label_f(p.Dsurface_Fric_kinetic))
      return false;
  }
  return r.ok();
}

/* ---- .POBJ sphere (ReadPObject_Sphere, :1082-1169) ---- */
bool read_pobj_sphere(Reader& r, int force_spec, int i, bdem_POBJ& p, Material& m) {
  r.word(); r.word(); r.word(); r.word(); r.word();   /* Title: name Descp: desc Geometry_Specfications */
  if (!r.label_f(p.radius)) return false;
  p.radius = (float)(p.radius * 0.500);
  double den;
  if (!r.label_d(den)) return false;
  p.mass = (float)((4.0 / 3.0) * 2.0 * std::asin(1.0) * std::pow((double)p.radius, 3.0) * den * 1E-6);
  return read_material(r, force_spec, i, p, m);
}

/* ---- .POBJ polyhedron (ReadPObject_Poly, :1224-1617) ---- */
bool read_pobj_poly(Reader& r, int force_spec, int i, bdem_POBJ& p, Material& m, bdem_float3& p_size,
                    Err& e) {
  r.word(); r.word(); r.word(); r.word();   /* Title: name Descp: desc */
  int format;
  if (!r.label_i(format)) return false;
  bdem_float3 offset = mk(0, 0, 0);
  if (format == 1 && !r.label_v3(offset)) return false;
  r.skip(3);
  if (!r.label_i(p.num_vertex)) return false;
  if (p.num_vertex < 1 || p.num_vertex > 32) return e.set("a polyhedron holds 1 to 32 vertices");
  p_size = mk(0, 0, 0);
  bdem_float3 sum = mk(0, 0, 0);
  for (int k = 0; k < p.num_vertex; ++k) {
    if (format == 0) {
      int idx;
      if (!r.i32(idx) || !r.v3(p.vertex[k])) return false;
    } else {
      if (!r.v3(p.vertex[k])) return false;
      p.vertex[k] = add(p.vertex[k], offset);
    }
    if (p.vertex[k].x > p_size.x) p_size.x = p.vertex[k].x;
    if (p.vertex[k].y > p_size.y) p_size.y = p.vertex[k].y;
    if (p.vertex[k].z > p_size.z) p_size.z = p.vertex[k].z;
    sum = add(sum, p.vertex[k]);
  }
  p.COM = scale(1.0f / p.num_vertex, sum);
  float vol = 0.0f;
  if (!r.label_i(p.num_faces)) return false;
  if (p.num_faces < 1 || p.num_faces > 32) return e.set("a polyhedron holds 1 to 32 faces");
  const bdem_float3* V = p.vertex;
  for (int j = 0; j < p.num_faces; ++j) {
    bdem_Surface& f = p.face[j];
    bdem_float3 fsum = mk(0, 0, 0);
    if (format == 0) {
      int idx;
      if (!r.i32(idx)) return false;
      r.word();
      if (!r.i32(f.num_vertex)) return false;
    } else if (!r.i32(f.num_vertex)) {
      return false;
    }
    if (f.num_vertex < 3 || f.num_vertex > 8) return e.set("a polyhedron face holds 3 to 8 vertices");
    for (int q = 0; q < f.num_vertex; ++q) {
      int o;
      if (!r.i32(o)) return false;
      if (format == 1 && q < f.num_vertex - 1) r.word();   /* the "," after each index */
      if (o < 1 || o > p.num_vertex) return e.set("face vertex index out of range");
      f.vertex_Order[q] = o - 1;
      fsum = add(fsum, V[o - 1]);
    }
    const int* O = f.vertex_Order;
    f.centroid = scale(1.0f / f.num_vertex, fsum);
    if (f.num_vertex == 3) {
      bdem_float3 V1 = sub(V[O[1]], V[O[2]]), V2 = sub(V[O[2]], V[O[0]]);
      bdem_float3 n = cross(V1, V2);
      f.normal = divs(n, length(n));
      f.area = 0.5f * length(n);
      bdem_float3 AD = sub(V[O[0]], p.COM), BD = sub(V[O[1]], p.COM), CD = sub(V[O[2]], p.COM);
      vol += std::sqrt(dot(AD, cross(BD, CD))) / 6.0000f;   /* sqrt of the triple product, as written */
    } else {
      const int nv = f.num_vertex;
      bdem_float3 V1 = sub(V[O[2]], V[O[3]]), V2 = sub(V[O[3]], V[O[0]]);
      bdem_float3 n = cross(V1, V2);
      f.normal = divs(n, length(n));
      V1 = sub(V[O[nv - 1]], f.centroid);
      V2 = sub(V[O[0]], f.centroid);
      float tot = 0.50f * length(cross(V1, V2));
      bdem_float3 AD = sub(V[O[nv - 1]], p.COM), BD = sub(V[O[0]], p.COM), CD = sub(f.centroid, p.COM);
      vol += dot(AD, cross(BD, CD)) / 6.0000f;
      for (int t = 0; t < nv - 1; ++t) {
        V1 = sub(V[O[t]], f.centroid);
        V2 = sub(V[O[t + 1]], f.centroid);
        AD = sub(V[O[t]], p.COM);
        BD = sub(V[O[t + 1]], p.COM);
        CD = sub(f.centroid, p.COM);
        vol += dot(AD, cross(BD, CD)) / 6.0000f;
        tot += 0.50f * length(cross(V1, V2));
      }
      f.area = tot;
    }
  }
  if (!r.label_i(p.num_edges)) return false;
  if (p.num_edges < 0 || p.num_edges > 32) return e.set("a polyhedron holds at most 32 edges");
  for (int k = 0; k < p.num_edges; ++k) {
    int a, b;
    if (!r.i32(a) || !r.i32(b)) return false;
    if (a < 1 || a > p.num_vertex || b < 1 || b > p.num_vertex) return e.set("edge vertex index out of range");
    p.edge[k].point_vindex[0] = (unsigned)(a - 1);
    p.edge[k].point_vindex[1] = (unsigned)(b - 1);
  }
  r.skip(3);
  float density;
  if (!r.label_f(density)) return false;
  p.mass = (float)(density * vol * 1E-6);
  r.word();
  for (int g = 0; g < 9; ++g)
    if (!r.f32(p.InertiaT[g])) return false;
  for (int g = 0; g < 9; ++g)
    if (p.InertiaT[g] > 0.0f) p.InertiaT[g] = (float)(1.0 / p.InertiaT[g]);
  /* Make_Particle_LocalInfo (:1186-1222): vertices about the COM, bounding radius */
  float dis = 0.0f;
  for (int k = 0; k < p.num_vertex; ++k) {
    p.Local_Vertex[k] = sub(p.vertex[k], p.COM);
    if (length(p.Local_Vertex[k]) > dis) dis = length(p.Local_Vertex[k]);
  }
  p.radius = dis;
  r.skip(3);
  return read_material(r, force_spec, i, p, m);
}

/* faces of a lifter from its vertices (ReadDObject :1699-1734 and the
 * rotations of Input/Processing_Mills.h:20-45, which always take the
 * quadrilateral form) */
void dobj_quad_face(bdem_Surface& f, const bdem_float3* V) {
  const int* O = f.vertex_Order;
  bdem_float3 V1 = sub(V[O[1]], V[O[0]]), V2 = sub(V[O[2]], V[O[1]]);
  bdem_float3 V3 = sub(V[O[3]], V[O[2]]), V4 = sub(V[O[0]], V[O[3]]);
  bdem_float3 n = cross(V2, V1);
  f.normal = divs(n, length(n));
  f.area = 0.500000f * length(n) + 0.500000f * length(cross(V3, V4));
}

void dobj_centroid(bdem_Surface& f, const bdem_float3* V) {
  bdem_float3 sum = mk(0, 0, 0);
  for (int q = 0; q < f.num_vertex; ++q) sum = add(sum, V[f.vertex_Order[q]]);
  f.centroid = scale(1.0f / f.num_vertex, sum);
}

/* ---- .DOBJ (ReadDObject, :1648-1750) ---- */
bool read_dobj(Reader& r, bdem_DOBJ& d, Err& e) {
  r.skip(4);
  int bt;
  if (!r.label_i(bt)) return false;
  d.btype = (unsigned)bt;
  if (bt == 1 && !r.label_v3(d.Axis)) return false;
  if (!r.label_v3(d.COM) || !r.label_f(d.boundR) || !r.label_i(d.num_vertex)) return false;
  if (d.num_vertex < 0 || d.num_vertex > 8) return e.set("a lifter holds at most 8 vertices");
  if (d.num_vertex == 0) return r.ok();
  for (int k = 0; k < d.num_vertex; ++k) {
    int idx;
    if (!r.i32(idx) || !r.v3(d.vertex[k])) return false;
  }
  if (!r.label_i(d.num_faces)) return false;
  if (d.num_faces < 0 || d.num_faces > 3) return e.set("a lifter holds at most 3 faces");
  for (int j = 0; j < d.num_faces; ++j) {
    bdem_Surface& f = d.faces[j];
    int idx;
    if (!r.i32(idx)) return false;
    r.word();
    /* "AreaTol: t"; the three unit-test volume objects leave it out (the
     * reference then reads the next label's number as the tolerance): 0.10,
     * the value the ball-mill lifter gives */
    if (r.peek() == "NUM_VERTEX:") {
      f.Tol = 0.10f;
    } else {
      r.word();
      if (!r.f32(f.Tol)) return false;
    }
    if (!r.label_i(f.num_vertex)) return false;
    if (f.num_vertex < 3 || f.num_vertex > 4) return e.set("a lifter face holds 3 or 4 vertices");
    for (int q = 0; q < f.num_vertex; ++q) {
      int o;
      if (!r.i32(o)) return false;
      if (o < 1 || o > d.num_vertex) return e.set("lifter face vertex index out of range");
      f.vertex_Order[q] = o - 1;
    }
    dobj_centroid(f, d.vertex);
    const int* O = f.vertex_Order;
    if (f.num_vertex == 3) {
      bdem_float3 V1 = sub(d.vertex[O[2]], d.vertex[O[0]]), V2 = sub(d.vertex[O[1]], d.vertex[O[2]]);
      bdem_float3 n = cross(V1, V2);
      f.normal = divs(n, length(n));
      f.area = 0.5f * length(n);
    } else {
      dobj_quad_face(f, d.vertex);
    }
  }
  if (!r.label_i(d.num_edges)) return false;
  if (d.num_edges < 0 || d.num_edges > 4) return e.set("a lifter holds at most 4 edges");
  for (int k = 0; k < d.num_edges; ++k) {
    int a, b;
    if (!r.i32(a) || !r.i32(b)) return false;
    if (a < 1 || a > d.num_vertex || b < 1 || b > d.num_vertex) return e.set("lifter edge index out of range");
    d.edge[k].point_vindex[0] = (unsigned)(a - 1);
    d.edge[k].point_vindex[1] = (unsigned)(b - 1);
  }
  return r.ok();
}

/* x' = cos a x + sin a y, y' = -sin a x + cos a y about `about`, with a
 * double angle and float coordinates (cos/sin of a double; the products and
 * the sum in double, rounded once to float), as the host rotations do. */
bdem_float3 turn(bdem_float3 v, bdem_float3 about, double c, double s) {
  bdem_float3 p = sub(v, about);
  /* x and y replaced, z kept, then the whole point moved back by `about` */
  bdem_float3 q = mk((float)(c * p.x + s * p.y), (float)(-s * p.x + c * p.y), v.z);
  return add(q, about);
}

/* Make_New_Lifter (:1913-1960): a copy of list[index] turned by `angle`
 * (a float angle, cos/sin of it in float as the float overloads give). */
void make_new_lifter(std::vector<bdem_DOBJ>& L, bdem_float3 about, int index, int d_index, float angle) {
  L[d_index] = L[index];
  bdem_DOBJ& d = L[d_index];
  const float c = std::cos(angle), s = std::sin(angle);
  for (int k = 0; k < d.num_vertex; ++k) {
    bdem_float3 p = sub(d.vertex[k], about);
    d.vertex[k] = add(mk(c * p.x + s * p.y, -s * p.x + c * p.y, d.vertex[k].z), about);
  }
  for (int j = 0; j < d.num_faces; ++j) {
    dobj_centroid(d.faces[j], d.vertex);
    dobj_quad_face(d.faces[j], d.vertex);
  }
  bdem_float3 p = sub(d.COM, about);
  d.COM = add(mk(c * p.x + s * p.y, -s * p.x + c * p.y, d.COM.z), about);
}

/* Repeat_Lifters (:1962-2028) */
void repeat_lifters(std::vector<bdem_DOBJ>& L, int types, int per_type, float R) {
  const bdem_float3 about = mk(R, R, 0.0f);
  const float theta = (float)(4.0 * std::asin(1.0f) / (float)per_type);
  int off = 0;
  if (types == 2) {
    make_new_lifter(L, about, 1, 2, 0.50f * theta);
    L[1] = L[2];
    off = 1;
  }
  int dindex = 1 + off;
  int ee = 0;
  for (int i = 0; i < types; ++i) {
    if (i == 0) ee = 1;
    for (int j = 1; j < per_type - 1 + ee; ++j) {
      make_new_lifter(L, about, i, dindex, theta * j);
      dindex += types == 2 ? 2 : 1;
    }
    dindex = 3;
  }
}

/* ---- .World (KSimulationData constructor, :27-656) ---- */
bool read_world(bdem_deck& D, const std::string& world_name, Err& e) {
  bdem_deck_info& I = D.info;
  Reader r;
  if (!r.open(D.dir + "/World/" + world_name + ".World")) return e.set(r.error());
  r.word();
  copy_str(I.title, sizeof(I.title), r.word());
  float dt = 0.0f;
  if (!r.label_i(I.force_spec) || !r.label_i(I.sim_type) || !r.label_i(I.particle_type)) return e.set(r.error());
  if (starts_with(r.peek(), "Delta_t")) {
    if (!r.label_f(dt)) return e.set(r.error());
  } else {
    I.dt_in_setup = 1;
  }
  if (I.force_spec != 0 && I.force_spec != 1) return e.set("ModelType must be 0 (TC) or 1 (KN)");
  if (I.sim_type < 0 || I.sim_type > 2) return e.set("Simulation_Type must be 0, 1 or 2");
  if (I.particle_type != 0 && I.particle_type != 1) return e.set("Particle_Type must be 0 or 1");
  r.skip(5);
  /* 1.1 world objects */
  int nw;
  if (!r.label_i(nw)) return e.set(r.error());
  if (nw < 1 || nw > MAX_WOBJ) return e.set("the number of world objects must be 1 to 8");
  D.wobj.assign(nw, bdem_WOBJ{});
  std::vector<std::string> wnames(nw);
  for (int i = 0; i < nw; ++i) {
    int idx;
    if (!r.i32(idx)) return e.set(r.error());
    wnames[i] = r.word();
    bdem_WOBJ& w = D.wobj[i];
    if (idx == 99) {
      w.is_rotating = true;
      bdem_float3& ax = w.cyl_geo.AXIS;
      if (!r.label_v3(ax) || !r.label_f(I.mill_rpm) || !r.label_i(I.total_revs)) return e.set(r.error());
      w.cyl_geo.radius = std::max(std::max(ax.x, ax.y), ax.z);
      ax = divs(ax, length(ax));
    }
    Reader wr;
    if (!wr.open(D.dir + "/SurfaceObjects/" + wnames[i] + ".WOBJ")) return e.set(wr.error());
    if (!read_wobj(wr, w, e)) return e.set(wr.error());
  }
  /* 1.2 particle objects */
  r.skip(3);
  int np;
  if (!r.label_i(np)) return e.set(r.error());
  if (np < 1 || np > MAX_POBJ) return e.set("the number of particle objects must be 1 to 6");
  I.num_pobj = np;
  D.pobj.assign(np, bdem_POBJ{});
  std::vector<Material> mats(np);
  bdem_float3 p_size[MAX_POBJ] = {};
  for (int i = 0; i < np; ++i) {
    int idx;
    if (!r.i32(idx)) return e.set(r.error());
    std::string name = r.word();
    if (!r.i32(D.per_object[i])) return e.set(r.error());
    Reader pr;
    if (!pr.open(D.dir + "/ParticleObjects/" + name + ".POBJ")) return e.set(pr.error());
    const bool ok = I.particle_type == 0
                        ? read_pobj_sphere(pr, I.force_spec, i, D.pobj[i], mats[i])
                        : read_pobj_poly(pr, I.force_spec, i, D.pobj[i], mats[i], p_size[i], e);
    if (!ok) return e.set(pr.error());
  }
  /* 1.3 dynamic objects (lifters or volume objects) */
  r.skip(3);
  int nd, repeat;
  if (!r.label_i(nd)) return e.set(r.error());
  r.word();
  if (!r.i32(repeat)) return e.set(r.error());
  I.dobj_types = nd;
  I.dobj_repeat = repeat;
  if (nd < 0 || nd > MAX_DOBJ || (repeat > 0 && nd * repeat > MAX_DOBJ))
    return e.set("at most 32 dynamic objects after repetition");
  if (repeat > 0 && nd > 2) return e.set("lifter repetition takes one or two lifter types");
  if (nd > 0) {
    D.dobj.assign(repeat > 0 ? nd * repeat : nd, bdem_DOBJ{});
    for (int i = 0; i < nd; ++i) {
      int idx;
      if (!r.i32(idx)) return e.set(r.error());
      std::string name = r.word();
      bdem_DOBJ& d = D.dobj[i];
      d.is_attached_DWorld = !(idx == 99 || idx == 98);
      d.is_translating = idx == 98;
      if (idx == 98 && !r.label_v3(d.velocity)) return e.set(r.error());
      Reader dr;
      const char* folder = I.sim_type == BDEM_SIM_BALLMILL ? "/LifterObjects/" : "/VolumeObjects/";
      if (!dr.open(D.dir + folder + name + ".DOBJ")) return e.set(dr.error());
      if (!read_dobj(dr, d, e)) return e.set(dr.error());
    }
    if (repeat > 0) repeat_lifters(D.dobj, nd, repeat, D.wobj[0].cyl_geo.radius);
  }
  I.num_dobj = (int)D.dobj.size();
  /* 2 simulation setup */
  r.skip(3);
  int ntot;
  bdem_float3 gravity;
  if (!r.label_i(ntot)) return e.set(r.error());
  if (I.dt_in_setup) {
    if (!starts_with(r.peek(), "Delta_t")) return e.set("no Delta_t after Particle_Type or Num_Particles");
    if (!r.label_f(dt)) return e.set(r.error());
  }
  if (!r.label_v3(gravity)) return e.set(r.error());
  r.word();
  int sum = 0;
  for (int i = 0; i < np; ++i) sum += D.per_object[i];
  if (sum != ntot) return e.set("Num_Particles differs from the sum of the particle object counts");
  if (ntot < 1) return e.set("Num_Particles must be positive");
  if (dt > 1E-1f || dt < 0.0f) return e.set("Delta_t must be in [0, 0.1]");
  int rotation;
  float roll_res, damp, vlimit;
  if (!r.label_i(rotation) || !r.label_f(roll_res) || !r.label_f(damp) || !r.label_f(vlimit))
    return e.set(r.error());
  if (!(roll_res >= 0.0f)) return e.set("RollRes must be >= 0");
  /* 3 grid */
  r.skip(3);
  bdem_float3 origin, world, cell;
  if (!r.label_v3(origin) || !r.label_v3(world) || !r.label_v3(cell)) return e.set(r.error());
  /* 4 initial particles */
  r.skip(3);
  bdem_InitConfig& G = D.init;
  if (!r.label_i(I.read_file)) return e.set(r.error());
  if (I.read_file == 1) {
    r.word();
    copy_str(I.start_file, sizeof(I.start_file), r.word());
    G.grid_type = 0;
  } else {
    if (!r.label_i(G.grid_type)) return e.set(r.error());
    if (G.grid_type < 2) {
      bdem_int3& n3 = G.num;
      r.word();
      if (!r.i32(n3.x) || !r.i32(n3.y) || !r.i32(n3.z)) return e.set(r.error());
      if (!r.label_v3(G.start) || !r.label_v3(G.space) || !r.label_v3(G.velocity[0])) return e.set(r.error());
    } else {
      if (!r.label_v3(G.launch_Vel)) return e.set(r.error());
      if (I.sim_type != BDEM_SIM_BALLMILL &&
          (!r.label_v3(G.fill_plane_start) || !r.label_v3(G.fill_plane_end)))
        return e.set(r.error());
      if (mats[0].cor > 0.35) return e.set("a fill needs COR <= 0.35");
      if (ntot % 2 != 0) return e.set("a fill needs an even particle count");
    }
  }
  /* 5 output options */
  r.skip(3);
  int disp;
  if (!r.label_i(disp)) return e.set(r.error());
  I.render = disp > 0;
  if (I.render) {
    if (!r.i32(I.color_type)) return e.set(r.error());
    if (I.color_type > 0 && !r.f32(I.color_bin)) return e.set(r.error());
  }
  if (!r.label_i(I.ortho) || !r.label_i(I.fps) || !r.label_i(I.gl_debug) || !r.label_i(I.write_steps))
    return e.set(r.error());
  if (I.write_steps > 0) {
    /* the format follows the count; SPH_90.World omits it (the reference's
     * stream fails there and loses the rest of the file): detail 3 */
    const std::string t = r.peek();
    if (!t.empty() && (std::isdigit((unsigned char)t[0]) || t[0] == '-' || t[0] == '+')) {
      if (!r.i32(I.output_format)) return e.set(r.error());
    } else {
      I.output_format = 3;
    }
  }
  int lim;
  if (!r.label_f(I.total_time) || !r.label_i(lim)) return e.set(r.error());
  I.is_run_limit = lim == 1;
  if (I.is_run_limit && !r.f32(I.run_time)) return e.set(r.error());
  if (!r.label_i(I.energy_steps) || !r.label_i(I.wall_steps)) return e.set(r.error());
  r.word();
  /* 6 simulation specific */
  if (I.sim_type == BDEM_SIM_BALLMILL) {
    r.skip(2);
    if (!r.label_i(I.power_revs) || !r.label_i(I.total_revs) || !r.label_f(I.mill_rpm)) return e.set(r.error());
    r.word();
  } else if (I.sim_type == BDEM_SIM_SILO) {
    r.skip(2);
    if (!r.label_i(I.manual_hatch)) return e.set(r.error());
    I.manual_hatch = I.manual_hatch != 0;
    if (!r.label_i(I.measure_flow) || !r.i32(I.flow_steps)) return e.set(r.error());
    I.measure_flow = I.measure_flow != 0;
    int kill;
    if (!r.label_i(kill)) return e.set(r.error());
    if (kill != 0 && kill != 1) return e.set("kill_after must be 0 or 1");
    I.kill_particle = kill;
  }
  if (!r.ok()) return e.set(r.error());

  /* spring constants: per type, then the pair matrix (:276-292) */
  double cor[MAX_POBJ], cor_s[MAX_POBJ], cor_d[MAX_POBJ], tc_ps[MAX_POBJ], tc_pd[MAX_POBJ];
  for (int i = 0; i < np; ++i) {
    if (I.force_spec == 0) {
      mats[i].tc_pp *= dt;
      mats[i].tc_ps *= dt;
      mats[i].tc_pd *= dt;
    }
    cor[i] = mats[i].cor;
    cor_s[i] = mats[i].cor_s;
    cor_d[i] = mats[i].cor_d;
    tc_ps[i] = mats[i].tc_ps;
    tc_pd[i] = mats[i].tc_pd;
  }
  const int mode = I.force_spec == 0 ? BDEM_SPRING_TC : BDEM_SPRING_KN;
  if (bdem::derive_springs(D.pobj.data(), np, mode, cor, cor_s, cor_d, mats[np - 1].tc_pp, tc_ps, tc_pd) !=
      BDEM_OK)
    return e.set("spring constants: every type needs a positive mass (and contact times in TC mode)");

  /* ---- Set_SimInfo (main.cpp:469-833) ---- */
  bdem_SimulationInfo& S = D.sim;
  S.Num_Particles = ntot;
  S.Num_WorldObjects = nw;
  S.Num_ParticleObjects = np;
  S.Num_ParticlesPerObject = D.per_object;
  S.particle_type = I.particle_type;
  S.InitalDelta_t = dt;
  S.Force_Field = gravity;
  S.Rotation = rotation != 0;
  S.Roll_Res = roll_res;
  S.Global_VisDamp = damp;
  S.Rotating_WObject = D.wobj[0].is_rotating;
  for (int i = 0; i < np; ++i) {
    bdem_POBJ& p = D.pobj[i];
    G.p_size[i] = I.particle_type == 1 ? p_size[i] : mk(p.radius * 2.0f, p.radius * 2.0f, p.radius * 2.0f);
    const float inv = 1.0f / (0.40f * p.radius * p.radius * p.mass);
    if (I.particle_type == 0) {
      p.InertiaT[0] = inv;
    } else {
      for (int g = 0; g < 9; ++g) p.InertiaT[g] = inv * p.InertiaT[g];
    }
  }
  G.threads_perBlock = 1024;
  G.use_file = I.read_file == 1;
  S.cellSize = cell;
  S.worldOrigin = origin;
  const bdem_float3 total = sub(world, origin);
  S.num_NNCells.x = (int)std::ceil(total.x / cell.x);
  S.num_NNCells.y = (int)std::ceil(total.y / cell.y);
  S.num_NNCells.z = (int)std::ceil(total.z / cell.z);
  S.TotnumCells = S.num_NNCells.x * S.num_NNCells.y * S.num_NNCells.z;
  S.unit_test = std::strcmp(I.title, "UNIT_TEST") == 0 ? 1 : 0;
  S.Vel_Limit = vlimit;
  const float R0 = D.wobj[0].cyl_geo.radius;
  I.rotate_about = mk(R0, R0, 0.0f);
  S.Simulation_Type = I.sim_type;
  if (I.sim_type == BDEM_SIM_BALLMILL) {
    S.isMillSim = true;
    S.EnergyCalc = I.power_revs > 0;
    I.steps_per_rev = (int)((60.0f / I.mill_rpm) / dt);
    if (I.total_time <= 0.0f) I.total_time = (float)(I.total_revs * I.steps_per_rev * dt);
    I.rad_per_step = -((I.mill_rpm * 2.000f * PI_MAIN) / 60.000f) * dt;
    I.vel_mill = (float)std::fabs((I.rad_per_step / dt) * R0 * 0.01);
    S.MillRotVel = -I.vel_mill;
    if (!D.dobj.empty()) {
      S.DObject_CylinderBoundR = R0 - (2.0f * D.dobj[0].COM.y);
      S.cent = mk(D.wobj[0].cyl_geo.center_bot_cap.x, D.wobj[0].cyl_geo.center_bot_cap.y, 0.0f);
    }
  } else if (I.sim_type == BDEM_SIM_SILO) {
    S.isMillSim = false;
    I.hatch_height = D.wobj[nw - 1].vertex[0].y;
    S.Silo_Kill_Height = I.hatch_height - 5.0f;
    S.Kill_Particles = I.kill_particle != 0;
    if (I.total_time == 0.0f) {
      if (!I.measure_flow) return e.set("a silo run needs total_Time or a flow counter");
      I.total_time = 60.0f;
    }
  } else if (S.Rotating_WObject) {
    I.steps_per_rev = (int)((60.0f / I.mill_rpm) / dt);
    if (I.total_time <= 0.0f) I.total_time = (float)(I.total_revs * I.steps_per_rev * dt);
    I.rad_per_step = -((I.mill_rpm * 2.000f * PI_MAIN) / 60.000f) * dt;
    I.vel_mill = (float)std::fabs((I.rad_per_step / dt) * (R0 - 100.0f) * 0.01);
    S.MillRotVel = -I.vel_mill;
  }
  I.rotating_world = S.Rotating_WObject ? 1 : 0;
  S.Integration_Type = 0;
  S.use_symmetry = false;
  S.sphere_orient = false;
  if (S.Rotating_WObject) S.Kill_Particles = false;
  S.max_size = mk(S.num_NNCells.x * S.cellSize.x, S.num_NNCells.y * S.cellSize.y, S.num_NNCells.z * S.cellSize.z);
  S.use_hist = false;
  I.num_wobj = nw;
  for (int i = 0; i < np && i < 8; ++i) I.num_per_pobj[i] = D.per_object[i];
  D.dobj_initial = D.dobj;
  return true;
}

/* ---- state files ---- */
bool read_header(Reader& r, int sim_type, int rotating_world, bdem_state_header& h) {
  r.line();
  r.word();
  if (!r.i32(h.spec_type)) return false;
  if (sim_type == BDEM_SIM_BALLMILL) {
    if (!r.label_i(h.rev_num) || !r.label_d(h.time) || !r.label_i(h.rev_complete) || !r.label_d(h.rot_angle))
      return false;
    for (int k = 0; k < 5; ++k)
      if (!r.label_f(h.energy[k])) return false;
  } else if (sim_type == BDEM_SIM_SILO) {
    if (!r.label_d(h.time) || !r.label_i(h.num_rem)) return false;
  } else if (rotating_world) {
    if (!r.label_d(h.time) || !r.label_i(h.rot_steps) || !r.label_d(h.rot_angle)) return false;
  } else {
    if (!r.label_d(h.time)) return false;
  }
  return true;
}

/* Rotate_VolObject_angle (Input/Processing_Mills.h:92-160): vertices from
 * the lifters as read, turned by the absolute angle; the COM from its
 * current position. */
void rotate_lifters_to(bdem_deck& D, double angle) {
  const double c = std::cos(angle), s = std::sin(angle);
  for (size_t k = 0; k < D.dobj.size(); ++k) {
    bdem_DOBJ& d = D.dobj[k];
    if (!d.is_attached_DWorld) continue;
    for (int v = 0; v < d.num_vertex; ++v) d.vertex[v] = turn(D.dobj_initial[k].vertex[v], D.info.rotate_about, c, s);
    for (int j = 0; j < d.num_faces; ++j) {
      dobj_centroid(d.faces[j], d.vertex);
      dobj_quad_face(d.faces[j], d.vertex);
    }
    d.COM = turn(d.COM, D.info.rotate_about, c, s);
  }
}

/* rand() as the reference's host calls it, never seeded: the C library's
 * default generator with seed 1.  glibc's default (TYPE_3) is an additive
 * lagged-Fibonacci generator of degree 31 and separation 3, as published
 * in its documentation and source: r[0] = seed, r[i] = 16807 r[i-1]
 * mod (2^31 - 1) for i < 31, r[i] = r[i-31] for 31 <= i < 34, then
 * r[i] = r[i-31] + r[i-3] (mod 2^32), and the k-th output is
 * r[k + 344] >> 1. */
class DefaultRand {
 public:
  DefaultRand() {
    int32_t r[34];
    r[0] = 1;
    for (int i = 1; i < 31; ++i) {
      const int64_t hi = r[i - 1] / 127773, lo = r[i - 1] % 127773;
      int64_t w = 16807 * lo - 2836 * hi;
      if (w < 0) w += 2147483647;
      r[i] = (int32_t)w;
    }
    for (int i = 31; i < 34; ++i) r[i] = r[i - 31];
    for (int i = 0; i < 34; ++i) ring_[i] = (uint32_t)r[i];
    k_ = 34;
    for (int i = 34; i < 344; ++i) next_raw();
  }
  int operator()() { return (int)(next_raw() >> 1); }

 private:
  uint32_t next_raw() {
    const uint32_t v = ring_[(k_ - 31) % 34] + ring_[(k_ - 3) % 34];
    ring_[k_ % 34] = v;
    ++k_;
    return v;
  }
  uint32_t ring_[34];
  uint64_t k_;
};
constexpr float RAND_MAX_F = 2147483647.0f;   /* (float)RAND_MAX, as the reference divides by it */

/* make_quaterion(theta_deg, x, y, z) then normalise (DataStructures/Quart.h:34-80) */
bdem_Quaterion quat_deg(float theta, float x, float y, float z) {
  theta = ((2.0f * 3.14152f) / 360.0f) * theta;   /* the header's own pi */
  bdem_Quaterion q{std::cos(theta / 2.0f), std::sin(theta / 2.0f) * x, std::sin(theta / 2.0f) * y,
                   std::sin(theta / 2.0f) * z};
  const float sq = q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z;
  const float n = sq > 0.0f ? 1.0f / std::sqrt(sq) : 0.0f;
  return bdem_Quaterion{q.w * n, q.x * n, q.y * n, q.z * n};
}

/* Set_Position_DefaultGrid / Set_Position_Random_Grid
 * (Device/DeviceInterface.cu:192-362, 435-601): the lattice from Start with
 * pitch p_size[0] + Space, walked y outer, z middle, x inner. */
int grid_fill(const bdem_deck& D, bdem_float3* pos, bdem_Quaterion* ornt, unsigned* ptype, int* pid) {
  const bdem_InitConfig& G = D.init;
  const int n = D.sim.Num_Particles, NX = G.num.x, NY = G.num.y, NZ = G.num.z;
  const bool orient = D.info.particle_type == 1 || D.sim.sphere_orient;
  if (NX < 1 || NY < 1 || NZ < 1 || (long long)NX * NY * NZ < n) return BDEM_EINVAL;
  const int np = D.info.num_pobj;
  DefaultRand rnd;
  float x = G.start.x, y = G.start.y, z = G.start.z;
  int sel = -1, selcount[8] = {0, 0, 0, 0, 0, 0, 0, 0}, typed = 0;
  for (int yd = 0; yd < NY; ++yd) {
    for (int zd = 0; zd < NZ; ++zd) {
      for (int xd = 0; xd < NX; ++xd) {
        const long long k = (long long)yd * NZ * NX + (long long)zd * NX + xd;
        if (k < n) {
          pos[k] = mk(x, y, z);
          if (ornt) ornt[k] = bdem_Quaterion{1.0f, 0.0f, 0.0f, 0.0f};
          if (G.grid_type == 1) {
            if (orient) {
              const float rw = (float)(((float)rnd() / RAND_MAX_F) * 30.0);
              const float rx = (float)rnd() / RAND_MAX_F;
              const float ry = (float)rnd() / RAND_MAX_F;
              const float rz = (float)rnd() / RAND_MAX_F;
              if (ornt) ornt[k] = quat_deg(rw, rx, ry, rz);
            }
            /* types in turn, skipping a type whose count is used up */
            if (++sel >= np) sel = 0;
            int pty = sel;
            if (selcount[pty] >= D.per_object[pty])
              for (int i = 0; i < np; ++i)
                if (selcount[i] <= D.per_object[i] - 1) { pty = i; break; }
            selcount[pty]++;
            ptype[k] = (unsigned)pty;
            pid[k] = (int)k;
          }
        }
        x = x + G.p_size[0].x + G.space.x;
      }
      x = G.start.x;
      z = z + G.p_size[0].z + G.space.z;
    }
    z = G.start.z;
    y = y + G.p_size[0].y + G.space.y;
  }
  if (G.grid_type == 0) {
    for (int i = 0; i < np; ++i)
      for (int j = 0; j < D.per_object[i]; ++j, ++typed) {
        ptype[typed] = (unsigned)i;
        pid[typed] = typed;
      }
  } else if (n > 1) {
    /* the shuffle of types and IDs (:562-575) */
    for (int i = 0; i < n; ++i) {
      const int a = rnd() % (n - 1), b = rnd() % (n - 1);
      std::swap(ptype[a], ptype[b]);
      std::swap(pid[a], pid[b]);
    }
  }
  return BDEM_OK;
}

void fmt_g(std::string& out, double v) {
  char b[48];
  std::snprintf(b, sizeof(b), "%g", v);
  out += b;
}

}  // namespace

extern "C" {

int bdem_ksim_read(const char* path, char* project, char* world, int cap, int* mode2d, int* device,
                   int* num_devices) {
  if (!path || !project || !world || cap < 1 || !mode2d || !device || !num_devices) return BDEM_EINVAL;
  Reader r;
  if (!r.open(path)) return BDEM_EINVAL;
  r.word();
  std::string p = r.word();
  r.word();
  std::string w = r.word();
  r.word();
  if (!r.i32(*mode2d)) return BDEM_EINVAL;
  r.word(); r.word();
  if (!r.i32(*device)) return BDEM_EINVAL;
  r.word();
  if (!r.i32(*num_devices)) return BDEM_EINVAL;
  copy_str(project, (size_t)cap, p);
  copy_str(world, (size_t)cap, w);
  return r.ok() ? BDEM_OK : BDEM_EINVAL;
}

int bdem_deck_load(bdem_deck** out, const char* project_dir, const char* world, int mode2d, char* err,
                   int err_cap) {
  if (err && err_cap > 0) err[0] = 0;
  if (!out) return BDEM_EINVAL;
  *out = nullptr;
  if (!project_dir || !world) return BDEM_EINVAL;
  bdem_deck* D = new bdem_deck();
  std::memset(&D->info, 0, sizeof(D->info));
  std::memset(&D->sim, 0, sizeof(D->sim));
  std::memset(&D->init, 0, sizeof(D->init));
  std::memset(D->per_object, 0, sizeof(D->per_object));
  D->dir = project_dir;
  Err e;
  if (!read_world(*D, world, e)) {
    if (err && err_cap > 0) copy_str(err, (size_t)err_cap, e.msg.empty() ? "deck error" : e.msg);
    delete D;
    return BDEM_EINVAL;
  }
  D->sim.Mode2D = mode2d == 1;
  *out = D;
  return BDEM_OK;
}

void bdem_deck_free(bdem_deck* deck) { delete deck; }

int bdem_deck_get(const bdem_deck* D, const bdem_deck_info** info, const bdem_SimulationInfo** sim,
                  const bdem_InitConfig** init, const bdem_WOBJ** wobj, const bdem_POBJ** pobj,
                  const bdem_DOBJ** dobj) {
  if (!D) return BDEM_EINVAL;
  if (info) *info = &D->info;
  if (sim) *sim = &D->sim;
  if (init) *init = &D->init;
  if (wobj) *wobj = D->wobj.data();
  if (pobj) *pobj = D->pobj.data();
  if (dobj) *dobj = D->dobj.empty() ? nullptr : D->dobj.data();
  return BDEM_OK;
}

int bdem_fill_random(int n, const float* radius, bdem_float3 lo, bdem_float3 hi, const float* cyl, float gap,
                     unsigned long long seed, int max_tries, bdem_float3* pos, int* failed) {
  if (failed) *failed = -1;
  if (n < 0 || (n > 0 && (!radius || !pos)) || max_tries < 1 || !(gap >= 0.0f)) return BDEM_EINVAL;
  if (n == 0) return BDEM_OK;
  float rmax = 0.0f;
  for (int k = 0; k < n; ++k) {
    if (!(radius[k] > 0.0f)) return BDEM_EINVAL;
    rmax = std::max(rmax, radius[k]);
  }
  /* a uniform grid of cells no smaller than the largest contact distance:
   * every sphere that can touch a candidate is in its 27 cells */
  const double cell = 2.0 * rmax + gap;
  const double span[3] = {double(hi.x) - lo.x, double(hi.y) - lo.y, double(hi.z) - lo.z};
  int dim[3];
  for (int a = 0; a < 3; ++a) {
    if (!(span[a] > 0.0)) return BDEM_EINVAL;
    dim[a] = std::max(1, (int)std::ceil(span[a] / cell));
  }
  const long long ncell = (long long)dim[0] * dim[1] * dim[2];
  if (ncell > (1ll << 31)) return BDEM_EINVAL;
  std::vector<int> head((size_t)ncell, -1), next((size_t)n, -1);
  unsigned long long state = seed;
  auto draw = [&]() {
    state += 0x9E3779B97F4A7C15ull;
    unsigned long long z = state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return double(z >> 40) * (1.0 / 16777216.0);
  };
  auto cell_of = [&](double v, int a, double o) {
    int c = (int)std::floor((v - o) / cell);
    return std::min(std::max(c, 0), dim[a] - 1);
  };
  for (int k = 0; k < n; ++k) {
    const double r = radius[k];
    const double a0 = lo.x + r, a1 = hi.x - r, b0 = lo.y + r, b1 = hi.y - r, c0 = lo.z + r, c1 = hi.z - r;
    if (a1 < a0 || b1 < b0 || c1 < c0) { if (failed) *failed = k; return BDEM_EINVAL; }
    bool placed = false;
    for (int t = 0; t < max_tries && !placed; ++t) {
      const double x = a0 + (a1 - a0) * draw(), y = b0 + (b1 - b0) * draw(), z = c0 + (c1 - c0) * draw();
      const float fx = float(x), fy = float(y), fz = float(z);
      if (cyl) {
        const double dx = double(fx) - cyl[0], dz = double(fz) - cyl[1], lim = double(cyl[2]) - r;
        if (lim < 0.0 || dx * dx + dz * dz > lim * lim) continue;
      }
      const int gx = cell_of(fx, 0, lo.x), gy = cell_of(fy, 1, lo.y), gz = cell_of(fz, 2, lo.z);
      bool clear = true;
      for (int dy = -1; dy <= 1 && clear; ++dy)
        for (int dz = -1; dz <= 1 && clear; ++dz)
          for (int dx = -1; dx <= 1 && clear; ++dx) {
            const int X = gx + dx, Y = gy + dy, Z = gz + dz;
            if (X < 0 || Y < 0 || Z < 0 || X >= dim[0] || Y >= dim[1] || Z >= dim[2]) continue;
            for (int j = head[((size_t)Y * dim[2] + Z) * dim[0] + X]; j >= 0; j = next[j]) {
              const double ex = double(pos[j].x) - fx, ey = double(pos[j].y) - fy, ez = double(pos[j].z) - fz;
              const double need = double(radius[j]) + r + gap;
              if (ex * ex + ey * ey + ez * ez < need * need) { clear = false; break; }
            }
          }
      if (!clear) continue;
      pos[k] = bdem_float3{fx, fy, fz};
      const size_t c = ((size_t)gy * dim[2] + gz) * dim[0] + gx;
      next[k] = head[c];
      head[c] = k;
      placed = true;
    }
    if (!placed) { if (failed) *failed = k; return BDEM_EINVAL; }
  }
  return BDEM_OK;
}

int bdem_deck_setup(const bdem_deck* D, bdem_ctx* ctx) {
  if (!D || !ctx) return BDEM_EINVAL;
  return bdem_set_simulation_data(ctx, D->wobj.data(), D->pobj.data(), D->dobj.empty() ? nullptr : D->dobj.data(),
                                  &D->sim, (int)D->wobj.size(), (int)D->pobj.size(), (int)D->dobj.size(),
                                  &D->init);
}

int bdem_state_read(const char* path, int sim_type, int rotating_world, int particle_type, int n,
                    bdem_state_header* hdr, bdem_float3* pos, bdem_Quaterion* ornt, bdem_float3* vel,
                    bdem_float3* avel, bdem_float3* acc, unsigned int* ptype, int* pid, int* rows) {
  if (!path || !hdr || !pos || !ptype || !pid || n < 0) return BDEM_EINVAL;
  std::memset(hdr, 0, sizeof(*hdr));
  if (rows) *rows = 0;
  Reader r;
  if (!r.open(path)) return BDEM_EINVAL;
  if (!read_header(r, sim_type, rotating_world, *hdr)) return BDEM_EINVAL;
  const int st = hdr->spec_type;
  if (st < 1 || st > 4) return BDEM_EINVAL;
  bdem_float3 dv;
  bdem_Quaterion dq;
  float dm;
  int k = 0;
  for (; k < n; ++k) {
    if (r.peek().empty()) break;
    int type_i;
    if (!r.i32(pid[k]) || !r.i32(type_i) || !r.v3(pos[k])) return BDEM_EINVAL;
    ptype[k] = (unsigned)type_i;
    if (st >= 2 && !r.v3(vel ? vel[k] : dv)) return BDEM_EINVAL;
    if (st == 4 && !r.v3(acc ? acc[k] : dv)) return BDEM_EINVAL;
    if (particle_type == 1) {
      bdem_Quaterion& q = ornt ? ornt[k] : dq;
      if (!r.f32(q.w) || !r.f32(q.x) || !r.f32(q.y) || !r.f32(q.z)) return BDEM_EINVAL;
    }
    if (st >= 2 && !r.v3(avel ? avel[k] : dv)) return BDEM_EINVAL;
    if (st == 4 && !(r.f32(dm) && r.f32(dm) && r.f32(dm))) return BDEM_EINVAL;
    if (st < 4 && acc) acc[k] = mk(0, 0, 0);
  }
  if (rows) *rows = k;
  return BDEM_OK;
}

int bdem_state_write(const char* path, int sim_type, int rotating_world, int particle_type, int n,
                     const bdem_state_header* hdr, const bdem_float3* pos, const bdem_Quaterion* ornt,
                     const bdem_float3* vel, const bdem_float3* avel, const bdem_float3* acc,
                     const unsigned int* ptype, const int* pid, const bdem_POBJ* pobj, int num_pobj) {
  if (!path || !hdr || !pos || !ptype || !pid || n < 0) return BDEM_EINVAL;
  const int st = hdr->spec_type;
  if (st < 1 || st > 5) return BDEM_EINVAL;
  if (st >= 2 && st <= 4 && (!vel || !avel)) return BDEM_EINVAL;
  if (st == 4 && !acc) return BDEM_EINVAL;
  if (particle_type == 1 && st <= 4 && !ornt) return BDEM_EINVAL;
  if (st == 5 && (!pobj || num_pobj < 1)) return BDEM_EINVAL;
  const bool poly = particle_type == 1;
  std::string o;
  switch (st) {
    case 1: o += poly ? "#  P_ID    PType_ID   Pos(x,y,z)cm  Orient(w,x,y,z)\n" : "#  P_ID    PType_ID   Pos(x,y,z)cm  \n"; break;
    case 2: o += poly ? "#  P_ID   PType_ID   Pos(x,y,z)cm    Vel(x,y,z)cm/s  Orient(w,x,y,z) AngVel(x,y,z)cm/s\n"
                      : "#  P_ID   PType_ID   Pos(x,y,z)cm    Vel(x,y,z)cm/s    AngVel(x,y,z)cm/s\n"; break;
    case 3: o += poly ? "#  P_ID   PType_ID   Pos(x,y,z)cm  Vel(x,y,z)cm/s    Orient(w,x,y,z)  AngVel(x,y,z)cm/s\n"
                      : "#  P_ID   PType_ID   Pos(x,y,z)cm  Vel(x,y,z)cm/s     AngVel(x,y,z)cm/s\n"; break;
    case 4: o += poly ? "#  P_ID   PType_ID   Pos(x,y,z)cm  Vel(x,y,z)cm/s  Acc(x,y,z)  Orient(w,x,y,z)  AngVel(x,y,z)cm/s  | MagVel(m/s)  MagAVel(m/s) MagAcc(m/s)\n"
                      : "#  P_ID   PType_ID   Pos(x,y,z)cm  Vel(x,y,z)cm/s  Acc(x,y,z)  AngVel(x,y,z)cm/s   | MagVel(m/s)  MagAVel(m/s) MagAcc(m/s)\n"; break;
    default: o += "# COM   Vertex\n"; break;
  }
  o += "Specfication_Detail: " + std::to_string(st) + "\n";
  if (sim_type == BDEM_SIM_BALLMILL) {
    o += " RevNum: " + std::to_string(hdr->rev_num) + "  Time: ";
    fmt_g(o, hdr->time);
    o += " RevComplete: " + std::to_string(hdr->rev_complete) + "  RotAngle: ";
    fmt_g(o, hdr->rot_angle);
    const char* lab[5] = {"  EParticle: ", " ELift ", " EDrum: ", " ENorm: ", " EShear: "};
    for (int k = 0; k < 5; ++k) { o += lab[k]; fmt_g(o, hdr->energy[k]); }
    o += "\n";
  } else if (sim_type == BDEM_SIM_SILO) {
    o += "  Time: "; fmt_g(o, hdr->time);
    o += " NumRem: " + std::to_string(hdr->num_rem) + "\n";
  } else if (rotating_world) {
    o += "  Time: "; fmt_g(o, hdr->time);
    o += "  RotSteps: " + std::to_string(hdr->rot_steps) + "  RotAngle: ";
    fmt_g(o, hdr->rot_angle);
    o += "\n";
  } else {
    o += "  Time: "; fmt_g(o, hdr->time); o += "\n";
  }
  auto v3 = [&](const char* sep0, const bdem_float3& v, const char* sep) {
    o += sep0; fmt_g(o, v.x); o += sep; fmt_g(o, v.y); o += sep; fmt_g(o, v.z);
  };
  for (int i = 0; i < n; ++i) {
    if (pid[i] == -1) continue;
    o += std::to_string(pid[i]) + "  " + std::to_string(ptype[i]) + "    ";
    fmt_g(o, pos[i].x); o += st == 1 || st == 2 ? "  " : " ";
    fmt_g(o, pos[i].y); o += st == 1 || st == 2 ? "  " : " ";
    fmt_g(o, pos[i].z);
    if (st == 1) {
      o += "    ";
      if (poly) {
        fmt_g(o, ornt[i].w); o += "  "; fmt_g(o, ornt[i].x); o += "  ";
        fmt_g(o, ornt[i].y); o += "  "; fmt_g(o, ornt[i].z);
      }
      o += "\n";
      continue;
    }
    if (st == 5) {
      o += "      ";
      const unsigned t = ptype[i] < (unsigned)num_pobj ? ptype[i] : 0u;
      for (int v = 0; v < pobj[t].num_vertex; ++v) {
        bdem_float3 p = add(pobj[t].vertex[v], pos[i]);
        fmt_g(o, p.x); o += "  "; fmt_g(o, p.y); o += "  "; fmt_g(o, p.z); o += "    ";
      }
      o += "\n";
      continue;
    }
    v3(st == 2 ? "    " : (st == 3 ? "       " : "             "), vel[i], st == 2 ? "  " : "   ");
    if (st == 4) v3("       ", acc[i], "   ");
    if (poly) {
      o += st == 2 ? "    " : "       ";
      fmt_g(o, ornt[i].w); o += "   "; fmt_g(o, ornt[i].x); o += "   ";
      fmt_g(o, ornt[i].y); o += " "; fmt_g(o, ornt[i].z);
    }
    v3("       ", avel[i], "   ");
    if (st == 4) {
      o += "       "; fmt_g(o, length(scale(0.010f, vel[i])));
      o += "   "; fmt_g(o, length(scale(0.010f, avel[i])));
      o += "   "; fmt_g(o, length(scale(0.010f, acc[i])));
    }
    o += "\n";
  }
  FILE* f = std::fopen(path, "wb");
  if (!f) return BDEM_EINVAL;
  const bool ok = std::fwrite(o.data(), 1, o.size(), f) == o.size();
  return (std::fclose(f) == 0 && ok) ? BDEM_OK : BDEM_EINVAL;
}

int bdem_deck_initial_state(const bdem_deck* D, bdem_float3* pos, bdem_Quaterion* ornt, unsigned int* ptype,
                            int* pid) {
  if (!D || !pos || !ptype || !pid) return BDEM_EINVAL;
  if (D->info.read_file == 1 || D->init.grid_type > 1) return BDEM_ENOTSUP;
  return grid_fill(*D, pos, ornt, ptype, pid);
}

int bdem_deck_load_state(bdem_deck* D, bdem_ctx* ctx, const char* path, bdem_state_header* hdr_out) {
  if (!D || !ctx) return BDEM_EINVAL;
  if (!path && D->info.read_file != 1) {
    /* no start file: the grid fill of Device_Set_SimulationData (:1771-1787) */
    const int n = D->sim.Num_Particles;
    std::vector<bdem_float3> pos(n);
    std::vector<bdem_Quaterion> q(n, bdem_Quaterion{1.0f, 0.0f, 0.0f, 0.0f});
    std::vector<unsigned> ptype(n);
    std::vector<int> pid(n);
    const int rc = bdem_deck_initial_state(D, pos.data(), q.data(), ptype.data(), pid.data());
    if (rc != BDEM_OK) return rc;
    if (hdr_out) std::memset(hdr_out, 0, sizeof(*hdr_out));
    return bdem_set_pos(ctx, pos.data(), D->info.particle_type == 1 ? q.data() : nullptr, ptype.data(), pid.data());
  }
  const std::string file = path ? std::string(path) : D->dir + "/" + D->info.start_file;
  const int n = D->sim.Num_Particles;
  std::vector<bdem_float3> pos(n), vel(n), avel(n), acc(n);
  std::vector<bdem_Quaterion> q(n, bdem_Quaterion{1.0f, 0.0f, 0.0f, 0.0f});
  std::vector<unsigned> ptype(n);
  std::vector<int> pid(n);
  bdem_state_header h;
  int rows = 0;
  int rc = bdem_state_read(file.c_str(), D->info.sim_type, D->info.rotating_world, D->info.particle_type, n, &h,
                           pos.data(), q.data(), vel.data(), avel.data(), acc.data(), ptype.data(), pid.data(),
                           &rows);
  if (rc != BDEM_OK) return rc;
  if (rows != n) return BDEM_EINVAL;   /* the reference reads Num_Particles rows */
  for (int i = 0; i < n; ++i)
    if (ptype[i] >= (unsigned)D->pobj.size()) return BDEM_EINVAL;
  if (hdr_out) *hdr_out = h;
  const bdem_Quaterion* qo = D->info.particle_type == 1 ? q.data() : nullptr;
  if (D->info.sim_type == BDEM_SIM_BALLMILL) {
    float powers[3] = {h.energy[0] * 1E4f, h.energy[1] * 1E4f, h.energy[2] * 1E4f};
    if ((rc = bdem_set_tallys(ctx, powers)) != BDEM_OK) return rc;
    if (h.rev_complete == 0 && !D->dobj.empty()) {
      rotate_lifters_to(*D, h.rot_angle);
      if ((rc = bdem_update_dobject_positions(ctx, D->dobj.data())) != BDEM_OK) return rc;
    }
  } else if (D->info.rotating_world && h.rot_angle != 0.0) {
    /* Rotate_WObject_angle (Input/Processing_Mills.h:164-240): not part of
     * this library (no shipped deck has a rotating world object) */
    return BDEM_ENOTSUP;
  }
  if (h.spec_type == 1) return bdem_set_pos(ctx, pos.data(), qo, ptype.data(), pid.data());
  if (h.spec_type == 2)
    return bdem_set_pos_velocity(ctx, pos.data(), qo, vel.data(), avel.data(), ptype.data(), pid.data());
  return bdem_set_system_state(ctx, pos.data(), qo, vel.data(), avel.data(), acc.data(), ptype.data(), pid.data());
}

}  // extern "C"
