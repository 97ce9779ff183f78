/*
 * bdem_types.h - plain-old-data types that cross the bdem C-ABI.
 *
 * Field order, sizes and alignment reproduce the structs the reference
 * passes to its device layer, so a caller that fills them for the reference
 * can hand the same bytes to this library:
 *   SimulationInfo, InitConfig, Surface, Edge_Def, PP_Parms, POBJ, DOBJ,
 *   Macro_Cylinder, WOBJ        DataStructures/KSimulationObjects.h:104-381
 *   Contact_Info                DataStructures/KSimulationObjects.h:18-21
 *   Quaterion                   Utilities/Common/Types.h:15-21
 * The reference's float3/int3 come from CUDA's vector_types.h (12 bytes,
 * 4-byte aligned); bdem_float3/bdem_int3 have that layout.  Units are the
 * reference's cm-kg-s.  The static_asserts at the bottom pin every offset.
 *
 * Everything here is C99 so that C, C++ and HIP translation units share it.
 */
#ifndef BDEM_TYPES_H_
#define BDEM_TYPES_H_

#include <stdint.h>
#include <stdbool.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bdem_float3 { float x, y, z; } bdem_float3;
typedef struct bdem_int3   { int   x, y, z; } bdem_int3;
typedef struct bdem_uint3  { unsigned int x, y, z; } bdem_uint3;

/* Utilities/Common/Types.h:15-21 (w first) */
typedef struct bdem_Quaterion { float w, x, y, z; } bdem_Quaterion;

/* KSimulationObjects.h:18-21, 16-byte aligned */
typedef struct bdem_Contact_Info {
  int num_contact, obj_id, snum, cont_type;
} __attribute__((aligned(16))) bdem_Contact_Info;

/* Sim_Type / SurfaceType enums: KSimulationObjects.h:49-54, 217-220 */
enum { BDEM_SIM_GENERAL = 0, BDEM_SIM_BALLMILL = 1, BDEM_SIM_SILO = 2 };
enum { BDEM_SURFACE_PLANE = 0, BDEM_SURFACE_CYLINDER = 1 };

/* KSimulationObjects.h:104-152 */
typedef struct bdem_SimulationInfo {
  int   particle_type;            /* 0 spheres, 1 polyhedra */
  int   Num_Particles;
  int   Num_ParticleObjects;
  int  *Num_ParticlesPerObject;   /* host pointer, never dereferenced by the library */
  int   Num_WorldObjects;
  int   TotnumCells;

  bool  Rotation;
  bool  Mode2D;
  bool  EnergyCalc;
  bool  isMillSim;
  bool  Kill_Particles;

  float InitalDelta_t;
  float MillRotVel;
  float Global_VisDamp;
  float Roll_Res;
  float Silo_Kill_Height;
  float Vel_Limit;
  float DObject_CylinderBoundR;

  bdem_int3   num_NNCells;
  bdem_float3 Force_Field;
  bdem_float3 worldOrigin;
  bdem_float3 cellSize;
  bdem_float3 cent;
  bdem_float3 Vel_Limit_Bounds;
  bdem_float3 max_size;

  int   Simulation_Type;          /* enum Sim_Type */
  int   Integration_Type;         /* 0 symplectic Euler (the only live path) */

  bool  use_symmetry;
  bool  sphere_orient;
  bool  dobject_hasVertex;
  bool  Rotating_WObject;
  bool  use_hist;

  int   unit_test;
} bdem_SimulationInfo;

/* KSimulationObjects.h:158-181 */
typedef struct bdem_InitConfig {
  bdem_int3   num;
  bdem_float3 space;
  bdem_float3 start;
  bdem_float3 p_size[16];
  bool        is_spheres;
  bool        use_file;
  int         grid_type;
  bdem_float3 launch_Vel;
  bdem_float3 fill_plane_start;
  bdem_float3 fill_plane_end;
  bdem_float3 velocity[12];
  bdem_float3 position[12];
  bdem_float3 ang_vel[12];
  bool        DebugFlag;
  int         use_device;
  bool        multi_gpu;
  int         threads_perBlock;
} bdem_InitConfig;

/* KSimulationObjects.h:224-227 */
typedef struct bdem_Edge_Def { unsigned int point_vindex[2]; } bdem_Edge_Def;

/* KSimulationObjects.h:241-251 */
typedef struct bdem_Surface {
  int         num_vertex;
  int         vertex_Order[8];
  float       area;               /* -1 marks an infinite plane */
  bdem_float3 normal;
  bdem_float3 centroid;
  float       Tol;                /* area tolerance in percent */
} bdem_Surface;

/* KSimulationObjects.h:261-267 */
typedef struct bdem_PP_Parms {
  float Kn, Cn, Fric_static, Fric_kinetic;
} bdem_PP_Parms;

/* KSimulationObjects.h:272-304 */
typedef struct bdem_POBJ {
  float         radius;           /* cm */
  float         mass;             /* kg */
  bdem_float3   COM;
  int           num_vertex;
  bdem_float3   Local_Vertex[32];
  int           num_edges;
  bdem_Edge_Def edge[32];
  int           num_faces;
  bdem_Surface  face[32];
  float         InertiaT[9];      /* spheres: InertiaT[0] = 1/(0.4 r^2 m) */
  bdem_float3   vertex[32];
  bdem_PP_Parms PP[6];
  float surface_Kn, surface_Cn, surface_Fric_static, surface_Fric_kinetic;
  float Dsurface_Kn, Dsurface_Cn, Dsurface_Fric_static, Dsurface_Fric_kinetic;
} bdem_POBJ;

/* KSimulationObjects.h:318-342 */
typedef struct bdem_DOBJ {
  unsigned int  btype;
  bdem_float3   Axis;
  float         boundR;
  bdem_float3   COM;
  int           num_vertex;
  bdem_float3   vertex[8];
  int           num_faces;
  bdem_Surface  faces[3];
  int           num_edges;
  bdem_Edge_Def edge[4];
  bdem_float3   velocity;
  bool          is_attached_DWorld;
  bool          is_translating;
} bdem_DOBJ;

/* KSimulationObjects.h:350-363 */
typedef struct bdem_Macro_Cylinder {
  float       height;
  float       radius;
  bool        has_top_cap;
  bool        has_bot_cap;
  bdem_float3 normal_cyl;
  bdem_float3 normal_top_cap;
  bdem_float3 normal_bot_cap;
  bdem_float3 center_top_cap;
  bdem_float3 center_bot_cap;
  bdem_float3 AXIS;               /* component mask, (1,1,1) - normal_bot_cap */
} bdem_Macro_Cylinder;

/* KSimulationObjects.h:370-381 */
typedef struct bdem_WOBJ {
  int                 surface_type;   /* enum SurfaceType */
  int                 num_surfaces;
  bdem_Surface        surfaces[16];
  int                 num_vertex;
  bdem_float3         vertex[16];
  bool                is_rotating;
  bdem_Macro_Cylinder cyl_geo;
} bdem_WOBJ;

/* ------------------------------------------------------------------------ */
/* Extensions with no counterpart in the reference.                          */
/* ------------------------------------------------------------------------ */

/* Contact laws.  BDEM_LAW_REFERENCE is the reference's pair/wall law
 * (Collision_Response.cuh:4-526): linear spring-dashpot normal force and a
 * velocity-ratio tangential force with no history.  The other two are the
 * textbook laws with a tangential spring carried across steps. */
enum {
  BDEM_LAW_REFERENCE     = 0,
  BDEM_LAW_LSD_COULOMB   = 1,   /* Cundall-Strack: kn, cn from e; kt spring; Coulomb cap */
  BDEM_LAW_HERTZ_MINDLIN = 2    /* Hertz normal, Mindlin no-slip tangential, Tsuji damping */
};

/* Per particle-type material for the history laws. */
typedef struct bdem_material {
  float kn;            /* LSD normal stiffness, kg/s^2 */
  float kt_over_kn;    /* LSD tangential stiffness ratio (2/7 for the box deck) */
  float restitution;   /* e, pair and wall */
  float mu;            /* sliding friction, pair and wall */
  float youngs;        /* Hertz E, kg cm^-1 s^-2 (1 Pa = 0.01) */
  float poisson;       /* Hertz nu */
  float pad0, pad1;
} bdem_material;

/* Analytic conical-frustum wall (a hopper).  Points on the axis are
 * apex_base + s*axis; the wall is the surface of radius r(s) between s0 and
 * s1 with r linear from r0 to r1.  Particles live inside it. */
typedef struct bdem_cone {
  bdem_float3 base;    /* point on the axis where s = 0 */
  bdem_float3 axis;    /* unit axis */
  float s0, s1;        /* axial extent */
  float r0, r1;        /* radius at s0 and s1 */
} bdem_cone;

#define BDEM_MAX_CONES 4

typedef struct bdem_ext_params {
  int   contact_law;          /* BDEM_LAW_* */
  int   per_particle_radius;  /* 1: radius/mass come from bdem_set_radii */
  int   num_cones;
  int   kill_check_interval;  /* steps between discharge checks; 0 = 1000 like the reference */
  bdem_material material[8];  /* indexed by particle type */
  bdem_cone     cone[BDEM_MAX_CONES];
  float wall_youngs;          /* Hertz wall E (0 = same as the particle) */
  float wall_poisson;
  float pad[2];
} bdem_ext_params;

/* The geometry of one convex polyhedron, for per-particle shapes (no
 * reference counterpart: the reference takes every polyhedron's geometry
 * from its type's POBJ, at most 8 types of at most 32 vertices, faces and
 * edges, KSimulationObjects.h:272-316).  The fields the polyhedra path
 * reads, named as in POBJ; a triangulated hull of V vertices has 2V - 4
 * faces and 3V - 6 edges, so every hull of up to 18 vertices fits.  vbound
 * is the largest |Local_Vertex| times 1.0001; bdem_set_poly_shapes
 * recomputes it and plane_slack. */
#define BDEM_SHAPE_MAX_VERTEX 32
#define BDEM_SHAPE_MAX_FACE 32
#define BDEM_SHAPE_MAX_EDGE 64
typedef struct bdem_poly_shape {
  float         radius;           /* bounding radius about COM (cm) */
  float         mass;             /* kg */
  bdem_float3   COM;
  int           num_vertex;
  bdem_float3   Local_Vertex[BDEM_SHAPE_MAX_VERTEX];
  int           num_edges;
  bdem_Edge_Def edge[BDEM_SHAPE_MAX_EDGE];
  int           num_faces;
  bdem_Surface  face[BDEM_SHAPE_MAX_FACE];
  float         InertiaT[9];      /* inverse inertia tensor about COM (body frame) */
  float         vbound;
  float         plane_slack;      /* largest height of a vertex above a face plane of its own
                                     shape (0 for a convex hull up to rounding), cm */
} bdem_poly_shape;

#ifdef __cplusplus
} /* extern "C" */
#endif

/* Layout pins: offsets of the reference structs on x86-64 / gfx950 with a
 * 12-byte, 4-aligned float3. */
#ifdef __cplusplus
#define BDEM_SA(c, m) static_assert(c, m)
#else
#define BDEM_SA(c, m) _Static_assert(c, m)
#endif
BDEM_SA(sizeof(bdem_float3) == 12, "float3");
BDEM_SA(sizeof(bdem_Quaterion) == 16, "Quaterion");
BDEM_SA(sizeof(bdem_Contact_Info) == 16, "Contact_Info");
BDEM_SA(offsetof(bdem_SimulationInfo, Num_ParticlesPerObject) == 16, "SimInfo ptr");
BDEM_SA(offsetof(bdem_SimulationInfo, Rotation) == 32, "SimInfo bools");
BDEM_SA(offsetof(bdem_SimulationInfo, InitalDelta_t) == 40, "SimInfo dt");
BDEM_SA(offsetof(bdem_SimulationInfo, num_NNCells) == 68, "SimInfo cells");
BDEM_SA(offsetof(bdem_SimulationInfo, max_size) == 140, "SimInfo max_size");
BDEM_SA(offsetof(bdem_SimulationInfo, Simulation_Type) == 152, "SimInfo type");
BDEM_SA(offsetof(bdem_SimulationInfo, use_symmetry) == 160, "SimInfo flags");
BDEM_SA(offsetof(bdem_SimulationInfo, unit_test) == 168, "SimInfo unit_test");
BDEM_SA(sizeof(bdem_SimulationInfo) == 176, "SimInfo size");
BDEM_SA(sizeof(bdem_InitConfig) == 720, "InitConfig size");
BDEM_SA(sizeof(bdem_Surface) == 68, "Surface size");
BDEM_SA(offsetof(bdem_POBJ, edge) == 412, "POBJ edge");
BDEM_SA(offsetof(bdem_POBJ, face) == 672, "POBJ face");
BDEM_SA(offsetof(bdem_POBJ, InertiaT) == 2848, "POBJ inertia");
BDEM_SA(offsetof(bdem_POBJ, PP) == 3268, "POBJ PP");
BDEM_SA(sizeof(bdem_POBJ) == 3396, "POBJ size");
BDEM_SA(sizeof(bdem_DOBJ) == 392, "DOBJ size");
BDEM_SA(sizeof(bdem_Macro_Cylinder) == 84, "Macro_Cylinder size");
BDEM_SA(offsetof(bdem_WOBJ, cyl_geo) == 1296, "WOBJ cyl_geo");
BDEM_SA(sizeof(bdem_WOBJ) == 1380, "WOBJ size");
BDEM_SA(offsetof(bdem_poly_shape, face) == 928, "poly_shape face");
BDEM_SA(sizeof(bdem_poly_shape) == 3148, "poly_shape size");

#undef BDEM_SA

#endif /* BDEM_TYPES_H_ */
