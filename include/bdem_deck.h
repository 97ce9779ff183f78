/*
 * bdem_deck.h - project decks and state files (host only, part of libbdem.so).
 *
 * Replaces the reference's input layer: the KSimulationData reader of
 * .World / .WOBJ / .POBJ / .DOBJ files (Input/KSimulationData.cpp:27-2028),
 * the derived simulation parameters of Set_SimInfo (main.cpp:469-833) and the
 * state-file reader and writer (Read_FromFile, Output_SystemState,
 * Input/FileIO.h:12-652).  A project directory holds World/, SurfaceObjects/,
 * ParticleObjects/, LifterObjects/ (mills) or VolumeObjects/ and the start
 * files, as the reference's Projects/<name>/ folders do.
 *
 * Differences, by design:
 *   - errors are returned (with a message) instead of exit(1), and a token
 *     that does not parse is an error instead of a silently failed stream;
 *   - both .World layouts are read: Delta_t right after Particle_Type (the
 *     one the reference parses) or in section 2 after Num_Particles (the one
 *     8 of the 13 shipped World files use, which the reference rejects);
 *   - a triangular polyhedron face gets the area 0.5 |n| (the reference
 *     overwrites it with a value left from an earlier face or uninitialised,
 *     Input/KSimulationData.cpp:1475);
 *   - SpecType 4 files are written with the acceleration columns the reader
 *     expects (the reference writer omits them, FileIO.h:232-250).
 */
#ifndef BDEM_DECK_H_
#define BDEM_DECK_H_

#include "bdem_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bdem_deck bdem_deck;

/* What a deck holds beyond the device tables: run control and output
 * options (KSimulationData members, Input/KSimulationData.h:95-165, and the
 * derived mill/silo values of main.cpp:683-806). */
typedef struct bdem_deck_info {
  char   title[128];
  int    force_spec;        /* 0: contact time and COR (TC), 1: stiffness and COR (KN) */
  int    sim_type;          /* BDEM_SIM_GENERAL / BALLMILL / SILO */
  int    particle_type;     /* 0 spheres, 1 polyhedra */
  int    dt_in_setup;       /* 1: Delta_t was read from section 2 */
  int    num_wobj, num_pobj, num_dobj;   /* num_dobj after lifter repetition */
  int    dobj_types;        /* DOBJ files named in the World file */
  int    dobj_repeat;       /* "Repeat:" (lifters per type; 0 = as listed) */
  int    num_per_pobj[8];
  int    read_file;         /* 1: the start state comes from start_file */
  char   start_file[256];
  /* section 5: output options */
  int    render, color_type;
  float  color_bin;
  int    ortho, fps, gl_debug, write_steps, output_format;
  float  total_time;        /* after the mill/silo defaults of Set_SimInfo */
  int    is_run_limit;
  float  run_time;
  int    energy_steps, wall_steps;
  /* mill (section 6) and its derived rotation */
  int    power_revs, total_revs;
  float  mill_rpm;
  int    steps_per_rev;     /* Num_Steps_PerRev = (int)((60/RPM)/dt) */
  double rad_per_step;      /* RadPerStep = -(RPM*2*PI/60)*dt, PI = 3.14159265 */
  float  vel_mill;          /* |RadPerStep/dt * R * 0.01| (R - 100 for a rotating world) */
  bdem_float3 rotate_about; /* (R, R, 0): the axis point the lifters turn about */
  /* silo (section 6) */
  int    manual_hatch, measure_flow, flow_steps, kill_particle;
  float  hatch_height;      /* y of the first vertex of the last world object */
  /* world object 0 rotates (index 99 in the World file) */
  int    rotating_world;
} bdem_deck_info;

/* Header of a state file.  Which fields are present depends on the
 * simulation type, as in Read_FromFile (FileIO.h:348-491). */
typedef struct bdem_state_header {
  int    spec_type;         /* 1 pos, 2 +vel, 3 +vel (+quat) +avel, 4 +vel +acc (+quat) +avel +magnitudes */
  double time;
  int    rev_num;           /* mill */
  int    rev_complete;      /* mill */
  double rot_angle;         /* mill, rotating world */
  float  energy[5];         /* mill: EParticle ELift EDrum ENorm EShear as written */
  int    num_rem;           /* silo */
  int    rot_steps;         /* rotating world */
} bdem_state_header;

/* Simulation.KSim (main.cpp:843-848): project folder name, world name,
 * RunMode (1 = 2D), GPU number, device count.  Strings are truncated to
 * `cap` bytes. */
int bdem_ksim_read(const char* path, char* project, char* world, int cap, int* mode2d, int* device,
                   int* num_devices);

/* Read <project_dir>/World/<world>.World and the object files it names,
 * derive the spring constants (Set_SpringConstants*, :887-1071), repeat the
 * lifters (Repeat_Lifters, :1962-2028) and the SimulationInfo / InitConfig
 * of Set_SimInfo (main.cpp:469-833).  mode2d is Simulation.KSim's RunMode.
 * On failure *out is NULL and err (if given) holds the reason. */
int bdem_deck_load(bdem_deck** out, const char* project_dir, const char* world, int mode2d, char* err,
                   int err_cap);
void bdem_deck_free(bdem_deck* deck);

/* Pointers into the deck (valid until bdem_deck_free); any may be NULL. */
int bdem_deck_get(const bdem_deck* deck, const bdem_deck_info** info, const bdem_SimulationInfo** sim,
                  const bdem_InitConfig** init, const bdem_WOBJ** wobj, const bdem_POBJ** pobj,
                  const bdem_DOBJ** dobj);

/* bdem_set_simulation_data with the deck's tables (the KDevice constructor,
 * main.cpp:941-946). */
int bdem_deck_setup(const bdem_deck* deck, bdem_ctx* ctx);

/* Read a state file (Read_FromFile, FileIO.h:327-611).  sim_type and
 * rotating_world select the header; particle_type whether orientations are
 * present.  Up to n rows are read into the arrays (any array may be NULL
 * except pos, ptype and pid); *rows receives how many rows were read. */
int bdem_state_read(const char* path, int sim_type, int rotating_world, int particle_type, int n,
                    bdem_state_header* hdr, bdem_float3* pos, bdem_Quaterion* ornt, bdem_float3* vel,
                    bdem_float3* avel, bdem_float3* acc, unsigned int* ptype, int* pid, int* rows);

/* Write a state file in the layout of Output_SystemState (FileIO.h:12-326)
 * with detail hdr->spec_type (1-4; 5 writes each polyhedron's vertices and
 * needs pobj).  Rows with pid == -1 are skipped, as in the reference. */
int bdem_state_write(const char* path, int sim_type, int rotating_world, int particle_type, int n,
                     const bdem_state_header* hdr, const bdem_float3* pos, const bdem_Quaterion* ornt,
                     const bdem_float3* vel, const bdem_float3* avel, const bdem_float3* acc,
                     const unsigned int* ptype, const int* pid, const bdem_POBJ* pobj, int num_pobj);

/* The initial state of a deck without a start file (Read_FromFile 0):
 * Set_Position_DefaultGrid (GridType 0) or Set_Position_Random_Grid
 * (GridType 1: for polyhedra a random turn of up to 30 degrees about an
 * unnormalised random axis in [0,1]^3, then a shuffle of types and IDs, all
 * drawn from the C library's unseeded rand() sequence in the reference's
 * order), Device/DeviceInterface.cu:192-601.
 * Num_Particles entries each; ornt may be NULL.  BDEM_ENOTSUP for a start
 * file or a mill fill (GridType 2). */
int bdem_deck_initial_state(const bdem_deck* deck, bdem_float3* pos, bdem_Quaterion* ornt, unsigned int* ptype,
                            int* pid);

/* Seeded random non-overlapping fill (random sequential addition; no
 * reference counterpart: the reference's fills are lattices,
 * Device/DeviceInterface.cu:192-601).  Sphere k (in index order) gets a
 * centre drawn uniformly from the box [lo + r_k, hi - r_k] (also within
 * radius cyl[2] - r_k of the vertical axis x = cyl[0], z = cyl[1] when cyl
 * is not NULL), redrawn until its distance to every placed centre is at
 * least r_j + r_k + gap, at most max_tries draws.  Coordinates come from
 * splitmix64: state += 0x9E3779B97F4A7C15, then the standard mix; a draw is
 * (z >> 40) * 2^-24, three per try (x, y, z).  Host only.  BDEM_EINVAL with
 * the index in *failed (may be NULL) when a sphere cannot be placed. */
int bdem_fill_random(int n, const float* radius, bdem_float3 lo, bdem_float3 hi, const float* cyl, float gap,
                     unsigned long long seed, int max_tries, bdem_float3* pos, int* failed);

/* Convex polyhedron shapes from point sets (no reference counterpart: the
 * reference reads each polyhedron's vertices, faces and inertia from its
 * POBJ file, Input/KSimulationData.cpp:1180-1530).  Shape s is the convex
 * hull of the npts[s] (4..max_pts, max_pts <= 32) points
 * pts[s * max_pts ...]: hull faces (coplanar points merged into one polygon
 * of at most 8 vertices) ordered counter-clockwise from outside, points
 * inside the hull dropped.  Derived as the reference's reader derives a
 * POBJ: COM = vertex mean, Local_Vertex = vertex - COM, radius = largest
 * |Local_Vertex|, face centroid = vertex mean, face normal =
 * normalise(cross(v1 - v2, v2 - v0)) (triangles) or
 * normalise(cross(v2 - v3, v3 - v0)) (polygons), area = centroid fan, all
 * in float; edges in face order; mass = density (kg/m^3) x volume (cm^3)
 * x 1e-6 and InertiaT = inverse of the solid's inertia tensor about COM,
 * from tetrahedra in double.  Every face gets Tol = face_tol.  Host only.
 * BDEM_EINVAL with the shape index in *failed (may be NULL) for a
 * degenerate (flat) point set or a hull beyond the table's limits. */
int bdem_poly_shapes_from_points(int count, const int* npts, const bdem_float3* pts, int max_pts, float density,
                                 float face_tol, bdem_poly_shape* out, int* failed);

/* Read a state file of this deck (path NULL: <project_dir>/<start_file>, or
 * the grid fill of bdem_deck_initial_state through bdem_set_pos when the
 * deck has no start file)
 * and hand it to the context as Read_FromFile does: SpecType 1 through
 * bdem_set_pos, 2 bdem_set_pos_velocity, 3 and 4 bdem_set_system_state; a
 * mill file's energies, times 1E4, through bdem_set_tallys.  A mill file of
 * an unfinished revolution (RevComplete 0) turns the lifters to its
 * RotAngle (Rotate_VolObject_angle, Input/Processing_Mills.h:92-160) and
 * uploads them.  hdr (may be NULL) receives the header. */
int bdem_deck_load_state(bdem_deck* deck, bdem_ctx* ctx, const char* path, bdem_state_header* hdr);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* BDEM_DECK_H_ */
