/*
 * bdem_abi.h - C ABI of the MI355X DEM step library (libbdem.so).
 *
 * The entry points replace the reference's device layer, the C++ free
 * functions of Device/DeviceInterface.h:16-95 that the KDevice facade
 * (Interface/KDevice.h:17-99) forwards to.  Differences, by design:
 *   - state lives in an opaque context (the reference keeps process globals,
 *     Device/DeviceInterface.cu:52-186), so several simulations can coexist;
 *   - every call returns a status code instead of calling exit(1)
 *     (Device/DeviceInterface.cu:1891-1899 and after every launch);
 *   - bdem_update_sim is asynchronous on the context's stream; results are
 *     visible after bdem_sync or any bdem_get_* call;
 *   - no allocation happens per step (the reference allocates and frees
 *     about a dozen buffers per step, Device/DeviceInterface.cu:1929-2057).
 * Units and array layouts are the reference's: cm-kg-s, AoS float3
 * (12 bytes) positions/velocities, Quaterion {w,x,y,z}.
 *
 * A C++ shim with the reference's own names and signatures
 * (Device_Set_SimulationData, Device_DEM_UpdateSim, ...) is declared in
 * bdem_device_compat.h.
 */
#ifndef BDEM_ABI_H_
#define BDEM_ABI_H_

#include "bdem_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define BDEM_OK            0
#define BDEM_EINVAL       -1   /* bad argument or configuration */
#define BDEM_EHIP         -2   /* HIP runtime error (message in bdem_last_error) */
#define BDEM_ENOTSUP      -3   /* entry point without a sphere-path implementation */
#define BDEM_ESTATE       -4   /* called in the wrong order (e.g. step before setup) */
#define BDEM_ENOMEM       -5   /* device allocation failed */

/* Warning bits reported by bdem_get_status (kernels never abort). */
#define BDEM_WARN_OUT_OF_GRID    1u  /* a live particle left the hash grid */
#define BDEM_WARN_LIST_OVERFLOW  2u  /* > 32 contacts on a particle (reference list size) */
#define BDEM_WARN_HIST_OVERFLOW  4u  /* more contacts than tangential-history slots */
#define BDEM_WARN_SORT           8u  /* internal: incremental sort consistency check failed */

typedef struct bdem_ctx bdem_ctx;

/* ---- lifecycle ---------------------------------------------------------- */
/* Create a context on HIP device `device`.  No reference counterpart (the
 * reference selects device 0 inside Device_Set_SimulationData,
 * Device/DeviceInterface.cu:1566). */
int bdem_create(bdem_ctx** out, int device);
/* Device_Clean, Device/DeviceInterface.h:76 (frees everything the context owns). */
int bdem_destroy(bdem_ctx* ctx);
/* Text of the last error on this context (never NULL). */
const char* bdem_last_error(const bdem_ctx* ctx);

/* ---- setup -------------------------------------------------------------- */
/* Device_Set_SimulationData, Device/DeviceInterface.h:16-19 (defined
 * Device/DeviceInterface.cu:1499-1816).  Allocates for sim->Num_Particles,
 * uploads the object tables and sets every particle's velocity to
 * init->velocity[0] (zero when init->use_file), angular velocity to zero.
 * Positions come from bdem_set_pos* (the reference's grid fillers are not
 * part of this library). */
int bdem_set_simulation_data(bdem_ctx* ctx, const bdem_WOBJ* wobj, const bdem_POBJ* pobj,
                             const bdem_DOBJ* dobj, const bdem_SimulationInfo* sim, int num_wobj,
                             int num_pobj, int num_dobj, const bdem_InitConfig* init);
/* Extension parameters (contact law, cones, per-particle radius); call after
 * bdem_set_simulation_data.  No reference counterpart. */
int bdem_set_ext_params(bdem_ctx* ctx, const bdem_ext_params* ext);

/* Re-upload SimulationInfo (time step, gravity, flags ...) between steps, as
 * the reference re-uploads SimParms after fills and kills
 * (Device/DeviceInterface.cu:1095, 2089).  Particle count and grid
 * dimensions are fixed at setup. */
int bdem_update_simulation_info(bdem_ctx* ctx, const bdem_SimulationInfo* sim);

/* ---- materials (host only, no context, no device) ------------------------- */
/* Spring constants from restitution, as the deck reader derives them:
 * Set_SpringConstantsParticleKN / Surface_KN / DSurfaceKN (mode
 * BDEM_SPRING_KN, Input/KSimulationData.cpp:955-1071: the caller sets
 * PP[i].Kn, PP[i].Fric_* of every type i, surface_Kn and Dsurface_Kn) and
 * Set_SpringConstantsParticle_TC / Surface_TC / DSurface_TC (mode
 * BDEM_SPRING_TC, :887-953: the caller sets PP[0].Fric_*; tc_* are contact
 * times in seconds, the deck value times the time step as at :1119-1121).
 * Fills the PP matrix row by row (:276-292) and the wall dampings; at most
 * 6 types (POBJ::PP[6]). */
#define BDEM_SPRING_TC 0
#define BDEM_SPRING_KN 1
int bdem_derive_spring_constants(bdem_POBJ* pobj, int num_pobj, int mode, const double* cor_pp,
                                 const double* cor_surface, const double* cor_dsurface,
                                 double tc_pp, double tc_ps, double tc_pd);

/* ---- state in ----------------------------------------------------------- */
/* Device_Set_Pos, Device/DeviceInterface.h:22-23 (ornt may be NULL for spheres). */
int bdem_set_pos(bdem_ctx* ctx, const bdem_float3* pos, const bdem_Quaterion* ornt,
                 const unsigned int* ptype, const int* pid);
/* Device_Set_Pos_Velocity, Device/DeviceInterface.h:26-28 */
int bdem_set_pos_velocity(bdem_ctx* ctx, const bdem_float3* pos, const bdem_Quaterion* ornt,
                          const bdem_float3* vel, const bdem_float3* avel,
                          const unsigned int* ptype, const int* pid);
/* Device_Set_System_State, Device/DeviceInterface.h:31-33 (acc is accepted and
 * ignored, as Set_PSystem_State ignores it, GPU_Memory_Transfer.cuh:397-440). */
int bdem_set_system_state(bdem_ctx* ctx, const bdem_float3* pos, const bdem_Quaterion* ornt,
                          const bdem_float3* vel, const bdem_float3* avel, const bdem_float3* acc,
                          const unsigned int* ptype, const int* pid);
/* Per-particle radius (cm) and mass (kg) in the current particle order;
 * requires ext->per_particle_radius.  No reference counterpart (the
 * reference takes both from the particle type). */
int bdem_set_radii(bdem_ctx* ctx, const float* radius, const float* mass);
/* Device_Set_Tallys, Device/DeviceInterface.h:36 */
int bdem_set_tallys(bdem_ctx* ctx, const float* tallys);

/* ---- state out (synchronous) -------------------------------------------- */
/* Device_Get_System_State, Device/DeviceInterface.h:42-47: all
 * Num_Particles slots in device order; acc (may be NULL) is zero. */
int bdem_get_system_state(bdem_ctx* ctx, bdem_float3* pos, bdem_Quaterion* ornt, bdem_float3* vel,
                          bdem_float3* avel, bdem_float3* acc, unsigned int* ptype, int* pid);
/* Device_Get_P_Positions, Device/DeviceInterface.h:50-51 */
int bdem_get_positions(bdem_ctx* ctx, bdem_float3* pos, bdem_Quaterion* ornt, unsigned int* ptype,
                       int* pid);
/* Device_Get_P_Velocity, Device/DeviceInterface.h:54-55 (avel written when flag == 1) */
int bdem_get_velocity(bdem_ctx* ctx, bdem_float3* vel, bdem_float3* avel, int flag,
                      unsigned int* ptype, int* pid);
/* Device_Get_W_Forces, Device/DeviceInterface.h:61: wall force of the last
 * step per particle (recorded only after bdem_record_wall_forces(ctx, 1));
 * contact points are not recorded and are returned as zero. */
int bdem_get_wall_forces(bdem_ctx* ctx, bdem_float3* wall_forces, bdem_float3* contact_points);
int bdem_record_wall_forces(bdem_ctx* ctx, int on);
/* Device_Get_Energy, Device/DeviceInterface.h:63: the three dissipation
 * tallies {particle-particle, particle-lifter, particle-surface} scaled 1E-4
 * like GetEnergy (GPU_Memory_Transfer.cuh:603-619), i.e. in J.  The step
 * adds to them while SimulationInfo::EnergyCalc is set (spheres, reference
 * law), one float atomic per contact as in Collision_Response.cuh. */
int bdem_get_energy(bdem_ctx* ctx, float* energy);
/* Ensemble statistics of the current state (synchronous).  Over the live
 * particles with y > min_y: how many (the silo's particles left above the
 * hatch, Silo_Stats, main.cpp:253-285), their mean and largest speed in m/s
 * (|v| * 0.01) and the P_ID of the fastest (Output_SystemState,
 * Input/FileIO.h:286-318).  Over the live particles whose centre lies in the
 * box [box_lo, box_hi) (both NULL: none): how many, and their solid
 * fraction, the sum of 4/3 pi r^3 over the box volume (spheres; -1 for
 * polyhedra), which has no reference counterpart. */
typedef struct bdem_particle_stats {
  int    count;
  int    max_pid;              /* -1 when count is 0 */
  double mean_speed;           /* m/s */
  float  max_speed;            /* m/s */
  int    box_count;
  double box_solid_fraction;
} bdem_particle_stats;
int bdem_get_particle_stats(bdem_ctx* ctx, float min_y, const bdem_float3* box_lo, const bdem_float3* box_hi,
                            bdem_particle_stats* out);
/* Point cloud of a polyhedra context (synchronous): for each particle in
 * the library's order, its vertices in the world frame, the local vertices
 * turned by the orientation plus the centre of mass (as RotatePolyhedra,
 * Device/Utilities.cuh:420-479, makes them for the contact search), in
 * verts[i * max_vertex + v]; slots past the type's count are zero.
 * num_vertex and pid (either may be NULL) receive each particle's count and
 * P_ID.  The reference writes only the unturned vertices (SpecType 5,
 * Input/FileIO.h:253-263).  BDEM_EINVAL for a sphere context or a
 * max_vertex below the largest vertex count. */
int bdem_get_poly_vertices(bdem_ctx* ctx, int max_vertex, bdem_float3* verts, int* num_vertex, int* pid);
/* Per-particle polyhedron shapes (no reference counterpart: the reference
 * takes every polyhedron's geometry from its type's POBJ row,
 * KSimulationObjects.h:272-316, read by ReadPObject_Poly,
 * Input/KSimulationData.cpp:1180-1530).  After this call the particle with
 * P_ID p has the geometry, bounding radius, mass and inverse inertia of
 * shapes[p]; its type still selects the pair, wall and lifter constants
 * (POBJ::PP and the surface constants).  count rows, each checked as a POBJ
 * row is but with the larger bdem_poly_shape limits; vbound is recomputed.
 * Every live P_ID of the current state and of later uploads must be below
 * count; the current state's radius and mass are rewritten from the rows.
 * Synchronous.  BDEM_EINVAL for a sphere context or a bad table (the
 * context then keeps its previous table). */
int bdem_set_poly_shapes(bdem_ctx* ctx, const bdem_poly_shape* shapes, int count);
/* Rows of the shape table the polyhedra path reads (the POBJ count, or the
 * count of bdem_set_poly_shapes) and whether rows are P_IDs (1) or types. */
int bdem_get_num_poly_shapes(bdem_ctx* ctx, int* count, int* by_pid);
/* Free-surface profile of the current state (synchronous), for the angle
 * of repose; no reference counterpart (the reference computes no angle).
 * [lo, hi) along axis 0 (x) or 2 (z) is cut into `bins` (1..4096) equal
 * bins; top[k] receives the highest particle top y + r among the live
 * particles whose centre falls in bin k (polyhedra: centre of mass plus the
 * bounding radius), NaN for an empty bin.  A centre's bin is
 * (int)((u - lo) * inv_w) with inv_w = (float)bins / (hi - lo), in float. */
int bdem_get_surface_profile(bdem_ctx* ctx, int axis, float lo, float hi, int bins, float* top);
/* bdem_get_surface_profile without any allocation, for callers that cannot
 * allocate device memory (serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  Same arguments and results; its scratch
 * (the bin keys and heights, 2 x 4096 words) is a file-scope device array
 * of the library, one per GPU, whose address is taken on the first call on
 * each GPU.  All calls on one GPU share that array: calls that overlap on
 * one GPU (from several threads or contexts) race.  bins above 4096 are
 * rejected (BDEM_EINVAL). */
int bdem_get_surface_profile_noalloc(bdem_ctx* ctx, int axis, float lo, float hi, int bins, float* top);
/* Host only: least-squares line h = intercept + slope * u through the
 * non-empty (non-NaN) bins of a profile, u at the bin centres
 * lo + (k + 0.5)(hi - lo)/bins, in double; angle_deg = atan|slope| in
 * degrees.  BDEM_EINVAL with fewer than two non-empty bins.  Any output
 * may be NULL. */
int bdem_surface_angle(const float* top, int bins, float lo, float hi, double* slope, double* intercept,
                       double* angle_deg);
/* Kernel warnings accumulated since the last call (BDEM_WARN_* bits); clears them. */
int bdem_get_status(bdem_ctx* ctx, unsigned int* warnings);
/* Particles removed by discharge checks so far. */
int bdem_get_kill_count(bdem_ctx* ctx, int* killed);
int bdem_num_particles(const bdem_ctx* ctx);
long long bdem_num_cells(const bdem_ctx* ctx);

/* ---- actions -------------------------------------------------------------- */
/* Device_Update_DObjectPostions, Device/DeviceInterface.h:72 */
int bdem_update_dobject_positions(bdem_ctx* ctx, const bdem_DOBJ* dobj);
/* Device_Update_WorldObjectPostions, Device/DeviceInterface.h:73 */
int bdem_update_world_object_positions(bdem_ctx* ctx, const bdem_WOBJ* wobj);
/* Device_OpenHatch, Device/DeviceInterface.h:75: upload, then drop the last
 * world object (Hatch, GPU_Memory_Transfer.cuh:703-706). */
int bdem_open_hatch(bdem_ctx* ctx, const bdem_WOBJ* wobj, int num_wobj);
/* Device_Reset_Energy, Device/DeviceInterface.h:77 */
int bdem_reset_energy(bdem_ctx* ctx, int num_values);
/* drum_rotation, Device/DeviceInterface.h:80 */
int bdem_drum_rotation(bdem_ctx* ctx, int cylinder_rotation);
/* Turn every lifter attached to the drum by rad_per_step about the point
 * (ax, ay, az) in the xy plane before each step, on the device: the
 * reference's UpdateSim does it on the host (Rotate_VolObject_Inc,
 * Input/Processing_Mills.h:7-60) and re-uploads the table every step
 * (Device_Update_DObjectPostions, Device/DeviceInterface.cu:2826-2831).
 * Same arithmetic (double angle, float coordinates, faces re-derived).
 * rad_per_step 0 turns it off.  No counterpart as an entry point. */
int bdem_set_lifter_rotation(bdem_ctx* ctx, double rad_per_step, float ax, float ay, float az);
/* The lifter table as it stands (synchronous); out may be NULL to ask for
 * the count.  No reference counterpart (the host owns the table there). */
int bdem_get_dobjects(bdem_ctx* ctx, bdem_DOBJ* out, int* num);
/* Start_Bin, Device/DeviceInterface.h:81 (a host flag in the reference) */
int bdem_start_bin(bdem_ctx* ctx);
/* Device_AddForce / SubForce / ZeroForce, Device/DeviceInterface.h:83-85:
 * interactive no-ops in the reference (they change a device global no kernel
 * reads); accepted and ignored. */
int bdem_add_force(bdem_ctx* ctx);
int bdem_sub_force(bdem_ctx* ctx);
int bdem_zero_force(bdem_ctx* ctx);

/* ---- the hot path --------------------------------------------------------- */
/* Device_DEM_UpdateSim, Device/DeviceInterface.h:95 (defined
 * Device/DeviceInterface.cu:1825-2513): one step, asynchronous. */
int bdem_update_sim(bdem_ctx* ctx);
/* nsteps consecutive bdem_update_sim calls.  Steady-state pairs of steps
 * (incremental sort, no discharge check, skin masks, stage timing or
 * recording among them; spheres) are replayed from a HIP graph captured
 * once per parity: the same kernels with the same arguments, fewer launch
 * costs.  No reference counterpart as an entry point. */
int bdem_step(bdem_ctx* ctx, int nsteps);
/* Turn the graph replay of bdem_step on (default) or off. */
int bdem_set_graphs(bdem_ctx* ctx, int on);
/* Wait for the context's stream. */
int bdem_sync(bdem_ctx* ctx);

/* Skin masks of the sphere step, in cm.  No reference counterpart: the
 * reference tests every candidate of the 27-cell stencil every step
 * (Collision_Detection_BroadPhase.cuh:402-513).  With skin > 0 the step
 * keeps, per particle and stencil row, which candidates are within
 * ra + rb + skin and tests only those until a particle enters or leaves the
 * row's cells or the displacement bound of the epoch reaches skin / 2; the
 * contacts and their order are the reference's either way.  skin <= 0
 * turns the masks off (the default: on the silo benchmark the unmasked step
 * is faster).  Masks need Nx == Nz (an injective hash) and are not used for
 * polyhedra. */
int bdem_set_skin(bdem_ctx* ctx, float skin);
/* The skin in use (0: masks off). */
float bdem_get_skin(const bdem_ctx* ctx);

/* ---- measurement --------------------------------------------------------- */
/* Record HIP events on the context's stream around each stage of the next
 * max_steps bdem_update_sim calls.  No reference counterpart (the reference
 * has no timers). */
int bdem_timing_begin(bdem_ctx* ctx, int max_steps);
/* Wait, then return the summed milliseconds of the 3 stages
 * {sort (hash if needed, resort or full sort, gather, cell table), lists
 * (skin-list marking and rebuilding; 0 when the lists are off), fused step
 * kernel} and the number of steps recorded; releases the events. */
int bdem_timing_end(bdem_ctx* ctx, double* stage_ms, int* nsteps);
/* Wait, then return how many particles the last sort had to move: the size
 * of the changed list for an incremental resort, N for a full sort.  No
 * reference counterpart. */
int bdem_sort_changed(bdem_ctx* ctx, unsigned int* changed);
/* Run a plain copy kernel (k_copy_u4: 16-byte loads and stores) over two
 * scratch buffers of `bytes` each, `repeats` times, and wait.  A kernel of
 * known traffic, to scale hardware byte counters (FETCH_SIZE/WRITE_SIZE).
 * Allocates and frees its scratch; never call it in a timed region.  No
 * reference counterpart. */
int bdem_calibration_copy(bdem_ctx* ctx, unsigned long long bytes, int repeats);

/* ---- stage outputs for parity checks (synchronous; state unchanged) -------- */
/* Keys and identity index of the current state (Kernel_SpatialDecomp_CalcPHash). */
int bdem_debug_hash(bdem_ctx* ctx, unsigned int* keys, int* idx);
/* Keys and index after the stable sort (SpatialDecomp_ThrustSort_ByHash). */
int bdem_debug_sorted(bdem_ctx* ctx, unsigned int* keys, int* idx);
/* cellStart / cellEnd of the sorted state (Kernel_SpatialDecomp_BinData),
 * bdem_num_cells entries each, 0xFFFFFFFF for empty cells. */
int bdem_debug_cells(bdem_ctx* ctx, unsigned int* cell_start, unsigned int* cell_end);
/* Broad_List (32 per particle) and NN_NUM of the sorted state
 * (Kernel_BroadCollisionDetection_NonSymmetry). */
int bdem_debug_contacts(bdem_ctx* ctx, unsigned int* list32, unsigned int* nn);
/* One bdem_update_sim that also reports the contacts the step kernel used
 * (the reference's Broad_List / NN_NUM format: 32 slots per particle in
 * sorted order, full count), for parity checks of the skin lists.  No
 * reference counterpart. */
int bdem_debug_step_contacts(bdem_ctx* ctx, unsigned int* list32, unsigned int* nn);
/* Pair force (with gravity), pair torque, wall force, wall torque of the
 * sorted state, N x 3 floats each; sorted_pid receives the P_ID order. */
int bdem_debug_forces(bdem_ctx* ctx, float* fpp, float* tpp, float* fw, float* tw, int* sorted_pid);
/* Work of the last polyhedra step (synchronous; no reference counterpart):
 * counts[0] = listed pairs (bounding spheres overlapping, at most 32 per
 * particle), counts[1] = pairs the two facing faces did not separate, on
 * which the wave-per-pair kernel ran the full separating-plane test and,
 * where no face separated them, the vertex-face and edge-edge searches.
 * BDEM_EINVAL for a sphere context. */
int bdem_debug_poly_queues(bdem_ctx* ctx, unsigned int* counts);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* BDEM_ABI_H_ */
