"""ctypes mirrors of include/bdem_types.h (the reference's boundary PODs).

Field order and sizes follow DataStructures/KSimulationObjects.h:104-381 and
Utilities/Common/Types.h:15-21; the sizes are checked against the C header's
static_asserts by the tests.
"""
import ctypes as C


class Float3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class Int3(C.Structure):
    _fields_ = [("x", C.c_int), ("y", C.c_int), ("z", C.c_int)]


class Quaterion(C.Structure):
    _fields_ = [("w", C.c_float), ("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class SimulationInfo(C.Structure):
    _fields_ = [
        ("particle_type", C.c_int),
        ("Num_Particles", C.c_int),
        ("Num_ParticleObjects", C.c_int),
        ("Num_ParticlesPerObject", C.POINTER(C.c_int)),
        ("Num_WorldObjects", C.c_int),
        ("TotnumCells", C.c_int),
        ("Rotation", C.c_bool),
        ("Mode2D", C.c_bool),
        ("EnergyCalc", C.c_bool),
        ("isMillSim", C.c_bool),
        ("Kill_Particles", C.c_bool),
        ("InitalDelta_t", C.c_float),
        ("MillRotVel", C.c_float),
        ("Global_VisDamp", C.c_float),
        ("Roll_Res", C.c_float),
        ("Silo_Kill_Height", C.c_float),
        ("Vel_Limit", C.c_float),
        ("DObject_CylinderBoundR", C.c_float),
        ("num_NNCells", Int3),
        ("Force_Field", Float3),
        ("worldOrigin", Float3),
        ("cellSize", Float3),
        ("cent", Float3),
        ("Vel_Limit_Bounds", Float3),
        ("max_size", Float3),
        ("Simulation_Type", C.c_int),
        ("Integration_Type", C.c_int),
        ("use_symmetry", C.c_bool),
        ("sphere_orient", C.c_bool),
        ("dobject_hasVertex", C.c_bool),
        ("Rotating_WObject", C.c_bool),
        ("use_hist", C.c_bool),
        ("unit_test", C.c_int),
    ]


class InitConfig(C.Structure):
    _fields_ = [
        ("num", Int3),
        ("space", Float3),
        ("start", Float3),
        ("p_size", Float3 * 16),
        ("is_spheres", C.c_bool),
        ("use_file", C.c_bool),
        ("grid_type", C.c_int),
        ("launch_Vel", Float3),
        ("fill_plane_start", Float3),
        ("fill_plane_end", Float3),
        ("velocity", Float3 * 12),
        ("position", Float3 * 12),
        ("ang_vel", Float3 * 12),
        ("DebugFlag", C.c_bool),
        ("use_device", C.c_int),
        ("multi_gpu", C.c_bool),
        ("threads_perBlock", C.c_int),
    ]


class EdgeDef(C.Structure):
    _fields_ = [("point_vindex", C.c_uint * 2)]


class Surface(C.Structure):
    _fields_ = [
        ("num_vertex", C.c_int),
        ("vertex_Order", C.c_int * 8),
        ("area", C.c_float),
        ("normal", Float3),
        ("centroid", Float3),
        ("Tol", C.c_float),
    ]


class PPParms(C.Structure):
    _fields_ = [("Kn", C.c_float), ("Cn", C.c_float), ("Fric_static", C.c_float),
                ("Fric_kinetic", C.c_float)]


SHAPE_MAX_VERTEX, SHAPE_MAX_FACE, SHAPE_MAX_EDGE = 32, 32, 64


class PolyShape(C.Structure):
    """bdem_poly_shape: one convex polyhedron's geometry (per-particle shapes)."""
    _fields_ = [
        ("radius", C.c_float),
        ("mass", C.c_float),
        ("COM", Float3),
        ("num_vertex", C.c_int),
        ("Local_Vertex", Float3 * SHAPE_MAX_VERTEX),
        ("num_edges", C.c_int),
        ("edge", EdgeDef * SHAPE_MAX_EDGE),
        ("num_faces", C.c_int),
        ("face", Surface * SHAPE_MAX_FACE),
        ("InertiaT", C.c_float * 9),
        ("vbound", C.c_float),
        ("plane_slack", C.c_float),
    ]


class POBJ(C.Structure):
    _fields_ = [
        ("radius", C.c_float),
        ("mass", C.c_float),
        ("COM", Float3),
        ("num_vertex", C.c_int),
        ("Local_Vertex", Float3 * 32),
        ("num_edges", C.c_int),
        ("edge", EdgeDef * 32),
        ("num_faces", C.c_int),
        ("face", Surface * 32),
        ("InertiaT", C.c_float * 9),
        ("vertex", Float3 * 32),
        ("PP", PPParms * 6),
        ("surface_Kn", C.c_float),
        ("surface_Cn", C.c_float),
        ("surface_Fric_static", C.c_float),
        ("surface_Fric_kinetic", C.c_float),
        ("Dsurface_Kn", C.c_float),
        ("Dsurface_Cn", C.c_float),
        ("Dsurface_Fric_static", C.c_float),
        ("Dsurface_Fric_kinetic", C.c_float),
    ]


class DOBJ(C.Structure):
    _fields_ = [
        ("btype", C.c_uint),
        ("Axis", Float3),
        ("boundR", C.c_float),
        ("COM", Float3),
        ("num_vertex", C.c_int),
        ("vertex", Float3 * 8),
        ("num_faces", C.c_int),
        ("faces", Surface * 3),
        ("num_edges", C.c_int),
        ("edge", EdgeDef * 4),
        ("velocity", Float3),
        ("is_attached_DWorld", C.c_bool),
        ("is_translating", C.c_bool),
    ]


class MacroCylinder(C.Structure):
    _fields_ = [
        ("height", C.c_float),
        ("radius", C.c_float),
        ("has_top_cap", C.c_bool),
        ("has_bot_cap", C.c_bool),
        ("normal_cyl", Float3),
        ("normal_top_cap", Float3),
        ("normal_bot_cap", Float3),
        ("center_top_cap", Float3),
        ("center_bot_cap", Float3),
        ("AXIS", Float3),
    ]


class WOBJ(C.Structure):
    _fields_ = [
        ("surface_type", C.c_int),
        ("num_surfaces", C.c_int),
        ("surfaces", Surface * 16),
        ("num_vertex", C.c_int),
        ("vertex", Float3 * 16),
        ("is_rotating", C.c_bool),
        ("cyl_geo", MacroCylinder),
    ]


LAW_REFERENCE = 0
LAW_LSD_COULOMB = 1
LAW_HERTZ_MINDLIN = 2


class Material(C.Structure):
    _fields_ = [("kn", C.c_float), ("kt_over_kn", C.c_float), ("restitution", C.c_float),
                ("mu", C.c_float), ("youngs", C.c_float), ("poisson", C.c_float),
                ("pad0", C.c_float), ("pad1", C.c_float)]


class Cone(C.Structure):
    _fields_ = [("base", Float3), ("axis", Float3), ("s0", C.c_float), ("s1", C.c_float),
                ("r0", C.c_float), ("r1", C.c_float)]


MAX_CONES = 4


class ExtParams(C.Structure):
    _fields_ = [
        ("contact_law", C.c_int),
        ("per_particle_radius", C.c_int),
        ("num_cones", C.c_int),
        ("kill_check_interval", C.c_int),
        ("material", Material * 8),
        ("cone", Cone * MAX_CONES),
        ("wall_youngs", C.c_float),
        ("wall_poisson", C.c_float),
        ("pad", C.c_float * 2),
    ]


SURFACE_PLANE = 0
SURFACE_CYLINDER = 1

SIM_GENERAL, SIM_BALLMILL, SIM_SILO = 0, 1, 2


class DeckInfo(C.Structure):
    """bdem_deck_info (include/bdem_deck.h)."""
    _fields_ = [
        ("title", C.c_char * 128),
        ("force_spec", C.c_int), ("sim_type", C.c_int), ("particle_type", C.c_int),
        ("dt_in_setup", C.c_int),
        ("num_wobj", C.c_int), ("num_pobj", C.c_int), ("num_dobj", C.c_int),
        ("dobj_types", C.c_int), ("dobj_repeat", C.c_int),
        ("num_per_pobj", C.c_int * 8),
        ("read_file", C.c_int),
        ("start_file", C.c_char * 256),
        ("render", C.c_int), ("color_type", C.c_int),
        ("color_bin", C.c_float),
        ("ortho", C.c_int), ("fps", C.c_int), ("gl_debug", C.c_int), ("write_steps", C.c_int),
        ("output_format", C.c_int),
        ("total_time", C.c_float),
        ("is_run_limit", C.c_int),
        ("run_time", C.c_float),
        ("energy_steps", C.c_int), ("wall_steps", C.c_int),
        ("power_revs", C.c_int), ("total_revs", C.c_int),
        ("mill_rpm", C.c_float),
        ("steps_per_rev", C.c_int),
        ("rad_per_step", C.c_double),
        ("vel_mill", C.c_float),
        ("rotate_about", Float3),
        ("manual_hatch", C.c_int), ("measure_flow", C.c_int), ("flow_steps", C.c_int),
        ("kill_particle", C.c_int),
        ("hatch_height", C.c_float),
        ("rotating_world", C.c_int),
    ]


class ParticleStats(C.Structure):
    """bdem_particle_stats (include/bdem_abi.h)."""
    _fields_ = [
        ("count", C.c_int),
        ("max_pid", C.c_int),
        ("mean_speed", C.c_double),
        ("max_speed", C.c_float),
        ("box_count", C.c_int),
        ("box_solid_fraction", C.c_double),
    ]


class StateHeader(C.Structure):
    """bdem_state_header (include/bdem_deck.h)."""
    _fields_ = [
        ("spec_type", C.c_int),
        ("time", C.c_double),
        ("rev_num", C.c_int),
        ("rev_complete", C.c_int),
        ("rot_angle", C.c_double),
        ("energy", C.c_float * 5),
        ("num_rem", C.c_int),
        ("rot_steps", C.c_int),
    ]

EXPECTED_SIZES = {
    "SimulationInfo": 176, "InitConfig": 720, "Surface": 68, "POBJ": 3396, "DOBJ": 392,
    "MacroCylinder": 84, "WOBJ": 1380,
}


def f3(x, y, z):
    return Float3(float(x), float(y), float(z))


class BatchTuning(C.Structure):
    """bdem_batch_tuning (include/bdem_batch.h)."""
    _fields_ = [("threads", C.c_int), ("waves_per_cu", C.c_int), ("grab", C.c_int), ("cell_factor", C.c_float),
                ("use_graph", C.c_int)]
