"""ctypes binding of libbdem.so (include/bdem_abi.h).

`Device` mirrors the reference's KDevice facade (Interface/KDevice.h:17-99):
one object per simulation, constructed from the object tables, with
IDevice_*-style methods mapped onto the C ABI.  Arrays are numpy, in the
reference layout (AoS float3, Quaterion w,x,y,z).  There is no CPU fallback:
constructing a Device without the built library raises.
"""
import ctypes as C
import os

import numpy as np

from . import types as T

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG, "libbdem.so")

BDEM_OK = 0
WARN_OUT_OF_GRID = 1
WARN_LIST_OVERFLOW = 2
WARN_HIST_OVERFLOW = 4
WARN_SORT = 8

_lib = None

_P = C.c_void_p
_f3p = C.POINTER(T.Float3)
_qp = C.POINTER(T.Quaterion)
_up = C.POINTER(C.c_uint)
_ip = C.POINTER(C.c_int)
_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)

_PROTOS = {
    "bdem_create": (C.c_int, [C.POINTER(_P), C.c_int]),
    "bdem_destroy": (C.c_int, [_P]),
    "bdem_last_error": (C.c_char_p, [_P]),
    "bdem_set_simulation_data": (C.c_int, [_P, C.POINTER(T.WOBJ), C.POINTER(T.POBJ), C.POINTER(T.DOBJ),
                                           C.POINTER(T.SimulationInfo), C.c_int, C.c_int, C.c_int,
                                           C.POINTER(T.InitConfig)]),
    "bdem_set_ext_params": (C.c_int, [_P, C.POINTER(T.ExtParams)]),
    "bdem_update_simulation_info": (C.c_int, [_P, C.POINTER(T.SimulationInfo)]),
    "bdem_set_pos": (C.c_int, [_P, _f3p, _qp, _up, _ip]),
    "bdem_set_pos_velocity": (C.c_int, [_P, _f3p, _qp, _f3p, _f3p, _up, _ip]),
    "bdem_set_system_state": (C.c_int, [_P, _f3p, _qp, _f3p, _f3p, _f3p, _up, _ip]),
    "bdem_set_radii": (C.c_int, [_P, _fp, _fp]),
    "bdem_set_tallys": (C.c_int, [_P, _fp]),
    "bdem_get_system_state": (C.c_int, [_P, _f3p, _qp, _f3p, _f3p, _f3p, _up, _ip]),
    "bdem_get_positions": (C.c_int, [_P, _f3p, _qp, _up, _ip]),
    "bdem_get_velocity": (C.c_int, [_P, _f3p, _f3p, C.c_int, _up, _ip]),
    "bdem_get_wall_forces": (C.c_int, [_P, _f3p, _f3p]),
    "bdem_record_wall_forces": (C.c_int, [_P, C.c_int]),
    "bdem_get_energy": (C.c_int, [_P, _fp]),
    "bdem_get_status": (C.c_int, [_P, _up]),
    "bdem_get_particle_stats": (C.c_int, [_P, C.c_float, _f3p, _f3p, C.POINTER(T.ParticleStats)]),
    "bdem_get_kill_count": (C.c_int, [_P, _ip]),
    "bdem_get_surface_profile": (C.c_int, [_P, C.c_int, C.c_float, C.c_float, C.c_int, _fp]),
    "bdem_get_surface_profile_noalloc": (C.c_int, [_P, C.c_int, C.c_float, C.c_float, C.c_int, _fp]),
    "bdem_get_poly_vertices": (C.c_int, [_P, C.c_int, _f3p, _ip, _ip]),
    "bdem_set_poly_shapes": (C.c_int, [_P, C.POINTER(T.PolyShape), C.c_int]),
    "bdem_get_num_poly_shapes": (C.c_int, [_P, _ip, _ip]),
    "bdem_surface_angle": (C.c_int, [_fp, C.c_int, C.c_float, C.c_float, _dp, _dp, _dp]),
    "bdem_sort_changed": (C.c_int, [_P, _up]),
    "bdem_calibration_copy": (C.c_int, [_P, C.c_ulonglong, C.c_int]),
    "bdem_num_particles": (C.c_int, [_P]),
    "bdem_num_cells": (C.c_longlong, [_P]),
    "bdem_update_dobject_positions": (C.c_int, [_P, C.POINTER(T.DOBJ)]),
    "bdem_update_world_object_positions": (C.c_int, [_P, C.POINTER(T.WOBJ)]),
    "bdem_open_hatch": (C.c_int, [_P, C.POINTER(T.WOBJ), C.c_int]),
    "bdem_reset_energy": (C.c_int, [_P, C.c_int]),
    "bdem_drum_rotation": (C.c_int, [_P, C.c_int]),
    "bdem_start_bin": (C.c_int, [_P]),
    "bdem_add_force": (C.c_int, [_P]),
    "bdem_sub_force": (C.c_int, [_P]),
    "bdem_zero_force": (C.c_int, [_P]),
    "bdem_update_sim": (C.c_int, [_P]),
    "bdem_step": (C.c_int, [_P, C.c_int]),
    "bdem_set_graphs": (C.c_int, [_P, C.c_int]),
    "bdem_sync": (C.c_int, [_P]),
    "bdem_timing_begin": (C.c_int, [_P, C.c_int]),
    "bdem_timing_end": (C.c_int, [_P, C.POINTER(C.c_double), _ip]),
    "bdem_debug_hash": (C.c_int, [_P, _up, _ip]),
    "bdem_debug_sorted": (C.c_int, [_P, _up, _ip]),
    "bdem_debug_cells": (C.c_int, [_P, _up, _up]),
    "bdem_debug_contacts": (C.c_int, [_P, _up, _up]),
    "bdem_debug_forces": (C.c_int, [_P, _fp, _fp, _fp, _fp, _ip]),
    "bdem_debug_step_contacts": (C.c_int, [_P, _up, _up]),
    "bdem_debug_poly_queues": (C.c_int, [_P, _up]),
    "bdem_set_lifter_rotation": (C.c_int, [_P, C.c_double, C.c_float, C.c_float, C.c_float]),
    "bdem_get_dobjects": (C.c_int, [_P, C.POINTER(T.DOBJ), _ip]),
    "bdem_set_skin": (C.c_int, [_P, C.c_float]),
    "bdem_get_skin": (C.c_float, [_P]),
    "bdem_derive_spring_constants": (C.c_int, [C.POINTER(T.POBJ), C.c_int, C.c_int, _dp, _dp, _dp,
                                               C.c_double, C.c_double, C.c_double]),
    # include/bdem_deck.h
    "bdem_ksim_read": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int, _ip, _ip, _ip]),
    "bdem_deck_load": (C.c_int, [C.POINTER(_P), C.c_char_p, C.c_char_p, C.c_int, C.c_char_p, C.c_int]),
    "bdem_deck_free": (None, [_P]),
    "bdem_deck_get": (C.c_int, [_P, C.POINTER(C.POINTER(T.DeckInfo)), C.POINTER(C.POINTER(T.SimulationInfo)),
                                C.POINTER(C.POINTER(T.InitConfig)), C.POINTER(C.POINTER(T.WOBJ)),
                                C.POINTER(C.POINTER(T.POBJ)), C.POINTER(C.POINTER(T.DOBJ))]),
    "bdem_deck_setup": (C.c_int, [_P, _P]),
    "bdem_state_read": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(T.StateHeader),
                                  _f3p, _qp, _f3p, _f3p, _f3p, _up, _ip, _ip]),
    "bdem_state_write": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(T.StateHeader),
                                   _f3p, _qp, _f3p, _f3p, _f3p, _up, _ip, C.POINTER(T.POBJ), C.c_int]),
    "bdem_deck_load_state": (C.c_int, [_P, _P, C.c_char_p, C.POINTER(T.StateHeader)]),
    "bdem_deck_initial_state": (C.c_int, [_P, _f3p, _qp, _up, _ip]),
    "bdem_fill_random": (C.c_int, [C.c_int, _fp, T.Float3, T.Float3, _fp, C.c_float, C.c_ulonglong, C.c_int,
                                   _f3p, _ip]),
    "bdem_poly_shapes_from_points": (C.c_int, [C.c_int, _ip, _f3p, C.c_int, C.c_float, C.c_float,
                                               C.POINTER(T.PolyShape), _ip]),
    # include/bdem_batch.h
    "bdem_batch_create": (C.c_int, [C.POINTER(_P), C.c_int]),
    "bdem_batch_destroy": (C.c_int, [_P]),
    "bdem_batch_last_error": (C.c_char_p, [_P]),
    "bdem_batch_add": (C.c_int, [_P, _P, _ip]),
    "bdem_batch_set_tuning": (C.c_int, [_P, C.POINTER(T.BatchTuning)]),
    "bdem_batch_get_tuning": (C.c_int, [_P, C.POINTER(T.BatchTuning)]),
    "bdem_batch_finalize": (C.c_int, [_P]),
    "bdem_batch_step": (C.c_int, [_P, C.c_int]),
    "bdem_batch_sync": (C.c_int, [_P]),
    "bdem_batch_autotune": (C.c_int, [_P, C.c_int, C.POINTER(T.BatchTuning), _dp]),
    "bdem_batch_num_instances": (C.c_int, [_P]),
    "bdem_batch_num_particles": (C.c_int, [_P, C.c_int]),
    "bdem_batch_instance_sim": (C.c_int, [_P, C.c_int, C.POINTER(T.SimulationInfo)]),
    "bdem_batch_get_state": (C.c_int, [_P, C.c_int, _f3p, _f3p, _f3p, _ip]),
    "bdem_batch_get_status": (C.c_int, [_P, _up]),
}

SPRING_TC, SPRING_KN = 0, 1


def derive_spring_constants(pobj, mode, cor_pp, cor_surface, cor_dsurface, tc=(0.0, 0.0, 0.0)):
    """bdem_derive_spring_constants on a list of POBJ (modified in place;
    host only, no device needed)."""
    lib = load()
    arr = (T.POBJ * len(pobj))(*pobj)
    c = [np.ascontiguousarray(x, np.float64) for x in (cor_pp, cor_surface, cor_dsurface)]
    rc = lib.bdem_derive_spring_constants(arr, len(pobj), int(mode), *[_ptr(x, C.c_double) for x in c],
                                          float(tc[0]), float(tc[1]), float(tc[2]))
    if rc != 0:
        raise BdemError(f"bdem_derive_spring_constants: status {rc}")
    for k in range(len(pobj)):
        pobj[k] = arr[k]
    return pobj


def fill_random(radius, lo, hi, seed, gap=0.01, cyl=None, max_tries=100000):
    """bdem_fill_random (host only): a seeded random non-overlapping fill of
    spheres of these radii (placed in index order) in the box [lo, hi],
    optionally inside the vertical cylinder cyl = (x, z, R)."""
    lib = load()
    r = np.ascontiguousarray(radius, np.float32)
    n = len(r)
    pos = np.zeros((n, 3), np.float32)
    c = None if cyl is None else np.ascontiguousarray(cyl, np.float32)
    failed = C.c_int(-1)
    rc = lib.bdem_fill_random(n, _ptr(r, C.c_float), T.f3(*lo), T.f3(*hi), _ptr(c, C.c_float), float(gap),
                              int(seed) & 0xFFFFFFFFFFFFFFFF, int(max_tries), _ptr(pos, T.Float3), C.byref(failed))
    if rc != 0:
        raise BdemError(f"bdem_fill_random: status {rc}, sphere {failed.value} could not be placed")
    return pos


def poly_shapes_from_points(points, density, face_tol=0.1):
    """bdem_poly_shapes_from_points (host only): one convex-hull shape per
    point set.  points: a list of (k, 3) arrays (4 <= k <= 32) or one
    (count, max_pts, 3) array with counts in a second element, i.e.
    (pts, npts).  Returns a ctypes array of PolyShape."""
    lib = load()
    if isinstance(points, tuple):
        pts, npts = points
        pts = np.ascontiguousarray(pts, np.float32)
        npts = np.ascontiguousarray(npts, np.int32)
    else:
        m = max(len(p) for p in points)
        pts = np.zeros((len(points), m, 3), np.float32)
        npts = np.zeros(len(points), np.int32)
        for k, p in enumerate(points):
            pts[k, :len(p)] = p
            npts[k] = len(p)
    count, max_pts = pts.shape[0], pts.shape[1]
    out = (T.PolyShape * count)()
    failed = C.c_int(-1)
    rc = lib.bdem_poly_shapes_from_points(count, _ptr(npts, C.c_int), _ptr(pts, T.Float3), max_pts,
                                          float(density), float(face_tol), out, C.byref(failed))
    if rc != 0:
        raise BdemError(f"bdem_poly_shapes_from_points: status {rc}, shape {failed.value} is degenerate or too large")
    return out


def surface_angle(top, lo, hi):
    """bdem_surface_angle (host only): least-squares line through the
    non-empty bins of a surface profile -> (slope, intercept, angle in degrees)."""
    lib = load()
    top = np.ascontiguousarray(top, np.float32)
    sl, ic, ang = C.c_double(), C.c_double(), C.c_double()
    rc = lib.bdem_surface_angle(_ptr(top, C.c_float), len(top), float(lo), float(hi), C.byref(sl), C.byref(ic),
                                C.byref(ang))
    if rc != 0:
        raise BdemError(f"bdem_surface_angle: status {rc}")
    return sl.value, ic.value, ang.value

EXPORTED = tuple(_PROTOS)


def load(path=LIB_PATH):
    """Load libbdem.so and declare every prototype; raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"libbdem.so not built at {path}: run __graft_entry__.build()")
    lib = C.CDLL(path)
    for name, (res, args) in _PROTOS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class BdemError(RuntimeError):
    pass


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype)) if a is not None else None


def _f3(a):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a


def _shaped(a, dtype, shape, name):
    """a as a contiguous array of dtype and exactly this shape; the C side
    reads Num_Particles entries, so a short buffer must not reach it."""
    a = np.ascontiguousarray(a, dtype=dtype)
    if a.shape != tuple(shape):
        raise ValueError(f"{name} must have shape {tuple(shape)}, got {a.shape}")
    return a


class Device:
    """One simulation context on one GPU (KDevice, Interface/KDevice.h:17-99)."""

    def __init__(self, wobj, pobj, sim, init=None, dobj=None, device=0):
        self.lib = load()
        self.h = _P()
        rc = self.lib.bdem_create(C.byref(self.h), int(device))
        if rc != BDEM_OK:
            raise BdemError(f"bdem_create failed ({rc})")
        self._wobj = (T.WOBJ * max(len(wobj), 1))(*wobj)
        self._pobj = (T.POBJ * len(pobj))(*pobj)
        self._dobj = (T.DOBJ * max(len(dobj or []), 1))(*(dobj or []))
        self.sim = sim
        self.init = init if init is not None else T.InitConfig()
        self._check(self.lib.bdem_set_simulation_data(
            self.h, self._wobj, self._pobj, self._dobj, C.byref(sim), len(wobj), len(pobj),
            len(dobj or []), C.byref(self.init)))
        self.n = sim.Num_Particles

    def _check(self, rc):
        if rc != BDEM_OK:
            msg = self.lib.bdem_last_error(self.h).decode()
            raise BdemError(f"bdem error {rc}: {msg}")

    def close(self):
        if self.h:
            self.lib.bdem_destroy(self.h)
            self.h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- setup ------------------------------------------------------------
    def set_ext(self, ext):
        self._ext = ext
        self._check(self.lib.bdem_set_ext_params(self.h, C.byref(ext)))

    def update_sim_info(self, sim):
        self.sim = sim
        self._check(self.lib.bdem_update_simulation_info(self.h, C.byref(sim)))

    def set_state(self, pos, ptype, pid, vel=None, avel=None, quat=None):
        n = self.n
        pos = _shaped(pos, np.float32, (n, 3), "pos")
        ptype = _shaped(ptype, np.uint32, (n,), "ptype")
        pid = _shaped(pid, np.int32, (n,), "pid")
        q = None if quat is None else _shaped(quat, np.float32, (n, 4), "quat")
        if vel is not None:
            vel = _shaped(vel, np.float32, (n, 3), "vel")
        if avel is not None:
            avel = _shaped(avel, np.float32, (n, 3), "avel")
        if vel is None:
            self._check(self.lib.bdem_set_pos(self.h, _ptr(pos, T.Float3), _ptr(q, T.Quaterion),
                                              _ptr(ptype, C.c_uint), _ptr(pid, C.c_int)))
        else:
            vel = _f3(vel)
            avel = _f3(avel if avel is not None else np.zeros_like(vel))
            self._check(self.lib.bdem_set_pos_velocity(
                self.h, _ptr(pos, T.Float3), _ptr(q, T.Quaterion), _ptr(vel, T.Float3),
                _ptr(avel, T.Float3), _ptr(ptype, C.c_uint), _ptr(pid, C.c_int)))

    def set_radii(self, radius, mass):
        r = _shaped(radius, np.float32, (self.n,), "radius")
        m = _shaped(mass, np.float32, (self.n,), "mass")
        self._check(self.lib.bdem_set_radii(self.h, _ptr(r, C.c_float), _ptr(m, C.c_float)))

    # -- state out --------------------------------------------------------
    def get_state(self):
        n = self.n
        pos = np.zeros((n, 3), np.float32)
        vel = np.zeros((n, 3), np.float32)
        avel = np.zeros((n, 3), np.float32)
        quat = np.zeros((n, 4), np.float32)
        ptype = np.zeros(n, np.uint32)
        pid = np.zeros(n, np.int32)
        self._check(self.lib.bdem_get_system_state(
            self.h, _ptr(pos, T.Float3), _ptr(quat, T.Quaterion), _ptr(vel, T.Float3),
            _ptr(avel, T.Float3), None, _ptr(ptype, C.c_uint), _ptr(pid, C.c_int)))
        return dict(pos=pos, vel=vel, avel=avel, quat=quat, ptype=ptype, pid=pid)

    def particle_stats(self, min_y=-np.inf, box=None):
        """Live particles above min_y (count, mean and largest speed in m/s,
        P_ID of the fastest) and the solid fraction of box = (lo, hi)."""
        out = T.ParticleStats()
        lo = hi = None
        if box is not None:
            lo, hi = T.f3(*box[0]), T.f3(*box[1])
        self._check(self.lib.bdem_get_particle_stats(self.h, float(np.float32(min_y)) if np.isfinite(min_y) else
                                                 float(min_y), None if lo is None else C.byref(lo),
                                                 None if hi is None else C.byref(hi), C.byref(out)))
        return dict(count=out.count, max_pid=out.max_pid, mean_speed=out.mean_speed, max_speed=out.max_speed,
                    box_count=out.box_count, box_solid_fraction=out.box_solid_fraction)

    def poly_vertices(self, max_vertex=None):
        """World-frame vertices of every polyhedron: (verts [n, max_vertex, 3],
        num_vertex [n], pid [n]) in the library's order."""
        n = self.n
        if max_vertex is None:
            rows = self._shapes if getattr(self, "_shapes", None) is not None else self._pobj
            max_vertex = max(p.num_vertex for p in rows)
        verts = np.zeros((n, int(max_vertex), 3), np.float32)
        nv = np.zeros(n, np.int32)
        pid = np.zeros(n, np.int32)
        self._check(self.lib.bdem_get_poly_vertices(self.h, int(max_vertex), _ptr(verts, T.Float3),
                                                    _ptr(nv, C.c_int), _ptr(pid, C.c_int)))
        return verts, nv, pid

    def set_poly_shapes(self, shapes):
        """bdem_set_poly_shapes: per-particle shapes, row p for P_ID p (a
        ctypes array of PolyShape, e.g. from poly_shapes_from_points)."""
        self._check(self.lib.bdem_set_poly_shapes(self.h, shapes, len(shapes)))
        self._shapes = shapes

    def num_poly_shapes(self):
        """(rows, by_pid) of the shape table the polyhedra path reads."""
        cnt, by = C.c_int(), C.c_int()
        self._check(self.lib.bdem_get_num_poly_shapes(self.h, C.byref(cnt), C.byref(by)))
        return cnt.value, bool(by.value)

    def surface_profile(self, lo, hi, bins, axis=0):
        """Highest particle top y + r per bin of [lo, hi) along x (axis 0) or
        z (axis 2); NaN for empty bins."""
        top = np.zeros(int(bins), np.float32)
        self._check(self.lib.bdem_get_surface_profile(self.h, int(axis), float(lo), float(hi), int(bins),
                                                      _ptr(top, C.c_float)))
        return top

    def surface_profile_noalloc(self, lo, hi, bins, axis=0):
        """surface_profile through the allocation-free entry point."""
        top = np.zeros(int(bins), np.float32)
        self._check(self.lib.bdem_get_surface_profile_noalloc(self.h, int(axis), float(lo), float(hi), int(bins),
                                                              _ptr(top, C.c_float)))
        return top

    def status(self):
        w = C.c_uint(0)
        self._check(self.lib.bdem_get_status(self.h, C.byref(w)))
        return w.value

    def kill_count(self):
        k = C.c_int(0)
        self._check(self.lib.bdem_get_kill_count(self.h, C.byref(k)))
        return k.value

    def sort_changed(self):
        """Particles the last sort moved (changed list size, N after a full sort)."""
        k = C.c_uint(0)
        self._check(self.lib.bdem_sort_changed(self.h, C.byref(k)))
        return k.value

    def calibration_copy(self, nbytes, repeats=3):
        """Plain copy kernel over nbytes, `repeats` times (counter calibration)."""
        self._check(self.lib.bdem_calibration_copy(self.h, int(nbytes), int(repeats)))

    def set_tallys(self, tallys):
        """Device_Set_Tallys: the three dissipation tallies (unscaled)."""
        t = _shaped(tallys, np.float32, (3,), "tallys")
        self._check(self.lib.bdem_set_tallys(self.h, _ptr(t, C.c_float)))

    def energy(self):
        """Device_Get_Energy: {particle, lifter, surface} tallies times 1E-4."""
        e = np.zeros(3, np.float32)
        self._check(self.lib.bdem_get_energy(self.h, _ptr(e, C.c_float)))
        return e

    def set_lifter_rotation(self, rad_per_step, about=(0.0, 0.0, 0.0)):
        """Turn the drum's lifters by rad_per_step before every step, on the device."""
        self._check(self.lib.bdem_set_lifter_rotation(self.h, float(rad_per_step), *[float(a) for a in about]))

    def get_dobj(self):
        """The lifter table as it stands (after any turning on the device)."""
        k = C.c_int(0)
        self._check(self.lib.bdem_get_dobjects(self.h, None, C.byref(k)))
        arr = (T.DOBJ * max(k.value, 1))()
        self._check(self.lib.bdem_get_dobjects(self.h, arr, C.byref(k)))
        return list(arr)[:k.value]

    def reset_energy(self):
        """Device_Reset_Energy(0): zero the three tallies."""
        self._check(self.lib.bdem_reset_energy(self.h, 0))

    def update_dobj(self, dobj):
        """Upload moved lifters (Device_Update_DObjectPostions)."""
        arr = (T.DOBJ * len(dobj))(*dobj)
        self._dobj = arr
        self._check(self.lib.bdem_update_dobject_positions(self.h, arr))

    def num_cells(self):
        return int(self.lib.bdem_num_cells(self.h))

    # -- hot path ---------------------------------------------------------
    def update_sim(self):
        self._check(self.lib.bdem_update_sim(self.h))

    def step(self, nsteps):
        self._check(self.lib.bdem_step(self.h, int(nsteps)))

    def sync(self):
        self._check(self.lib.bdem_sync(self.h))

    def set_graphs(self, on):
        """Replay steady-state step pairs from a captured HIP graph (default on)."""
        self._check(self.lib.bdem_set_graphs(self.h, 1 if on else 0))

    def open_hatch(self):
        self._check(self.lib.bdem_open_hatch(self.h, None, 0))

    def drum_rotation(self, on):
        self._check(self.lib.bdem_drum_rotation(self.h, 1 if on else 0))

    # -- measurement ------------------------------------------------------
    STAGES = ("sort", "lists", "step")

    def timing_begin(self, max_steps):
        self._check(self.lib.bdem_timing_begin(self.h, int(max_steps)))

    def timing_end(self):
        ms = (C.c_double * len(self.STAGES))()
        n = C.c_int(0)
        self._check(self.lib.bdem_timing_end(self.h, ms, C.byref(n)))
        return dict(zip(self.STAGES, list(ms))), n.value

    # -- stage outputs ----------------------------------------------------
    def debug_hash(self):
        k = np.zeros(self.n, np.uint32)
        i = np.zeros(self.n, np.int32)
        self._check(self.lib.bdem_debug_hash(self.h, _ptr(k, C.c_uint), _ptr(i, C.c_int)))
        return k, i

    def debug_sorted(self):
        k = np.zeros(self.n, np.uint32)
        i = np.zeros(self.n, np.int32)
        self._check(self.lib.bdem_debug_sorted(self.h, _ptr(k, C.c_uint), _ptr(i, C.c_int)))
        return k, i

    def debug_cells(self):
        nc = self.num_cells()
        s = np.zeros(nc, np.uint32)
        e = np.zeros(nc, np.uint32)
        self._check(self.lib.bdem_debug_cells(self.h, _ptr(s, C.c_uint), _ptr(e, C.c_uint)))
        return s, e

    def debug_poly_queues(self):
        """(listed pairs, pairs left for the narrow search) of the last polyhedra step."""
        out = np.zeros(2, np.uint32)
        self._check(self.lib.bdem_debug_poly_queues(self.h, _ptr(out, C.c_uint)))
        return int(out[0]), int(out[1])

    def debug_contacts(self):
        lst = np.zeros(self.n * 32, np.uint32)
        nn = np.zeros(self.n, np.uint32)
        self._check(self.lib.bdem_debug_contacts(self.h, _ptr(lst, C.c_uint), _ptr(nn, C.c_uint)))
        return lst.reshape(self.n, 32), nn

    def debug_step_contacts(self):
        """One step, returning the contacts the step kernel used (32 slots per
        particle in sorted order, full counts)."""
        lst = np.zeros(self.n * 32, np.uint32)
        nn = np.zeros(self.n, np.uint32)
        self._check(self.lib.bdem_debug_step_contacts(self.h, _ptr(lst, C.c_uint), _ptr(nn, C.c_uint)))
        return lst.reshape(self.n, 32), nn

    def set_skin(self, skin):
        self._check(self.lib.bdem_set_skin(self.h, float(skin)))

    def skin(self):
        return float(self.lib.bdem_get_skin(self.h))

    def debug_forces(self):
        out = [np.zeros((self.n, 3), np.float32) for _ in range(4)]
        pid = np.zeros(self.n, np.int32)
        self._check(self.lib.bdem_debug_forces(self.h, *[_ptr(a, C.c_float) for a in out],
                                               _ptr(pid, C.c_int)))
        return out[0], out[1], out[2], out[3], pid


class Batch:
    """Many independent sphere simulations advanced together
    (include/bdem_batch.h).  Each is added from a configured Device whose
    state has been uploaded and not yet stepped; the Device is not used by
    the batch afterwards."""

    def __init__(self, device=0):
        self.lib = load()
        self.h = _P()
        rc = self.lib.bdem_batch_create(C.byref(self.h), int(device))
        if rc != BDEM_OK:
            raise BdemError(f"bdem_batch_create failed ({rc})")
        self.sizes = []

    def _check(self, rc):
        if rc != BDEM_OK:
            msg = self.lib.bdem_batch_last_error(self.h).decode()
            raise BdemError(f"bdem batch error {rc}: {msg}")

    def close(self):
        if self.h:
            self.lib.bdem_batch_destroy(self.h)
            self.h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, dev):
        k = C.c_int(-1)
        self._check(self.lib.bdem_batch_add(self.h, dev.h, C.byref(k)))
        self.sizes.append(dev.n)
        return k.value

    def set_tuning(self, threads=256, waves_per_cu=0, grab=1, cell_factor=1.0, use_graph=1):
        t = T.BatchTuning(int(threads), int(waves_per_cu), int(grab), float(cell_factor), int(use_graph))
        self._check(self.lib.bdem_batch_set_tuning(self.h, C.byref(t)))

    def tuning(self):
        t = T.BatchTuning()
        self._check(self.lib.bdem_batch_get_tuning(self.h, C.byref(t)))
        return dict(threads=t.threads, waves_per_cu=t.waves_per_cu, grab=t.grab, cell_factor=t.cell_factor,
                    use_graph=t.use_graph)

    def finalize(self):
        self._check(self.lib.bdem_batch_finalize(self.h))

    def step(self, nsteps):
        self._check(self.lib.bdem_batch_step(self.h, int(nsteps)))

    def sync(self):
        self._check(self.lib.bdem_batch_sync(self.h))

    def autotune(self, steps=10):
        t = T.BatchTuning()
        ms = C.c_double(0.0)
        self._check(self.lib.bdem_batch_autotune(self.h, int(steps), C.byref(t), C.byref(ms)))
        return dict(threads=t.threads, waves_per_cu=t.waves_per_cu, grab=t.grab, cell_factor=t.cell_factor,
                    use_graph=t.use_graph), ms.value

    def instance_sim(self, k):
        s = T.SimulationInfo()
        self._check(self.lib.bdem_batch_instance_sim(self.h, int(k), C.byref(s)))
        return s

    def get_state(self, k):
        n = self.sizes[k]
        pos = np.zeros((n, 3), np.float32)
        vel = np.zeros((n, 3), np.float32)
        avel = np.zeros((n, 3), np.float32)
        pid = np.zeros(n, np.int32)
        self._check(self.lib.bdem_batch_get_state(self.h, int(k), _ptr(pos, T.Float3), _ptr(vel, T.Float3),
                                                  _ptr(avel, T.Float3), _ptr(pid, C.c_int)))
        return dict(pos=pos, vel=vel, avel=avel, pid=pid)

    def status(self):
        w = C.c_uint(0)
        self._check(self.lib.bdem_batch_get_status(self.h, C.byref(w)))
        return w.value
