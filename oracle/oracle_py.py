"""ctypes wrapper of the CPU restatement (libbdem_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by the tests, the smoke check and the
CPU-baseline leg of the benchmark, never by the product.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbdem_oracle.so")

_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"oracle not built at {LIB_PATH}: run make -C oracle")
    lib = C.CDLL(LIB_PATH)
    P = C.c_void_p
    fp = C.POINTER(C.c_float)
    dp = C.POINTER(C.c_double)
    up = C.POINTER(C.c_uint)
    ip = C.POINTER(C.c_int)
    lib.bo_create.restype = P
    lib.bo_create.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p]
    lib.bo_destroy.argtypes = [P]
    lib.bo_set_state.argtypes = [P, C.c_int, fp, fp, fp, fp, up, ip, fp, fp]
    lib.bo_get_state.argtypes = [P, dp, dp, dp, dp, up, ip]
    lib.bo_num_cells.restype = C.c_longlong
    lib.bo_num_cells.argtypes = [P]
    lib.bo_hash.argtypes = [P, up, ip]
    lib.bo_sorted.argtypes = [P, up, ip]
    lib.bo_cells.argtypes = [P, up, up]
    lib.bo_contacts.argtypes = [P, up, up]
    lib.bo_forces.argtypes = [P, dp, dp, dp, dp, ip]
    lib.bo_step.argtypes = [P, C.c_int]
    lib.bo_time_steps.restype = C.c_double
    lib.bo_time_steps.argtypes = [P, C.c_int]
    lib.bo_set_active_world_objects.argtypes = [P, C.c_int]
    lib.bo_set_drum_rotation.argtypes = [P, C.c_int]
    lib.bo_set_pair_override.argtypes = [P, C.c_double, C.c_double, C.c_double, C.c_double]
    lib.bo_set_poly_shapes.argtypes = [P, C.c_void_p, C.c_int]
    lib.bo_set_dobj.argtypes = [P, C.c_void_p, C.c_int]
    lib.bo_lifter_types.argtypes = [P, ip]
    lib.bo_set_lifter_types.argtypes = [P, ip]
    lib.bo_warnings.argtypes = [P]
    lib.bo_history.argtypes = [P, C.c_int, ip, dp]
    lib.bo_poly_pair.argtypes = [P, C.c_int, C.c_int, dp]
    lib.bo_angular_acc.argtypes = [P, C.c_int, dp, dp]
    lib.bo_poly_walls.argtypes = [P, C.c_int, dp]
    lib.bo_get_dobj.argtypes = [P, C.c_void_p]
    lib.bo_lifter_contacts.argtypes = [P]
    lib.bo_lifter_forces.argtypes = [P, dp, dp, ip]
    lib.bo_turn_lifters.argtypes = [P, C.c_double, C.c_float, C.c_float, C.c_float]
    lib.bo_run.argtypes = [P, C.c_int, C.c_double, C.c_float, C.c_float, C.c_float]
    lib.bo_energy.argtypes = [P, fp, dp]
    lib.bo_set_tallys.argtypes = [P, fp]
    lib.bo_history_law.argtypes = [P, C.c_int, C.c_int, C.c_int, dp, C.c_double, dp, C.c_double, C.c_double,
                                   dp, dp]
    _lib = lib
    return lib


def _p(a, t):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


class Oracle:
    """CPU restatement of one simulation in float (bits=32) or double (bits=64)."""

    def __init__(self, deck, bits=32):
        import ctypes
        self.lib = load()
        self.deck = deck
        self._w = (deck["wobj"][0].__class__ * len(deck["wobj"]))(*deck["wobj"])
        self._p = (deck["pobj"][0].__class__ * len(deck["pobj"]))(*deck["pobj"])
        ext = deck.get("ext")
        self.h = self.lib.bo_create(bits, ctypes.byref(deck["sim"]), self._w, len(deck["wobj"]),
                                    self._p, len(deck["pobj"]),
                                    ctypes.byref(ext) if ext is not None else None)
        self.n = deck["sim"].Num_Particles
        self.bits = bits
        if deck.get("dobj"):
            self.set_dobj(deck["dobj"])
        if deck.get("shapes") is not None:
            self.set_poly_shapes(deck["shapes"])
        self.set_state(deck)

    def set_state(self, deck):
        f = lambda a: None if a is None else np.ascontiguousarray(a, np.float32)
        self._keep = [f(deck["pos"]), f(deck["vel"]), f(deck.get("avel")), f(deck.get("quat")),
                      np.ascontiguousarray(deck["ptype"], np.uint32),
                      np.ascontiguousarray(deck["pid"], np.int32), f(deck.get("radius")),
                      f(deck.get("mass"))]
        k = self._keep
        self.lib.bo_set_state(self.h, self.n, _p(k[0], C.c_float), _p(k[1], C.c_float),
                              _p(k[2], C.c_float), _p(k[3], C.c_float), _p(k[4], C.c_uint),
                              _p(k[5], C.c_int), _p(k[6], C.c_float), _p(k[7], C.c_float))

    def close(self):
        if self.h:
            self.lib.bo_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def state(self):
        n = self.n
        pos = np.zeros((n, 3)); vel = np.zeros((n, 3)); avel = np.zeros((n, 3)); quat = np.zeros((n, 4))
        pt = np.zeros(n, np.uint32); pid = np.zeros(n, np.int32)
        self.lib.bo_get_state(self.h, _p(pos, C.c_double), _p(vel, C.c_double), _p(avel, C.c_double),
                              _p(quat, C.c_double), _p(pt, C.c_uint), _p(pid, C.c_int))
        return dict(pos=pos, vel=vel, avel=avel, quat=quat, ptype=pt, pid=pid)

    def num_cells(self):
        return int(self.lib.bo_num_cells(self.h))

    def hash(self):
        k = np.zeros(self.n, np.uint32); i = np.zeros(self.n, np.int32)
        self.lib.bo_hash(self.h, _p(k, C.c_uint), _p(i, C.c_int))
        return k, i

    def sorted(self):
        k = np.zeros(self.n, np.uint32); i = np.zeros(self.n, np.int32)
        self.lib.bo_sorted(self.h, _p(k, C.c_uint), _p(i, C.c_int))
        return k, i

    def cells(self):
        nc = self.num_cells()
        s = np.zeros(nc, np.uint32); e = np.zeros(nc, np.uint32)
        self.lib.bo_cells(self.h, _p(s, C.c_uint), _p(e, C.c_uint))
        return s, e

    def contacts(self):
        l = np.zeros(self.n * 32, np.uint32); nn = np.zeros(self.n, np.uint32)
        self.lib.bo_contacts(self.h, _p(l, C.c_uint), _p(nn, C.c_uint))
        return l.reshape(self.n, 32), nn

    def forces(self):
        out = [np.zeros((self.n, 3)) for _ in range(4)]
        pid = np.zeros(self.n, np.int32)
        self.lib.bo_forces(self.h, *[_p(a, C.c_double) for a in out], _p(pid, C.c_int))
        return out[0], out[1], out[2], out[3], pid

    def lifter_forces(self):
        """Lifter force and torque of the current state, by P_ID."""
        fl, tl = np.zeros((self.n, 3)), np.zeros((self.n, 3))
        pid = np.zeros(self.n, np.int32)
        self.lib.bo_lifter_forces(self.h, _p(fl, C.c_double), _p(tl, C.c_double), _p(pid, C.c_int))
        order = np.argsort(pid, kind="stable")
        return fl[order], tl[order]

    def step(self, n=1):
        self.lib.bo_step(self.h, int(n))

    def time_steps(self, n):
        return float(self.lib.bo_time_steps(self.h, int(n)))

    def set_dobj(self, dobj):
        """Lifter table (list of DOBJ), as Device_Update_DObjectPostions uploads it."""
        arr = (dobj[0].__class__ * len(dobj))(*dobj) if dobj else None
        self._d = arr
        self.lib.bo_set_dobj(self.h, arr, len(dobj) if dobj else 0)

    def lifter_contacts(self):
        """Polyhedra that touched a lifter in the last step."""
        return int(self.lib.bo_lifter_contacts(self.h))

    def get_dobj(self):
        """The lifter table as it stands (after any turning)."""
        d = self.deck["dobj"]
        arr = (d[0].__class__ * len(d))()
        self.lib.bo_get_dobj(self.h, arr)
        return list(arr)

    def turn_lifters(self, rad, about):
        """Rotate_VolObject_Inc once."""
        self.lib.bo_turn_lifters(self.h, float(rad), *[float(a) for a in about])

    def run(self, nsteps, rad=0.0, about=(0.0, 0.0, 0.0)):
        """nsteps steps, each after turning the lifters by rad (the mill loop of UpdateSim)."""
        self.lib.bo_run(self.h, int(nsteps), float(rad), *[float(a) for a in about])

    def energy(self):
        """({particle, lifter, surface} tallies x 1E-4 as GetEnergy returns them,
        the same increments summed in double, unscaled)."""
        e = np.zeros(3, np.float32)
        x = np.zeros(3, np.float64)
        self.lib.bo_energy(self.h, _p(e, C.c_float), _p(x, C.c_double))
        return e, x

    def set_tallys(self, t):
        t = np.ascontiguousarray(t, np.float32)
        self.lib.bo_set_tallys(self.h, _p(t, C.c_float))

    def lifter_types(self):
        t = np.zeros(self.n, np.int32)
        self.lib.bo_lifter_types(self.h, _p(t, C.c_int))
        return t

    def set_lifter_types(self, t):
        t = np.ascontiguousarray(t, np.int32)
        self.lib.bo_set_lifter_types(self.h, _p(t, C.c_int))

    def set_pair_override(self, kn, cn, fk, rr):
        """Exact pair constants of type pair (0, 0) and rolling resistance."""
        self.lib.bo_set_pair_override(self.h, float(kn), float(cn), float(fk), float(rr))

    def set_poly_shapes(self, shapes):
        """Per-particle polyhedron shapes (a ctypes array of bdem_poly_shape):
        row p for P_ID p."""
        self._shapes = shapes
        self.lib.bo_set_poly_shapes(self.h, C.cast(shapes, C.c_void_p), len(shapes))

    def set_drum_rotation(self, on):
        self.lib.bo_set_drum_rotation(self.h, 1 if on else 0)

    def set_active_world_objects(self, k):
        self.lib.bo_set_active_world_objects(self.h, int(k))

    def warnings(self):
        return int(self.lib.bo_warnings(self.h))

    def history(self, pid):
        ids = np.zeros(16, np.int32); xi = np.zeros((16, 3))
        c = self.lib.bo_history(self.h, int(pid), _p(ids, C.c_int), _p(xi, C.c_double))
        return ids[:max(c, 0)], xi[:max(c, 0)]

    def poly_pair(self, a, b):
        """Contact of slot a with slot b and the force/torque on a (fresh
        handles: slots are input order until the first stage runs)."""
        out = np.zeros(16)
        if self.lib.bo_poly_pair(self.h, int(a), int(b), _p(out, C.c_double)) != 0:
            raise ValueError("poly_pair needs a polyhedra deck and valid slots")
        return dict(hit=bool(out[0]), cls=int(out[1]), dis=out[2], n=out[3:6], cp=out[6:9],
                    F=out[9:12], T=out[12:15])

    def angular_acc(self, i, tau):
        t = np.ascontiguousarray(tau, np.float64)
        out = np.zeros(3)
        if self.lib.bo_angular_acc(self.h, int(i), _p(t, C.c_double), _p(out, C.c_double)) != 0:
            raise ValueError("angular_acc needs a polyhedra deck")
        return out

    def poly_walls(self, i):
        out = np.zeros(6)
        if self.lib.bo_poly_walls(self.h, int(i), _p(out, C.c_double)) != 0:
            raise ValueError("poly_walls needs a polyhedra deck")
        return out[:3], out[3:]

    def history_law(self, n, delta, vc, rstar, mstar, xi=(0.0, 0.0, 0.0), ta=0, tb=0, wall=False):
        """One evaluation of the history law (pair constants of types ta, tb,
        or the wall constants of ta): returns (force, spring kept)."""
        nn = np.ascontiguousarray(n, np.float64)
        v = np.ascontiguousarray(vc, np.float64)
        x = np.array(xi, np.float64)
        F = np.zeros(3)
        rc = self.lib.bo_history_law(self.h, 1 if wall else 0, int(ta), int(tb), _p(nn, C.c_double), float(delta),
                                     _p(v, C.c_double), float(rstar), float(mstar), _p(x, C.c_double),
                                     _p(F, C.c_double))
        if rc != 0:
            raise ValueError("history_law needs a history-law deck and types in [0, 8)")
        return F, x
