"""Project decks and state files (include/bdem_deck.h).

`Deck` reads a reference project (World/, SurfaceObjects/, ParticleObjects/,
LifterObjects/ or VolumeObjects/ and its start files) with the library's C++
reader, the replacement of KSimulationData (Input/KSimulationData.cpp) and
Set_SimInfo (main.cpp:469-833).  `read_state` / `write_state` are
Read_FromFile / Output_SystemState (Input/FileIO.h).  Everything here is
host code: no device is needed until `Deck.device()`.
"""
import ctypes as C
import os

import numpy as np

from . import lib as L
from . import types as T


def _copy(ptr, n, cls):
    return [cls.from_buffer_copy(ptr[k]) for k in range(n)]


def read_ksim(path):
    """Simulation.KSim (main.cpp:843-848) -> (project, world, mode2d, device, num_devices)."""
    lib = L.load()
    p, w = C.create_string_buffer(256), C.create_string_buffer(256)
    m, d, nd = C.c_int(), C.c_int(), C.c_int()
    rc = lib.bdem_ksim_read(os.fsencode(path), p, w, 256, C.byref(m), C.byref(d), C.byref(nd))
    if rc != 0:
        raise L.BdemError(f"cannot read {path}")
    return p.value.decode(), w.value.decode(), m.value, d.value, nd.value


class Deck:
    """One project deck, as the reference reads it."""

    def __init__(self, project_dir, world, mode2d=0):
        self.lib = L.load()
        self.dir = os.fspath(project_dir)
        self.h = L._P()
        err = C.create_string_buffer(512)
        rc = self.lib.bdem_deck_load(C.byref(self.h), os.fsencode(self.dir), world.encode(), int(mode2d), err, 512)
        if rc != 0:
            raise L.BdemError(f"deck {world}: {err.value.decode()}")
        self.world = world
        self._refresh()

    def _refresh(self):
        info = C.POINTER(T.DeckInfo)()
        sim = C.POINTER(T.SimulationInfo)()
        init = C.POINTER(T.InitConfig)()
        w = C.POINTER(T.WOBJ)()
        p = C.POINTER(T.POBJ)()
        d = C.POINTER(T.DOBJ)()
        self.lib.bdem_deck_get(self.h, C.byref(info), C.byref(sim), C.byref(init), C.byref(w), C.byref(p),
                               C.byref(d))
        self.info = T.DeckInfo.from_buffer_copy(info[0])
        self.sim = T.SimulationInfo.from_buffer_copy(sim[0])
        self.init = T.InitConfig.from_buffer_copy(init[0])
        self.wobj = _copy(w, self.info.num_wobj, T.WOBJ)
        self.pobj = _copy(p, self.info.num_pobj, T.POBJ)
        self.dobj = _copy(d, self.info.num_dobj, T.DOBJ) if self.info.num_dobj else []

    def close(self):
        if self.h:
            self.lib.bdem_deck_free(self.h)
            self.h = L._P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def start_path(self):
        return os.path.join(self.dir, self.info.start_file.decode()) if self.info.read_file == 1 else None

    def device(self, device=0):
        """A context set up from the deck's tables (the KDevice constructor)."""
        return L.Device(self.wobj, self.pobj, self.sim, init=self.init, dobj=self.dobj or None, device=device)

    def load_state(self, dev, path=None):
        """Read a state file (default: the deck's start file) into `dev` as
        Read_FromFile does; returns the header.  May turn this deck's lifters
        (a mill file of an unfinished revolution)."""
        hdr = T.StateHeader()
        rc = self.lib.bdem_deck_load_state(self.h, dev.h, None if path is None else os.fsencode(path),
                                           C.byref(hdr))
        if rc != 0:
            raise L.BdemError(f"bdem_deck_load_state: status {rc}: {dev.lib.bdem_last_error(dev.h).decode()}")
        self._refresh()
        return hdr

    def read_state(self, path=None):
        path = self.start_path if path is None else path
        return read_state(path, self.info.sim_type, self.info.rotating_world, self.info.particle_type,
                          self.sim.Num_Particles)

    def initial_state(self):
        """The grid fill of a deck without a start file (bdem_deck_initial_state):
        positions, orientations, types and IDs; every velocity is the deck's
        Velocity line and every spin zero, as Device_Set_SimulationData sets them."""
        n = self.sim.Num_Particles
        pos = np.zeros((n, 3), np.float32)
        quat = np.tile(np.array([1, 0, 0, 0], np.float32), (n, 1))
        ptype = np.zeros(n, np.uint32)
        pid = np.zeros(n, np.int32)
        rc = self.lib.bdem_deck_initial_state(self.h, L._ptr(pos, T.Float3), L._ptr(quat, T.Quaterion),
                                              L._ptr(ptype, C.c_uint), L._ptr(pid, C.c_int))
        if rc != 0:
            raise L.BdemError(f"bdem_deck_initial_state: status {rc}")
        v0 = self.init.velocity[0]
        vel = np.tile(np.array([v0.x, v0.y, v0.z], np.float32), (n, 1))
        hdr = T.StateHeader()
        return dict(header=hdr, rows=n, pos=pos, vel=vel, avel=np.zeros((n, 3), np.float32),
                    acc=np.zeros((n, 3), np.float32), quat=quat, ptype=ptype, pid=pid)

    def scenario(self, path=None):
        """The deck with its start state (or grid fill) in the dict form the
        parity tests and the oracle take (sim, wobj, pobj, dobj, pos, vel,
        avel, ptype, pid, quat), plus the mill rotation (lifter_about,
        rad_per_step)."""
        st = self.read_state(path) if (path is not None or self.info.read_file == 1) else self.initial_state()
        n = self.sim.Num_Particles
        if st["rows"] != n:
            raise L.BdemError(f"state file holds {st['rows']} of {n} rows")
        d = dict(name=self.world, sim=self.sim, wobj=self.wobj, pobj=self.pobj, ext=None,
                 pos=st["pos"], vel=st["vel"], avel=st["avel"], ptype=st["ptype"], pid=st["pid"],
                 quat=st["quat"] if self.info.particle_type == 1 else None, radius=None, mass=None,
                 header=st["header"])
        if self.dobj:
            d["dobj"] = self.dobj
        if self.info.sim_type == T.SIM_BALLMILL:
            a = self.info.rotate_about
            d["lifter_about"] = (a.x, a.y, a.z)
            d["rad_per_step"] = self.info.rad_per_step
        return d


def run_silo(dev, deck, seconds, open_hatch=True):
    """The silo branch of UpdateSim (main.cpp:85-98) with Silo_Stats
    (main.cpp:253-285): open the hatch, then every flow_steps steps count the
    particles left above it and record (time s, percent discharged).
    Returns the list of records.  `deck` is the Deck the context came from."""
    info = deck.info
    if info.sim_type != 2:
        raise L.BdemError("not a silo deck")
    if open_hatch:
        dev.open_hatch()
    dt = deck.sim.InitalDelta_t
    flow = max(int(info.flow_steps), 1)
    total = int(round(seconds / dt))
    n0 = dev.particle_stats(info.hatch_height)["count"]
    out, done = [], 0
    while done < total:
        k = min(flow, total - done)
        dev.step(k)
        done += k
        rem = dev.particle_stats(info.hatch_height)["count"]
        out.append((done * dt, 100.0 * (n0 - rem) / n0 if n0 else 0.0))
    return out


def read_state(path, sim_type, rotating_world, particle_type, n):
    lib = L.load()
    hdr = T.StateHeader()
    pos = np.zeros((n, 3), np.float32)
    vel = np.zeros((n, 3), np.float32)
    avel = np.zeros((n, 3), np.float32)
    acc = np.zeros((n, 3), np.float32)
    quat = np.tile(np.array([1, 0, 0, 0], np.float32), (n, 1))
    ptype = np.zeros(n, np.uint32)
    pid = np.zeros(n, np.int32)
    rows = C.c_int()
    rc = lib.bdem_state_read(os.fsencode(path), int(sim_type), int(rotating_world), int(particle_type), int(n),
                             C.byref(hdr), L._ptr(pos, T.Float3), L._ptr(quat, T.Quaterion), L._ptr(vel, T.Float3),
                             L._ptr(avel, T.Float3), L._ptr(acc, T.Float3), L._ptr(ptype, C.c_uint),
                             L._ptr(pid, C.c_int), C.byref(rows))
    if rc != 0:
        raise L.BdemError(f"cannot read state file {path}")
    k = rows.value
    return dict(header=hdr, rows=k, pos=pos[:k], vel=vel[:k], avel=avel[:k], acc=acc[:k], quat=quat[:k],
                ptype=ptype[:k], pid=pid[:k])


def write_state(path, header, sim_type, rotating_world, particle_type, pos, ptype, pid, vel=None, avel=None,
                acc=None, quat=None, pobj=None):
    lib = L.load()
    n = len(pos)
    arrs = [None if a is None else np.ascontiguousarray(a, np.float32) for a in (pos, vel, avel, acc, quat)]
    ptype = np.ascontiguousarray(ptype, np.uint32)
    pid = np.ascontiguousarray(pid, np.int32)
    parr = (T.POBJ * len(pobj))(*pobj) if pobj else None
    rc = lib.bdem_state_write(os.fsencode(path), int(sim_type), int(rotating_world), int(particle_type), n,
                              C.byref(header), L._ptr(arrs[0], T.Float3), L._ptr(arrs[4], T.Quaterion),
                              L._ptr(arrs[1], T.Float3), L._ptr(arrs[2], T.Float3), L._ptr(arrs[3], T.Float3),
                              L._ptr(ptype, C.c_uint), L._ptr(pid, C.c_int), parr, len(pobj) if pobj else 0)
    if rc != 0:
        raise L.BdemError(f"cannot write state file {path}")
