"""CPU checks of the boundary: struct layouts, exported symbols, host logic."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import bdem
from bdem import scenarios as S
from bdem import types as T

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("bdem_abi.h", "bdem_deck.h", "bdem_batch.h", "bdem_adjoint.h")]


def header_functions():
    text = "".join(open(h).read() for h in HEADERS)
    return sorted(set(re.findall(r"^\s*(?:int|void|float|const char\*|long long)\s+(bdem_\w+)\s*\(", text, re.M)))


def test_struct_sizes_match_reference_layout():
    for name, size in T.EXPECTED_SIZES.items():
        assert ctypes.sizeof(getattr(T, name)) == size, name
    assert ctypes.sizeof(T.Quaterion) == 16
    assert T.SimulationInfo.InitalDelta_t.offset == 40
    assert T.SimulationInfo.num_NNCells.offset == 68
    assert T.POBJ.PP.offset == 3268
    assert T.WOBJ.cyl_geo.offset == 1296
    assert ctypes.sizeof(T.BatchTuning) == 20
    assert ctypes.sizeof(T.PolyShape) == 3148
    assert T.PolyShape.face.offset == 928


def test_library_exports_every_declared_entry_point():
    names = header_functions()
    assert len(names) >= 40
    lib = ctypes.CDLL(bdem.lib.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding declares exactly the header's entry points
    from bdem import adjoint
    assert sorted(bdem.lib.EXPORTED + adjoint.EXPORTED) == names


def test_compat_shim_exports_reference_names():
    lib = ctypes.CDLL(bdem.lib.LIB_PATH)
    out = subprocess.run(["nm", "-D", "--defined-only", bdem.lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    for ref_name in ("Device_Set_SimulationData", "Device_DEM_UpdateSim", "Device_Get_System_State",
                     "Device_Set_Pos", "Device_Set_Pos_Velocity", "Device_Set_System_State",
                     "Device_Get_P_Positions", "Device_Get_P_Velocity", "Device_Get_W_Forces",
                     "Device_Get_Energy", "Device_OpenHatch", "Device_Clean", "Device_Reset_Energy",
                     "Device_Update_DObjectPostions", "Device_Update_WorldObjectPostions",
                     "Device_Set_Tallys", "drum_rotation", "Start_Bin"):
        assert re.search(r"\b_Z\d+" + ref_name, out), ref_name
    del lib


def test_creating_a_context_without_gpu_fails_cleanly():
    lib = bdem.lib.load()
    h = ctypes.c_void_p()
    rc = lib.bdem_create(ctypes.byref(h), 0)
    if rc == 0:  # a GPU is present
        lib.bdem_destroy(h)
        pytest.skip("GPU present")
    assert rc in (-1, -2)
    assert lib.bdem_update_sim(None) == -1


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_silo_fill_is_non_overlapping_and_inside():
    d = S.silo_deck(20000)
    p = d["pos"].astype(np.float64)
    g = S.silo_geometry()
    rho = np.hypot(p[:, 0] - S.SILO_AXIS_XZ, p[:, 2] - S.SILO_AXIS_XZ)
    assert p[:, 1].min() > S.SILO_Y0
    assert np.all(rho + 0.4 < S.SILO_R)
    hop = p[:, 1] < g["y1"]
    rc = S.SILO_OUTLET_R + (p[hop, 1] - S.SILO_Y0) * np.tan(np.radians(S.SILO_HALF_ANGLE))
    assert np.all(rho[hop] < rc)
    # lattice spacing exceeds the largest diameter: no initial overlaps
    from scipy.spatial import cKDTree
    dmin, _ = cKDTree(p).query(p, k=2)
    assert dmin[:, 1].min() >= 0.8
    assert d["radius"].max() <= 0.4 and d["radius"].min() >= 0.2


def test_splitmix_stream_is_deterministic_and_uniform():
    a = S.uniform(3, S.STREAM_RADIUS, 100000)
    b = S.uniform(3, S.STREAM_RADIUS, 100000)
    assert np.array_equal(a, b)
    assert 0.49 < a.mean() < 0.51 and a.min() >= 0.0 and a.max() < 1.0
    assert not np.array_equal(a, S.uniform(3, S.STREAM_VEL, 100000))


KDEVICE_CALLER = r"""
// A caller written against the reference's device API (Interface/KDevice.h
// style): it must compile and link against libbdem.so unchanged.
#include "bdem_device_compat.h"
class KDevice {
 public:
  KDevice(WOBJ* w, POBJ* p, DOBJ* d, int nw, int np, int nd, SimulationInfo* s, InitConfig* c) {
    Device_Set_SimulationData(w, p, d, s, nw, np, nd, c);
  }
  void IDevice_DEM_UpdateSim() { Device_DEM_UpdateSim(); }
  void IDevice_Get_System_State(float3* p, Quaterion* q, float3* v, float3* w, float3* a, uint* t, int* id) {
    Device_Get_System_State(p, q, v, w, a, t, id);
  }
  void IDevice_OpenHatch(WOBJ* w, int n) { Device_OpenHatch(w, n); }
  void clean() { Device_Clean(); }
};
int main(int argc, char**) {
  if (argc > 5) {  // never executed: link check only
    KDevice k(nullptr, nullptr, nullptr, 0, 0, 0, nullptr, nullptr);
    k.IDevice_DEM_UpdateSim();
    k.IDevice_Get_System_State(nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
    k.IDevice_OpenHatch(nullptr, 0);
    k.clean();
    Device_Set_Pos(nullptr, nullptr, nullptr, nullptr);
    drum_rotation(true);
  }
  return 0;
}
"""


def test_reference_style_caller_compiles_and_links(tmp_path):
    src = tmp_path / "caller.cpp"
    src.write_text(KDEVICE_CALLER)
    pkg = os.path.dirname(bdem.lib.LIB_PATH)
    exe = tmp_path / "caller"
    subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                    "-L", pkg, "-lbdem", f"-Wl,-rpath,{pkg}"], check=True, capture_output=True)
    assert exe.exists()


def test_binding_rejects_misshapen_state_before_calling_the_library():
    # no context is needed: the checks run before any call into C
    dev = object.__new__(bdem.Device)
    dev.n = 4
    dev.h = None
    ok = dict(pos=np.zeros((4, 3)), ptype=np.zeros(4), pid=np.arange(4))
    with pytest.raises(ValueError, match="pos"):
        dev.set_state(np.zeros((3, 3)), ok["ptype"], ok["pid"])
    with pytest.raises(ValueError, match="ptype"):
        dev.set_state(ok["pos"], np.zeros(5), ok["pid"])
    with pytest.raises(ValueError, match="pid"):
        dev.set_state(ok["pos"], ok["ptype"], np.arange(3))
    with pytest.raises(ValueError, match="quat"):
        dev.set_state(ok["pos"], ok["ptype"], ok["pid"], quat=np.zeros((4, 3)))
    with pytest.raises(ValueError, match="vel"):
        dev.set_state(ok["pos"], ok["ptype"], ok["pid"], vel=np.zeros((2, 3)))
    with pytest.raises(ValueError, match="avel"):
        dev.set_state(ok["pos"], ok["ptype"], ok["pid"], vel=np.zeros((4, 3)), avel=np.zeros(12))
    with pytest.raises(ValueError, match="radius"):
        dev.set_radii(np.ones(3), np.ones(4))
    with pytest.raises(ValueError, match="mass"):
        dev.set_radii(np.ones(4), np.ones((4, 1)))
    with pytest.raises(ValueError, match="tallys"):
        dev.set_tallys(np.zeros(2))


def test_random_fill_is_seeded_and_non_overlapping():
    """bdem_fill_random: every sphere inside the box (and cylinder), no pair
    closer than r_i + r_j + gap (all pairs, brute force), the same seed gives
    the same fill and another seed a different one."""
    rad = S.uniform(3, S.STREAM_RADIUS, 600, 0.2, 0.6).astype(np.float32)
    lo, hi = (0.0, 0.0, 0.0), (12.0, 14.0, 12.0)
    a = bdem.lib.fill_random(rad, lo, hi, seed=77, gap=0.01, cyl=(6.0, 6.0, 5.9))
    b = bdem.lib.fill_random(rad, lo, hi, seed=77, gap=0.01, cyl=(6.0, 6.0, 5.9))
    c = bdem.lib.fill_random(rad, lo, hi, seed=78, gap=0.01, cyl=(6.0, 6.0, 5.9))
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    p = a.astype(np.float64)
    r = rad.astype(np.float64)
    assert np.all(p - r[:, None] >= -1e-6) and np.all(p + r[:, None] <= np.array(hi) + 1e-6)
    assert np.all(np.hypot(p[:, 0] - 6.0, p[:, 2] - 6.0) + r <= 5.9 + 1e-5)
    d = np.linalg.norm(p[:, None, :] - p[None, :, :], axis=2)
    need = r[:, None] + r[None, :] + 0.01
    np.fill_diagonal(d, np.inf)
    assert np.all(d >= need - 1e-5)


def test_random_fill_reports_an_impossible_fill():
    rad = np.full(50, 1.0, np.float32)
    with pytest.raises(bdem.BdemError, match="could not be placed"):
        bdem.lib.fill_random(rad, (0, 0, 0), (4, 4, 4), seed=1, max_tries=200)


def test_batch_generator_follows_the_configuration():
    decks = S.batch_box_decks(count=4, seed=4, nmin=1024, nmax=3000)
    for d in decks:
        n = d["sim"].Num_Particles
        assert 1024 <= n <= 3000
        r = d["radius"]
        assert r.min() >= 0.2 and r.max() <= 0.6
        prm = d["params"]
        assert 5e4 <= prm["kn"] <= 5e5 and 0.3 <= prm["e"] <= 0.9 and 0.2 <= prm["mu"] <= 0.8
        vol = (4.0 / 3.0) * np.pi * np.sum(r.astype(np.float64) ** 3)
        assert 0.2 < vol / prm["L"] ** 3 < 0.4
        assert d["sim"].cellSize.x == pytest.approx(2.0 * r.max())


def test_batch_without_gpu_fails_cleanly():
    lib = bdem.lib.load()
    h = ctypes.c_void_p()
    rc = lib.bdem_batch_create(ctypes.byref(h), 0)
    if rc == 0:
        lib.bdem_batch_destroy(h)
        pytest.skip("a GPU is present")
    assert rc != 0 and not h.value
