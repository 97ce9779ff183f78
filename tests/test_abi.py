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
HEADERS = [os.path.join(ROOT, "include", h) for h in ("bdem_abi.h", "bdem_deck.h")]


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


def test_library_exports_every_declared_entry_point():
    names = header_functions()
    assert len(names) >= 40
    lib = ctypes.CDLL(bdem.lib.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding declares exactly the header's entry points
    assert sorted(bdem.lib.EXPORTED) == names


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
