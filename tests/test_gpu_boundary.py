"""The drop-in boundary on the GPU: the reference-named Device_* API driving
a deck end to end, and the status-code contract of the C ABI (a rejected
upload leaves the context as it was)."""
import os
import subprocess

import numpy as np
import pytest

import bdem
from bdem import scenarios as S
from bdem import types as T
from helpers import make_device, make_pair, by_pid, rel_err

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
COMPAT_RUN = os.path.join(ROOT, "softx-d-15-00085_amd", "compat_run")


def _states_equal(a, b):
    a, b = by_pid(a), by_pid(b)
    assert np.array_equal(a["pid"], b["pid"])
    for k in ("pos", "vel", "avel"):
        assert np.array_equal(a[k], b[k]), k


def test_rejected_upload_leaves_the_context_usable():
    """bdem_set_pos with a P_ID beyond the history table returns BDEM_EINVAL
    before anything reaches the device; stepping on afterwards gives exactly
    the run that never saw the call."""
    deck = S.random_deck(2000, seed=41, L=10.0, law="lsd")
    print("[case] rejected upload: lsd random pack n=2000", flush=True)
    a = make_device(deck)
    a.step(3)
    bad_pid = deck["pid"].copy()
    bad_pid[17] = deck["sim"].Num_Particles + 5
    with pytest.raises(bdem.BdemError, match="P_ID"):
        a.set_state(deck["pos"] + 1.0, deck["ptype"], bad_pid, vel=deck["vel"])
    bad_type = deck["ptype"].copy()
    bad_type[3] = 7
    with pytest.raises(bdem.BdemError, match="type"):
        a.set_state(deck["pos"] + 1.0, bad_type, deck["pid"], vel=deck["vel"])
    a.step(5)
    b = make_device(deck)
    b.step(8)
    _states_equal(a.get_state(), b.get_state())
    a.close()
    b.close()


def test_non_unit_quaternion_is_taken_as_given():
    """The reference turns vertices by q v q* as the state gives q (its
    unit-test start files hold (1, 0.5, 0, 0)), so a non-unit quaternion is
    accepted; the separating-plane shortcut's vertex bound grows with |q|^2
    and the step equals the restatement's, which has no shortcut."""
    deck = S.poly_deck(200, seed=43, L=10.0)
    deck["quat"] = deck["quat"].copy()
    deck["quat"][::7] *= 1.3
    print("[case] non-unit quaternions: 200 cubes, every 7th |q| = 1.3", flush=True)
    dev, ora = make_pair(deck)
    lg, ng = dev.debug_contacts()
    lo, no = ora.contacts()
    assert np.array_equal(ng, no)
    dev.step(5)
    ora.step(5)
    a, b = by_pid(dev.get_state()), by_pid(ora.state())
    for k in ("vel", "avel", "quat"):
        assert rel_err(a[k], b[k], 1e-3) <= 1e-5, k
    dev.close()
    ora.close()


def _write_dump(path, deck, steps, hatch_at=-1):
    n = deck["sim"].Num_Particles
    quat = np.zeros((n, 4), np.float32)
    quat[:, 0] = 1.0
    with open(path, "wb") as f:
        f.write(b"BDEMDUMP")
        f.write(np.array([len(deck["wobj"]), len(deck["pobj"]), 0, steps, hatch_at], np.int32).tobytes())
        f.write(bytes(deck["sim"]))
        f.write(bytes(T.InitConfig()))
        for w in deck["wobj"]:
            f.write(bytes(w))
        for p in deck["pobj"]:
            f.write(bytes(p))
        f.write(np.ascontiguousarray(deck["pos"], np.float32).tobytes())
        f.write(quat.tobytes())
        f.write(np.ascontiguousarray(deck["vel"], np.float32).tobytes())
        f.write(np.ascontiguousarray(deck["avel"], np.float32).tobytes())
        f.write(np.ascontiguousarray(deck["ptype"], np.uint32).tobytes())
        f.write(np.ascontiguousarray(deck["pid"], np.int32).tobytes())


def _read_dump(path, n):
    raw = open(path, "rb").read()
    off = 0

    def take(dt, cols):
        nonlocal off
        cnt = n * cols
        a = np.frombuffer(raw, dt, cnt, off).reshape(n, cols) if cols > 1 else np.frombuffer(raw, dt, cnt, off)
        off += cnt * np.dtype(dt).itemsize
        return a.copy()

    pos = take(np.float32, 3)
    quat = take(np.float32, 4)
    vel = take(np.float32, 3)
    avel = take(np.float32, 3)
    ptype = take(np.uint32, 1)
    pid = take(np.int32, 1)
    return dict(pos=pos, quat=quat, vel=vel, avel=avel, ptype=ptype, pid=pid)


@pytest.mark.parametrize("hatch_at", [-1, 10])
def test_reference_named_api_drives_a_deck(tmp_path, hatch_at):
    """A main.cpp-style program (compat_run) calls Device_Set_SimulationData,
    Device_Set_System_State, Device_DEM_UpdateSim x 25 (Device_OpenHatch
    before step 10 in the second case) and Device_Get_System_State through
    the reference-named shim; its state equals the C-ABI run bit for bit."""
    assert os.path.exists(COMPAT_RUN), "build the library first (make -C softx-d-15-00085_amd/csrc)"
    deck = S.random_deck(1500, seed=44, L=10.0)
    steps = 25
    print(f"[case] compat_run: reference-law pack n=1500, {steps} steps, hatch_at={hatch_at}", flush=True)
    fin, fout = tmp_path / "in.dump", tmp_path / "out.dump"
    _write_dump(fin, deck, steps, hatch_at)
    r = subprocess.run([COMPAT_RUN, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = _read_dump(fout, deck["sim"].Num_Particles)
    dev = make_device(deck)
    for s in range(steps):
        if s == hatch_at:
            dev.open_hatch()
        dev.update_sim()
    want = dev.get_state()
    dev.close()
    _states_equal(got, want)
    if hatch_at >= 0:
        assert "Hatch Opened" in r.stdout
