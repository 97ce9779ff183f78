"""The adjoint of the drum step (include/bdem_adjoint.h) on the GPU.

A settled bed in a small rotating drum is run 20 steps; the gradient of a
seeded linear functional of the observables with respect to theta =
(friction, restitution, kn, rolling resistance) from the hand-written
adjoint (FP32) must match central finite differences of the FP64
restatement from the same state to 1e-4 relative, on a horizon whose
contact set the perturbations do not change (checked).  Also: the adjoint
run's forward is the plain step, and the torch.autograd.Function returns the
adjoint's gradient.
"""
import math

import numpy as np
import pytest

try:   # imported before any GPU work of this module
    import torch
except ImportError:   # pragma: no cover
    torch = None

import bdem
from bdem import adjoint as ADJ
from bdem import scenarios as S
from bdem import types as T
from helpers import by_pid, make_device
from oracle_py import Oracle

pytestmark = pytest.mark.gpu

THETA0 = np.array([0.4, 0.5, 1e5, 0.002])
# the finite-difference check: a rolling resistance that acts within its
# 20 steps, and relative steps that put each perturbation's effect well
# above the FP64 runs' rounding (kn dominates the loss; the friction term is
# linear in mu; rr scales the spin)
THETA_FD = np.array([0.4, 0.5, 1e5, 0.05])
FD_STEP = np.array([1e-3, 1e-4, 1e-4, 1e-3])
# Relative tolerance per parameter.  Over a horizon without contact changes
# the loss moves by ~1e-11 of itself per unit of relative rolling
# resistance, which FP64 central differences resolve only to ~4e-2 (their
# spread over steps 1e-5 .. 1e-3); the other three are resolved to ~1e-5
# at the steps used.
FD_TOL = np.array([1e-4, 1e-4, 1e-4, 1e-2])


def _settled_drum(n=600, R=4.0, L=3.0, settle=30000, seed=21):
    deck = S.drum_deck(n=n, seed=seed, R=R, L=L, rpm=30.0, kn=1e5, e=0.5, mu=0.4, roll_res=0.002)
    dev = make_device(deck)
    dev.drum_rotation(True)
    dev.step(settle)
    st = by_pid(dev.get_state())
    dev.close()
    d = dict(deck)
    d["pos"], d["vel"], d["avel"] = st["pos"], st["vel"], st["avel"]
    return d


def _pair_constants(pobj, theta, exact=False):
    """bdem_adj_set_params' derivation of the pair constants, in Python
    (exact: kn kept in double instead of rounded to the table's float)."""
    mu, e, kn, rr = theta
    m = 1.0 / (1.0 / pobj.mass + 1.0 / pobj.mass)
    es = e - 0.03808 if 0.30 <= e <= 0.90 else e
    kf = float(kn) if exact else float(np.float32(kn))
    l = math.log(es)
    cn = -2.0 * l * math.sqrt(kf * m) / math.sqrt(l * l + math.pi * math.pi)
    return kf, cn, mu, rr


def _with_theta(deck, theta):
    """The deck with theta's pair constants and rolling resistance."""
    d = dict(deck)
    p = T.POBJ.from_buffer_copy(bytes(deck["pobj"][0]))
    kf, cn, mu, rr = _pair_constants(p, theta)
    p.PP[0].Kn, p.PP[0].Cn, p.PP[0].Fric_kinetic = kf, cn, mu
    d["pobj"] = [p]
    sim = T.SimulationInfo.from_buffer_copy(bytes(deck["sim"]))
    sim.Roll_Res = rr
    d["sim"] = sim
    return d


def _oracle_obs(deck, theta, nsteps, window, bits=64):
    """FP64 restatement's observables with theta's constants in double."""
    d = _with_theta(deck, theta)
    ora = Oracle(d, bits=bits)
    ora.set_pair_override(*_pair_constants(deck["pobj"][0], theta, exact=True))
    ora.set_drum_rotation(True)
    mass = np.asarray(deck["mass"], np.float64)
    obs, lists = [], []
    for s in range(nsteps):
        ora.step(1)
        st = ora.state()
        m = mass[st["pid"]]
        x, y = st["pos"][:, 0], st["pos"][:, 1]
        M = m.sum()
        cx, cy = (m * x).sum() / M, (m * y).sum() / M
        cxx = (m * x * x).sum() / M - cx * cx
        cyy = (m * y * y).sum() / M - cy * cy
        cxy = (m * x * y).sum() / M - cx * cy
        obs.append((cx, cy, 0.5 * math.atan2(2 * cxy, cxx - cyy)))
        _, nn = ora.contacts()
        lists.append(nn.copy())
    ora.close()
    o = np.asarray(obs).reshape(nsteps // window, window, 3).mean(axis=1).reshape(-1)
    return o, lists


def _lively_drum(n=500, R=4.0, seed=21):
    """Neighbours overlapping by up to 0.02 cm (lattice pitch 0.97 x 2 rmax),
    the outer ring pressed 0.003 cm into the drum wall and the first layer
    0.03 cm into an end plate, velocities up to 5 cm/s and spins up to
    10 rad/s: contacts open and close by whole fractions of a step's motion.
    (A resting bed under the reference law's 1e-5 cm offset chatters at the
    contact threshold instead, and an FP32 and an FP64 run part ways there.)"""
    deck = S.drum_deck(n=n, seed=seed, R=R, L=3.0, rpm=30.0, kn=1e5, e=0.5, mu=0.4, roll_res=0.002,
                       vmax=5.0, spacing=0.97)
    pos = deck["pos"].astype(np.float64)
    r = deck["radius"].astype(np.float64)
    c = np.array([R, R])
    rho = np.linalg.norm(pos[:, :2] - c, axis=1)
    k = (R + 0.003 - r[np.argmax(rho + r)]) / rho.max()
    pos[:, :2] = c + (pos[:, :2] - c) * k
    pos[:, 2] -= pos[:, 2].min() - 0.18
    deck["pos"] = pos.astype(np.float32)
    deck["avel"] = S.uniform(seed, 7, 3 * n, -10.0, 10.0).reshape(n, 3).astype(np.float32)
    return deck


def _gpu_contact_totals(deck, nsteps):
    """Contacts the step kernel used at each step (plain context)."""
    dev = make_device(_with_theta(deck, THETA0))
    dev.drum_rotation(True)
    tot = []
    for _ in range(nsteps):
        _, nn = dev.debug_step_contacts()
        tot.append(int(nn.sum()))
    dev.close()
    return tot


def test_gradient_matches_fp64_central_differences():
    """A lively bed (lattice pitch below the diameters, velocities up to
    20 cm/s): contacts open and close by whole fractions of a step's motion,
    so the FP32 run and the FP64 runs see the same contact sets; a resting
    bed under the reference law's 1e-5 cm offset chatters at the contact
    threshold instead."""
    nsteps, window = 20, 5
    deck = _lively_drum()
    print(f"[case] adjoint vs FP64 FD: drum bed n={deck['sim'].Num_Particles}, {nsteps} steps", flush=True)
    dev = make_device(deck)
    dev.drum_rotation(True)
    run = ADJ.AdjointRun(dev, nsteps, window, checkpoint_every=10)
    obs = run.forward(THETA_FD)
    c = S.uniform(99, 0, len(obs), -1.0, 1.0)
    g_adj = run.backward(c)
    o64, lists0 = _oracle_obs(deck, THETA_FD, nsteps, window)
    print(f"observables GPU vs FP64 max abs diff {np.max(np.abs(obs - o64)):.3e}", flush=True)
    assert np.allclose(obs, o64, rtol=0, atol=1e-5)
    # the FP32 run and the FP64 runs cross no contact event differently: the
    # contacts each GPU step used (its input state) against the FP64 run's
    # contacts of the state each step produced
    gpu_tot = _gpu_contact_totals(deck, nsteps)
    print(f"contacts per step {gpu_tot}", flush=True)
    assert gpu_tot[1:] == [int(x.sum()) for x in lists0[:-1]]
    g_fd = np.zeros(4)
    for i in range(4):
        h = FD_STEP[i] * THETA_FD[i]
        tp, tm = THETA_FD.copy(), THETA_FD.copy()
        tp[i] += h
        tm[i] -= h
        op, lp = _oracle_obs(deck, tp, nsteps, window)
        om, lm = _oracle_obs(deck, tm, nsteps, window)
        for a, b in zip(lp, lm):
            assert np.array_equal(a, b), "the perturbation changed the contact set: shorten the horizon"
        g_fd[i] = (c @ op - c @ om) / (2 * h)
    scaled_adj, scaled_fd = g_adj * THETA_FD, g_fd * THETA_FD
    print(f"theta * dJ/dtheta  adjoint {scaled_adj}\n                   FP64 FD {scaled_fd}", flush=True)
    rel = np.abs(scaled_adj - scaled_fd) / np.maximum(np.abs(scaled_fd), 1e-12)
    print(f"relative error per parameter {rel}", flush=True)
    for i in range(4):
        assert rel[i] <= FD_TOL[i], (i, rel[i])
    run.close()
    dev.close()


def test_adjoint_forward_is_the_plain_step():
    deck = _settled_drum(n=400, settle=5000, seed=22)
    print("[case] adjoint forward vs plain steps, n=400, 60 steps", flush=True)
    dev = make_device(deck)
    dev.drum_rotation(True)
    run = ADJ.AdjointRun(dev, 60, 20, checkpoint_every=20)
    run.forward(THETA0)
    a = by_pid(dev.get_state())
    plain = make_device(_with_theta(deck, THETA0))
    plain.drum_rotation(True)
    plain.step(60)
    b = by_pid(plain.get_state())
    for k in ("pos", "vel", "avel"):
        assert np.array_equal(a[k], b[k]), k
    run.close()
    dev.close()
    plain.close()


@pytest.mark.skipif(torch is None, reason="torch is not importable")
def test_torch_function_returns_the_adjoint_gradient():
    deck = _settled_drum(n=400, settle=5000, seed=23)
    print("[case] torch.autograd.Function over the adjoint, n=400, 40 steps", flush=True)
    dev = make_device(deck)
    dev.drum_rotation(True)
    run = ADJ.AdjointRun(dev, 40, 10, checkpoint_every=20)
    target = torch.as_tensor(run.forward(THETA0 * np.array([1.1, 0.9, 1.2, 0.5])))
    Obs = ADJ.observables_function()
    theta = torch.tensor(THETA0, dtype=torch.float64, requires_grad=True)
    obs = Obs.apply(theta, run)
    loss = ((obs - target) ** 2).mean()
    loss.backward()
    g = run.backward((2.0 / len(obs)) * (obs.detach().numpy() - target.numpy()))
    print(f"loss {loss.item():.3e} grad {theta.grad.numpy()}", flush=True)
    assert np.all(np.isfinite(theta.grad.numpy()))
    assert np.allclose(theta.grad.numpy(), g, rtol=1e-12, atol=0)
    run.close()
    dev.close()
