"""Gradients of a drum run with respect to the contact parameters
(include/bdem_adjoint.h), and a torch.autograd.Function over them.

theta = (kinetic friction, restitution, normal stiffness, rolling
resistance).  The observables are, per window of steps, the window's mean
bed centre of mass (x, y) and bed inclination (radians).
"""
import ctypes as C

import numpy as np

from . import lib as L

_P = C.c_void_p
_dp = C.POINTER(C.c_double)

_PROTOS = {
    "bdem_adj_create": (C.c_int, [C.POINTER(_P), _P, C.c_int, C.c_int, C.c_int]),
    "bdem_adj_destroy": (C.c_int, [_P]),
    "bdem_adj_last_error": (C.c_char_p, [_P]),
    "bdem_adj_num_observables": (C.c_int, [_P]),
    "bdem_adj_set_params": (C.c_int, [_P, _dp]),
    "bdem_adj_forward": (C.c_int, [_P, _dp]),
    "bdem_adj_backward": (C.c_int, [_P, _dp, _dp]),
    "bdem_adj_timing": (C.c_int, [_P, _dp, _dp]),
}
EXPORTED = tuple(_PROTOS)


def _lib():
    lib = L.load()
    if not getattr(lib, "_adj_declared", False):
        for name, (res, args) in _PROTOS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        lib._adj_declared = True
    return lib


def _dptr(a):
    return a.ctypes.data_as(_dp)


class AdjointRun:
    """A checkpointed run of `nsteps` steps of a configured drum Device,
    started from the Device's current (uploaded, not yet stepped) state."""

    def __init__(self, dev, nsteps, window, checkpoint_every=100):
        self.lib = _lib()
        self.dev = dev          # the run steps this context; keep it alive
        self.h = _P()
        rc = self.lib.bdem_adj_create(C.byref(self.h), dev.h, int(nsteps), int(window), int(checkpoint_every))
        if rc != L.BDEM_OK:
            raise L.BdemError(f"bdem_adj_create failed ({rc})")
        self.nobs = int(self.lib.bdem_adj_num_observables(self.h))
        self.theta = None

    def _check(self, rc):
        if rc != L.BDEM_OK:
            raise L.BdemError(f"bdem adjoint error {rc}: {self.lib.bdem_adj_last_error(self.h).decode()}")

    def close(self):
        if self.h:
            self.lib.bdem_adj_destroy(self.h)
            self.h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, theta):
        th = np.ascontiguousarray(theta, np.float64).reshape(4)
        self._check(self.lib.bdem_adj_set_params(self.h, _dptr(th)))
        self.theta = th.copy()

    def forward(self, theta=None):
        if theta is not None:
            self.set_params(theta)
        obs = np.zeros(self.nobs, np.float64)
        self._check(self.lib.bdem_adj_forward(self.h, _dptr(obs)))
        return obs

    def backward(self, cot):
        cot = np.ascontiguousarray(cot, np.float64).reshape(self.nobs)
        g = np.zeros(4, np.float64)
        self._check(self.lib.bdem_adj_backward(self.h, _dptr(cot), _dptr(g)))
        return g

    def timing(self):
        f, b = C.c_double(0.0), C.c_double(0.0)
        self._check(self.lib.bdem_adj_timing(self.h, C.byref(f), C.byref(b)))
        return f.value, b.value


def observables_function():
    """torch.autograd.Function mapping theta (4, float64) to the run's
    observables; backward runs the adjoint of the forward it follows."""
    import torch

    class DrumObservables(torch.autograd.Function):
        @staticmethod
        def forward(ctx, theta, run):
            obs = run.forward(theta.detach().cpu().numpy().astype(np.float64))
            ctx.run = run
            ctx.theta = run.theta
            return torch.as_tensor(obs, dtype=torch.float64)

        @staticmethod
        def backward(ctx, grad_obs):
            run = ctx.run
            if run.theta is None or not np.array_equal(run.theta, ctx.theta):
                run.forward(ctx.theta)   # another forward ran since: restore this one's checkpoints
            g = run.backward(grad_obs.detach().cpu().numpy().astype(np.float64))
            return torch.as_tensor(g, dtype=torch.float64), None

    return DrumObservables
