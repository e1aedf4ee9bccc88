"""Differentiation form: gradients of a fixed-dt rollout of the CTU+CT step
through the hand-written adjoint kernels (include/athmi_adjoint.h), exposed
to PyTorch as a torch.autograd.Function.

The rollout is: set_bvals, then N times (integrate_2d with the fixed dt,
set_bvals) -- the main loop of main.c:221-318 with new_dt bypassed. States
are float64 CUDA tensors of shape (11, Nx2+8, Nx1+8) holding the Grid fields
(d, M1, M2, M3, E, B1c, B2c, B3c, B1i, B2i, B3i) including ghost zones.

    h = Rollout(desc, dt=2e-4, nsteps=256, every=16)
    sN = rollout(s0, h)                 # differentiable
    loss = 0.5 * ((sN[0] - target) ** 2)[interior].sum()
    loss.backward()                     # s0.grad via the adjoint

Import this module (it imports torch and initialises its HIP runtime) before
the first use of libathmi.so in a process: torch ships its own HIP runtime
under the same soname as the one the library links, and whichever is loaded
first serves the whole process.
"""
import ctypes

import numpy as np
import torch

from . import AthmiError, _Desc, _check, lib, FLUXES, VARIANTS

_D, _I, _P = ctypes.c_double, ctypes.c_int, ctypes.c_void_p


def _bind():
    if not torch.cuda.is_available():
        raise AthmiError("torch sees no GPU: import athena_cuda_amd.adjoint (which imports "
                         "torch) before the first use of libathmi.so in this process")
    L = lib()
    if not getattr(L, "_adjoint_bound", False):
        L.athmi_adjoint_create.argtypes = [ctypes.POINTER(_Desc), _D, _I, _I,
                                           ctypes.POINTER(_P)]
        L.athmi_adjoint_destroy.argtypes = [_P]
        L.athmi_adjoint_set_stream.argtypes = [_P, _P]
        L.athmi_adjoint_forward.argtypes = [_P, _P, _P]
        L.athmi_adjoint_backward.argtypes = [_P, _P, _P]
        L.athmi_adjoint_generation.argtypes = [_P, ctypes.POINTER(ctypes.c_ulonglong)]
        L.athmi_adjoint_set_deterministic.argtypes = [_P, _I]
        L.athmi_adjoint_every.argtypes = [_P, ctypes.POINTER(_I)]
        for n in ("create", "destroy", "set_stream", "forward", "backward", "generation",
                  "set_deterministic", "every"):
            getattr(L, "athmi_adjoint_" + n).restype = _I
        L._adjoint_bound = True
    return L


class Rollout:
    """Device handle of one fixed-dt rollout with sqrt(N) checkpointing."""

    def __init__(self, desc, dt, nsteps, every=None, deterministic=False):
        """every: checkpoint spacing; None lets the library choose (one
        segment, every step stored and none re-run, when that fits in the
        free device memory, else sqrt(nsteps))."""
        L = _bind()
        every = int(every) if every else 0
        bc = desc.get("bc", (4, 4, 4, 4))
        flux = desc.get("flux", "roe")
        d = _Desc(int(desc["nx1"]), int(desc["nx2"]), desc["x1min"], desc["x1max"],
                  desc["x2min"], desc["x2max"], desc["gamma"], desc.get("cour_no", 0.4),
                  *[int(b) for b in bc], FLUXES[flux] if isinstance(flux, str) else int(flux),
                  VARIANTS["exact"])
        h = _P()
        _check(L.athmi_adjoint_create(ctypes.byref(d), dt, nsteps, every, ctypes.byref(h)),
               "athmi_adjoint_create")
        self._h = h
        if deterministic:
            _check(L.athmi_adjoint_set_deterministic(h, 1), "athmi_adjoint_set_deterministic")
        self.deterministic = bool(deterministic)
        self.desc = dict(desc)
        k = ctypes.c_int()
        _check(L.athmi_adjoint_every(h, ctypes.byref(k)), "athmi_adjoint_every")
        self.dt, self.nsteps, self.every = dt, nsteps, k.value
        self.shape = (11, int(desc["nx2"]) + 8, int(desc["nx1"]) + 8)

    def close(self):
        if getattr(self, "_h", None):
            lib().athmi_adjoint_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _stream(self):
        # the handle runs on its own stream: wait for torch's pending work on
        # the inputs; the library synchronises its stream before returning
        import torch
        torch.cuda.current_stream().synchronize()

    def _check_tensor(self, t):
        import torch
        if not (t.is_cuda and t.dtype == torch.float64 and tuple(t.shape) == self.shape):
            raise AthmiError(f"expected a float64 CUDA tensor of shape {self.shape}")
        return t.contiguous()

    def forward(self, state0):
        import torch
        s0 = self._check_tensor(state0)
        out = torch.empty_like(s0)
        self._stream()
        _check(lib().athmi_adjoint_forward(self._h, _P(s0.data_ptr()), _P(out.data_ptr())),
               "athmi_adjoint_forward")
        return out

    def generation(self):
        """Completed forward passes of this handle; its checkpoints belong to
        the latest one."""
        gen = ctypes.c_ulonglong()
        _check(lib().athmi_adjoint_generation(self._h, ctypes.byref(gen)),
               "athmi_adjoint_generation")
        return gen.value

    def backward(self, grad_n):
        import torch
        g = self._check_tensor(grad_n)
        out = torch.empty_like(g)
        self._stream()
        _check(lib().athmi_adjoint_backward(self._h, _P(g.data_ptr()), _P(out.data_ptr())),
               "athmi_adjoint_backward")
        return out


def _function():
    import torch

    class RolloutFn(torch.autograd.Function):
        # The handle keeps the checkpoints of its latest forward pass only, so
        # a graph whose forward pass was followed by another on the same
        # handle cannot be reversed: its backward raises instead of reversing
        # the other trajectory.
        @staticmethod
        def forward(ctx, state0, handle):
            out = handle.forward(state0)
            ctx.handle = handle
            ctx.generation = handle.generation()
            return out

        @staticmethod
        def backward(ctx, grad_n):
            if ctx.handle.generation() != ctx.generation:
                raise AthmiError(
                    "rollout backward: another forward pass ran on this Rollout handle after "
                    "the one being differentiated (the handle keeps only the latest "
                    "checkpoints); use one handle per pending graph")
            return ctx.handle.backward(grad_n), None

    return RolloutFn


_FN = None


def rollout(state0, handle):
    """Differentiable rollout: state after handle.nsteps fixed-dt steps. The
    backward pass must come before the next rollout on the same handle (it
    raises AthmiError otherwise)."""
    global _FN
    if _FN is None:
        _FN = _function()
    return _FN.apply(state0, handle)


# ---------------------------------------------------------------------------
# The sensitivity problem: periodic [0,1]^2, gamma = 5/3, rho = 1 + a density
# perturbation of 16 Fourier modes with |k| <= 4 and random phases (their sum
# scaled to amplitude 1e-2), P = 1, v = 0, B = (0.5, 0.2, 0), face fields
# uniform.
GAMMA = 1.6666666666666667


def density_modes(seed, n1, n2, amp=1e-2, nmodes=16, kmax=4):
    rng = np.random.default_rng(seed)
    ks = [(kx, ky) for kx in range(-kmax, kmax + 1) for ky in range(0, kmax + 1)
          if 0 < kx * kx + ky * ky <= kmax * kmax and (ky > 0 or kx > 0)]
    pick = rng.choice(len(ks), size=nmodes, replace=False)
    phases = rng.uniform(0.0, 2.0 * np.pi, nmodes)
    x = (np.arange(n1) + 0.5) / n1
    y = (np.arange(n2) + 0.5) / n2
    X, Y = np.meshgrid(x, y)
    f = np.zeros((n2, n1))
    for m, idx in enumerate(pick):
        kx, ky = ks[idx]
        f += np.cos(2.0 * np.pi * (kx * X + ky * Y) + phases[m])
    return amp * f / nmodes


def sensitivity_desc(n=512, flux="roe"):
    return dict(nx1=n, nx2=n, x1min=0.0, x1max=1.0, x2min=0.0, x2max=1.0, gamma=GAMMA,
                cour_no=0.4, bc=(4, 4, 4, 4), flux=flux)


def sensitivity_state(seed, n=512, b=(0.5, 0.2, 0.0), pres=1.0):
    """Dense (11, n+8, n+8) float64 initial state (ghosts left 0: set_bvals
    fills them) and the interior density."""
    s = np.zeros((11, n + 8, n + 8))
    rho = 1.0 + density_modes(seed, n, n)
    inner = (slice(4, n + 4), slice(4, n + 4))
    s[0][inner] = rho
    s[4][inner] = pres / (GAMMA - 1.0) + 0.5 * (b[0] ** 2 + b[1] ** 2 + b[2] ** 2)
    for q, v in ((5, b[0]), (6, b[1]), (7, b[2])):
        s[q][inner] = v
    s[8][4:n + 4, 4:n + 5] = b[0]   # B1i incl. the face at ie+1
    s[9][4:n + 5, 4:n + 4] = b[1]   # B2i incl. the face at je+1
    s[10][inner] = b[2]
    return s, rho
