"""Potential fitting: energies, forces and effective fields of frozen bcc Fe
configurations as differentiable functions of the DD05 and Fe exchange/Landau
parameters, on the GPU through libspilady_hip.so's model entry points.

The reference's potential (DD05.cpp:6-164, JijFe.cpp:5-43) is fixed at
compile time and has no gradient; this module has no reference counterpart.
The forward pass is the reference's calculate_force_energy / calculate_spin
physics with the potential's coefficients taken from a parameter vector
theta (22 entries, layout in include/spilady_hip.h), tabulated as
potential_CPU.cpp:61-170 does. The backward pass is a hand-written adjoint
kernel: the vector-Jacobian product of (E, F, H) with respect to theta, exact
for the tabulated model, recomputing pair geometry rather than storing it.

    model = PotentialModel(frozen_configs())
    theta = torch.tensor(model.default_params(), requires_grad=True)
    E, F, H = model(theta)                     # ModelFunction.apply
    loss = mse_loss(E, F, H, targets, natoms)
    loss.backward()                            # theta.grad from the adjoint
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence, Tuple

import numpy as np

from . import abi, configs
from .engine import BatchEngine, Engine, SldError, load

NPARAM = 22
PARAM_NAMES = ([f"a{k}" for k in range(6)] + [f"V{k}" for k in range(8)] + ["aa", "bb", "J_A", "J_rc"]
               + ["Landau_A", "Landau_B", "Landau_C", "Landau_D"])


def default_params() -> np.ndarray:
    """The reference's parameter values (DD05.cpp, JijFe.cpp)."""
    L = load()
    th = np.zeros(NPARAM)
    rc = L.sld_model_default_params(th.ctypes.data_as(C.POINTER(C.c_double)))
    if rc != 0:
        raise SldError(rc, "sld_model_default_params")
    return th


def frozen_configs(n_unit: int = 16, seeds: Sequence[int] = range(200, 232), sigma: float = 0.08,
                   tilt_max_deg: float = 30.0):
    """Frozen bcc Fe configurations: n_unit^3 cubic cells each, Gaussian
    displacements of sigma (A) per component and spin tilts uniform in
    [0, tilt_max_deg] with a uniform azimuth, one per seed."""
    return [configs.make_bcc_fe(n_unit, 300.0, int(s), tilt_max_deg=tilt_max_deg, displace_sigma=sigma)
            for s in seeds]


def perturbed_params(theta0: np.ndarray, rel: float = 0.02, seed: int = 999) -> np.ndarray:
    """theta0 with each entry scaled by (1 + rel * N(0, 1)) (zeros stay zero)."""
    rng = np.random.Generator(np.random.Philox(seed))
    return theta0 * (1.0 + rel * rng.normal(size=theta0.shape))


class PotentialModel:
    """The frozen configurations on one device context (one replica each).
    Atom arrays are compact: configuration k's atoms follow those of
    configurations 0..k-1, in each configuration's own order."""

    def __init__(self, atoms_list, variables: Optional[dict] = None):
        if not atoms_list:
            raise ValueError("no configurations")
        v = dict(configs.fe_variables(300.0, 1, thermostats=False))
        if variables:
            v.update(variables)
        v["eltemp"] = 0
        plist = [abi.params_from_variables(v, a.box, a.r.shape[0], replica=k) for k, a in enumerate(atoms_list)]
        if len(plist) == 1:
            self.engine = Engine(plist[0], atoms_list[0])
        else:
            self.engine = BatchEngine(plist, atoms_list)
        self.L = self.engine.L
        self.h = self.engine.h
        self.nconf = len(atoms_list)
        self.natoms = np.array([a.r.shape[0] for a in atoms_list], dtype=np.int64)
        self.ntot = int(self.L.sld_model_natoms(self.h))
        assert self.ntot == int(self.natoms.sum())
        self._theta: Optional[np.ndarray] = None

    # --- raw entry points (numpy) ---
    def default_params(self) -> np.ndarray:
        return default_params()

    def set_params(self, theta) -> None:
        th = np.ascontiguousarray(np.asarray(theta, dtype=np.float64).reshape(NPARAM))
        if self._theta is not None and np.array_equal(th, self._theta):
            return
        self.engine._rc(self.L.sld_model_set_params(self.h, th.ctypes.data_as(C.POINTER(C.c_double))))
        self._theta = th.copy()

    def eval(self) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
        E = np.zeros(self.nconf)
        F = np.zeros((self.ntot, 3))
        H = np.zeros((self.ntot, 3))
        self.engine._rc(self.L.sld_model_eval(self.h, E.ctypes.data, F.ctypes.data, H.ctypes.data))
        return E, F, H

    def vjp(self, gE=None, gF=None, gH=None) -> np.ndarray:
        def ptr(a, shape):
            if a is None:
                return None, None
            a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(shape))
            return a, a.ctypes.data
        gE, pE = ptr(gE, (self.nconf,))
        gF, pF = ptr(gF, (self.ntot, 3))
        gH, pH = ptr(gH, (self.ntot, 3))
        g = np.zeros(NPARAM)
        self.engine._rc(self.L.sld_model_vjp(self.h, pE, pF, pH, g.ctypes.data))
        return g

    # --- torch ---
    def __call__(self, theta):
        """(E, F, H) as torch tensors on the GPU, differentiable in theta."""
        return ModelFunction.apply(theta, self)

    def close(self):
        self.engine.close()


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("the potential model runs on the GPU only (no device visible)")
    return torch


def _make_function():
    import torch

    class ModelFunction(torch.autograd.Function):
        """theta -> (E[nconf], F[natoms, 3], H[natoms, 3]); backward is the
        adjoint kernel (sld_model_vjp)."""

        @staticmethod
        def forward(ctx, theta, model: PotentialModel):
            t = _torch()
            dev = t.device("cuda", t.cuda.current_device())
            model.set_params(theta.detach().to("cpu", t.float64).numpy())
            E = t.empty(model.nconf, dtype=t.float64, device=dev)
            F = t.empty((model.ntot, 3), dtype=t.float64, device=dev)
            H = t.empty((model.ntot, 3), dtype=t.float64, device=dev)
            t.cuda.synchronize()
            model.engine._rc(model.L.sld_model_eval(model.h, E.data_ptr(), F.data_ptr(), H.data_ptr()))
            ctx.model = model
            ctx.save_for_backward(theta.detach())
            return E, F, H

        @staticmethod
        def backward(ctx, gE, gF, gH):
            t = _torch()
            (theta,) = ctx.saved_tensors
            model: PotentialModel = ctx.model
            model.set_params(theta.to("cpu", t.float64).numpy())
            dev = t.device("cuda", t.cuda.current_device())

            def dptr(g):
                if g is None:
                    return None, None
                g = g.to(device=dev, dtype=t.float64).contiguous()
                return g, g.data_ptr()
            gE, pE = dptr(gE)
            gF, pF = dptr(gF)
            gH, pH = dptr(gH)
            grad = t.empty(NPARAM, dtype=t.float64, device=dev)
            t.cuda.synchronize()
            model.engine._rc(model.L.sld_model_vjp(model.h, pE, pF, pH, grad.data_ptr()))
            return grad.to(device=theta.device, dtype=theta.dtype).view_as(theta), None

    return ModelFunction


class _Lazy:
    """ModelFunction is created on first use so that importing this module
    does not import torch."""

    def __init__(self):
        self._f = None

    def apply(self, *a):
        if self._f is None:
            self._f = _make_function()
        return self._f.apply(*a)


ModelFunction = _Lazy()


def mse_loss(E, F, H, targets, natoms, w_e: float = 1.0, w_f: float = 1.0, w_h: float = 1.0):
    """w_e mean((E - E*) / N)^2 + w_f mean |F - F*|^2 + w_h mean |H - H*|^2
    (energies per atom of each configuration; torch or numpy)."""
    Et, Ft, Ht = targets
    de = (E - Et) / natoms
    return w_e * (de * de).mean() + w_f * ((F - Ft) ** 2).mean() + w_h * ((H - Ht) ** 2).mean()


def mse_upstream(E, F, H, targets, natoms, w_e: float = 1.0, w_f: float = 1.0, w_h: float = 1.0):
    """Gradients of mse_loss with respect to E, F, H (numpy)."""
    Et, Ft, Ht = targets
    n = np.asarray(natoms, dtype=np.float64)
    gE = w_e * 2.0 * (E - Et) / (n * n) / E.size
    gF = w_f * 2.0 * (F - Ft) / F.size
    gH = w_h * 2.0 * (H - Ht) / H.size
    return gE, gF, gH
