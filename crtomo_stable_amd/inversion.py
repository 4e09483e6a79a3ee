"""Complex Gauss-Newton inversion on the device: the reference's inversion loop in its
fixed-lambda mode, every heavy step a library call.

Per iteration (src/inv.f90:332-956 with a lambda schedule instead of the lambda search:
one lambda per step, ``lamfix`` (blam0) first and a constant factor after, as the
reference's fixed-lambda mode ``llamf`` does, src/inv.f90:733-736):

1. bsensi on the device for the current model, blam0 at the first step;
2. update (RHS, cgfac, cjg, ``par += dpar``, src/update.f90) and ``sigma = exp(par)``
   (src/bsigma.f90);
3. forward on the new model (kompab..vre for every (electrode, wavenumber) system,
   rtrafo, bvolti) and its misfit ``nrmsd = sqrt(sum_i wmatd_i |d_i - f_i|^2 / nanz)``
   (src/dmisft.f90:85-121, no polarity changes, all data used); a step that does not
   lower the misfit is retried at half length once (step control);
4. stop at the target misfit (``rmsmin``) or when no step lowers the misfit.

The CG form (Jacobi scaling as the reference, or the library's FSAI / learned
preconditioners, FP64 or FP32 Jacobian copy) is whatever the context is set to; the
loop records per iteration the misfit, lambda, CG iterations and the device time of
each phase, so that forms can be compared at equal final misfit.
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np


def _sync():
    ctypes.CDLL("libamdhip64.so").hipDeviceSynchronize()


def nrmsd(dat, sigmaa, wmatd) -> float:
    """src/dmisft.f90:85-121 (all data valid, no difference inversion, not FPI)."""
    psi2 = wmatd * np.abs(dat - sigmaa) ** 2
    return float(np.sqrt(psi2.sum() / dat.size))


@dataclass
class Iteration:
    it: int
    rms: float
    lam: float = 0.0
    ncg: int = 0
    ms: dict = field(default_factory=dict)


@dataclass
class Result:
    iterations: List[Iteration]
    final_rms: float
    par: np.ndarray
    lamfix: float

    def step_ms(self) -> List[float]:
        """Wall time of every Gauss-Newton step (Jacobian, update, the forward of the
        new model and its misfit, the half-step forward where one was needed),
        milliseconds."""
        return [sum(v for k, v in i.ms.items() if k != "step") for i in self.iterations if "update" in i.ms]


def gauss_newton(ctx, grid, dat, wmatd, par0, nit: int, lamfix: Optional[float] = None, ltri: int = 0,
                 alfx: float = 1.0, alfz: float = 1.0, eps: float = 1e-12, ncgmax: Optional[int] = None,
                 rmsmin: float = 1.0, lam_factor: float = 1.0, on_update=None) -> Result:
    """Run at most ``nit`` Gauss-Newton steps from the model ``par0`` (log conductivity)
    on the context's grid and survey (rwaven and precal done).

    lambda: ``lamfix`` (blam0 of the first Jacobian when None) at the first step, then
    multiplied by ``lam_factor`` per step (1: the reference's fixed-lambda mode,
    llamf = 1; 0.5: its first lambda trial factor fstart).  A step whose misfit is not
    below the previous one is retried once at half length (the reference's step
    control, src/inv.f90:780-790, ``step = 0.5``); if that fails too the loop ends at
    the previous model.  ``on_update(ctx, it)`` is called after every update (the
    learned preconditioner's training hook)."""
    wdfak = np.ones(dat.size, np.int32)
    manz = grid.elanz
    ncgmax = manz // 10 if ncgmax is None else ncgmax  # src/rall.f90:735-738
    par = np.array(par0, np.complex128)
    ctx.bsmatm(ltri, alfx, alfz, nx=getattr(grid, "nx", 0), nz=getattr(grid, "nz", 0))

    def forward(p):
        ctx.set_sigma(np.exp(p))
        ctx.forward()
        _, sa = ctx.voltages(1)
        return nrmsd(dat, sa, wmatd)

    its: List[Iteration] = []
    _sync()
    t0 = time.perf_counter()
    rms = forward(par)
    _sync()
    its.append(Iteration(it=0, rms=rms, ms={"forward": 1e3 * (time.perf_counter() - t0)}))
    lam = None
    for it in range(1, nit + 1):
        if rms <= rmsmin:
            break
        ms = {}
        t0 = time.perf_counter()
        ctx.sens(1)
        _sync()
        ms["sens"] = 1e3 * (time.perf_counter() - t0)
        t0 = time.perf_counter()
        if lamfix is None:
            lamfix = ctx.blam0(wmatd, wdfak, alfx, alfz)
        lam = lamfix if lam is None else lam * lam_factor
        r = ctx.update(lam, eps, ncgmax, wmatd, wdfak, dat, par)
        _sync()
        ms["update"] = 1e3 * (time.perf_counter() - t0)
        if on_update is not None:
            on_update(ctx, it)
        t0 = time.perf_counter()
        rms_new = forward(r["par"])
        _sync()
        ms["forward"] = 1e3 * (time.perf_counter() - t0)
        step = 1.0
        if rms_new >= rms:
            # step control: half the model update, one more forward
            t0 = time.perf_counter()
            p_half = par + 0.5 * (r["par"] - par)
            rms_half = forward(p_half)
            _sync()
            ms["forward_half_step"] = 1e3 * (time.perf_counter() - t0)
            step = 0.5
            if rms_half >= rms:
                its.append(Iteration(it=it, rms=rms, lam=lam, ncg=r["ncg"], ms=ms))
                ctx.set_sigma(np.exp(par))
                break
            r["par"], rms_new = p_half, rms_half
        par, rms = r["par"], rms_new
        its.append(Iteration(it=it, rms=rms, lam=lam, ncg=r["ncg"], ms=ms))
        its[-1].ms["step"] = step
    return Result(iterations=its, final_rms=rms, par=par, lamfix=lamfix)


def synthetic_data(ctx, sigma_true, noise_rel: float = 0.03, noise_abs: float = 1e-4, seed: int = 11):
    """Voltages of ``sigma_true`` with ``noise_rel`` relative + ``noise_abs`` V Gaussian
    noise on the magnitude (the C3 survey's 3 % + 1e-4 V), as data -ln V and weights
    1 / (relative error)^2 (src/rdati.f90:254: stabw0, stabm0)."""
    ctx.set_sigma(sigma_true)
    ctx.forward()
    v, _ = ctx.voltages(1)
    rng = np.random.default_rng(seed)
    err = noise_rel * np.abs(v) + noise_abs
    vo = v + err * rng.standard_normal(v.size) * v / np.abs(v)
    return -np.log(vo), 1.0 / (err / np.abs(v)) ** 2
