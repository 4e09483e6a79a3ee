"""Error returns of the forward, voltage and normal-equation entry points, through the
per-context path and the batched graph path, against the codes the reference's
routines set (src/error.txt):

* 27 zero diagonal: scalab stops on |a_ii| = 0 (src/scalab.f90:54-58); a model of
  zero conductivity makes every diagonal entry vanish, and the oracle reports 27;
* 19 band width exceeded: the grid's element edges span more than mb
  (src/relem.f90 / cutmck band check), refused when the grid is set;
* 82 zero voltage: bvolti takes -log(V) and stops on V = 0 (src/bvolti.f90:80-88),
  here a measurement whose two potential electrodes coincide;
* 901 (library: call order / state): a Jacobian and a regularisation installed for
  different cell counts are refused by every normal-equation entry point before any
  kernel runs.
The inputs are the ones the host or the kernels turn into these codes; every kernel
still runs to its end on them.
"""
import ctypes as C

import numpy as np
import pytest

from crtomo_stable_amd import configs, grid as G, lib
from crtomo_stable_amd.lib import CRTError

pytestmark = pytest.mark.gpu


def _setup(ctx, w, sigma=None):
    ctx.set_grid(w.grid)
    ctx.set_survey(w.survey)
    ctx.rwaven()
    ctx.precal()
    ctx.set_sigma(w.sigma if sigma is None else sigma)


def test_zero_diagonal(ctx, oracle_lib):
    O = oracle_lib
    w = configs.small_quad()
    zero = np.zeros(w.grid.elanz, np.complex128)
    kwn, _, _, _ = O.rwaven(w.grid, w.survey)
    elbg, relbg, kg, lb = O.precal(w.grid, kwn)
    with pytest.raises(RuntimeError, match="errnr 27"):
        O.forward(w.grid, zero, elbg, relbg, kg, lb)
    _setup(ctx, w, zero)
    with pytest.raises(CRTError) as ei:
        ctx.forward()
    print(f"[errors] zero diagonal: {ei.value}", flush=True)
    assert ei.value.code == 27
    # the context is usable again with a proper model
    ctx.set_sigma(w.sigma)
    ctx.forward()


def test_zero_diagonal_batched():
    """The same error through the batched graph launch, with a healthy model beside it:
    the batch returns 27 and only the failing model loses its potentials."""
    good, bad = configs.small_quad(), configs.small_quad(nx=30, nz=10, nelec=8)
    cs = []
    try:
        for w, sig in ((good, good.sigma), (bad, np.zeros(bad.grid.elanz, np.complex128))):
            c = lib.Context(0)
            cs.append(c)
            _setup(c, w, sig)
        with lib.Batch(cs) as b:
            with pytest.raises(CRTError) as ei:
                b.forward()
            assert ei.value.code == 27
            cs[0].voltages(1)  # the healthy model's potentials are there
            with pytest.raises(CRTError) as e2:
                cs[1].voltages(1)
            assert e2.value.code == 901
            # with a proper model in its place the same graph replays cleanly
            cs[1].set_sigma(bad.sigma)
            b.forward()
            cs[1].voltages(1)
    finally:
        for c in cs:
            c.close()


def test_band_exceeded(ctx):
    w = configs.small_quad()
    g = w.grid
    rnr = g.rnr if g.rnr.size else np.zeros(1, np.int32)
    p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    rc = ctx.lib.crt_set_grid(ctx._h, g.sanz, g.mb - 1, g.typanz, p(g.typ), p(g.nelanz), p(g.selanz), g.maxn,
                              p(g.nrel), p(g.x), p(g.y), p(rnr), g.eanz, p(g.enr))
    assert rc == 19
    ctx.set_grid(g)  # the true band width is accepted


def test_zero_voltage(ctx):
    w = configs.small_quad()
    q = w.survey.electrodes()[:6].copy()
    q[3, 3] = q[3, 2]  # measurement 4: both potential electrodes at one node
    s = G.Survey(strnr=q[:, 1] * 10000 + q[:, 0], vnr=q[:, 3] * 10000 + q[:, 2])
    _setup(ctx, w)
    ctx.set_survey(s)
    ctx.forward()
    with pytest.raises(CRTError) as ei:
        ctx.voltages(1)
    assert ei.value.code == 82
    volt, _ = ctx.voltages(0)  # bvolt (CRMod) has no logarithm and no stop
    assert volt[3] == 0.0


def test_mismatched_jacobian_and_regularization(ctx):
    """J for one cell count, R for another: crt_update, crt_cjg and crt_buncert refuse
    (901) instead of reading past either set of buffers; crt_blam0 needs J only."""
    w = configs.small_quad()
    g = w.grid
    ctx.set_grid(g)
    manz, nanz = g.elanz, 20
    rng = np.random.default_rng(0)
    sens = rng.standard_normal((manz, nanz)) + 1j * rng.standard_normal((manz, nanz))
    ctx.set_sens(sens)
    # a regularisation for a grid of half the cells
    half = manz // 2
    nx = g.nx
    smatm = np.ones((3, half))
    ctx.set_regularization(0, smatm, nx=nx)
    wmatd = np.ones(nanz)
    wdfak = np.ones(nanz, np.int32)
    dat = rng.standard_normal(nanz) + 1j * rng.standard_normal(nanz)
    sa = np.zeros(nanz, np.complex128)
    par = np.zeros(manz, np.complex128)
    with pytest.raises(CRTError) as e1:
        ctx.update(1.0, 1e-4, 10, wmatd, wdfak, dat, par, sigmaa=sa)
    assert e1.value.code == 901
    with pytest.raises(CRTError) as e2:
        ctx.cjg(1.0, 1e-4, 10, wmatd, wdfak, np.ones(manz), np.ones(manz, np.complex128))
    assert e2.value.code == 901
    with pytest.raises(CRTError) as e3:
        ctx.buncert(1.0, wmatd, wdfak)
    assert e3.value.code == 901
    assert ctx.blam0(wmatd, wdfak) > 0.0
    # the matching regularisation makes the same call work
    ctx.set_regularization(0, np.ones((3, manz)), nx=nx)
    r = ctx.update(1.0, 1e-4, 10, wmatd, wdfak, dat, par, sigmaa=sa)
    assert np.all(np.isfinite(r["dpar"]))
