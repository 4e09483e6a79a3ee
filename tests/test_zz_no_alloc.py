"""crt_cjg_noalloc, the CG entry point that allocates nothing (no counterpart in the
reference; for callers that cannot allocate device memory): its buffers live in the
library's static device scratch, its kernels and their order are crt_cjg's.

Two solves in a row, the second on a smaller system (so the scratch holds the first
call's larger contents when the second starts): each against crt_cjg on the same
resident J and regularisation (dpar, cgres and ncg identical: the same kernels in the
same order) and, converged (eps 1e-30), against the oracle's cjggra at 1e-10.  A solve
whose buffers exceed the scratch is refused with 904 before any launch.
"""
import numpy as np
import pytest

from crtomo_stable_amd import configs
from crtomo_stable_amd.lib import CRTError

pytestmark = pytest.mark.gpu


def _system(O, w, ltri, seed):
    g, s = w.grid, w.survey
    kwn, kwi, _, _ = O.rwaven(g, s)
    elbg, relbg, kg, lb = O.precal(g, kwn)
    kpot = O.forward(g, w.sigma, elbg, relbg, kg, lb)
    volt, sa = O.bvolti(g, s, O.rtrafo(kpot, kwi))
    sens = O.bsensi(g, s, kwi, kpot, elbg, w.sigma, volt)
    rng = np.random.default_rng(seed)
    dat = sa + 0.02 * (rng.standard_normal(sa.size) + 1j * 0.1 * rng.standard_normal(sa.size))
    wmatd = 1.0 / (0.03 + 0.01 * rng.random(sa.size)) ** 2
    wdfak = np.ones(sa.size, np.int32)
    par = np.log(w.sigma) + 0.05 * rng.standard_normal(w.sigma.size)
    if ltri == 0:
        smatm, nachbar = O.bsmatmreg(g, g.nx, g.nz, 1.0, 1.0), None
    else:
        nachbar = O.bnachbar(g)
        smatm = O.bsmatmtri(g, nachbar, 1.0, 1.0)
    return dict(sens=sens, sa=sa, dat=dat, wmatd=wmatd, wdfak=wdfak, par=par, smatm=smatm, nachbar=nachbar)


def test_cjg_noalloc_matches_cjg(ctx, oracle_lib):
    O = oracle_lib
    cases = [
        (configs.c1(scale=0.2, nelec=12, nmeas=60, seed=7), 1, 2.0),  # larger system first
        (configs.small_quad(), 0, 5.0),                                 # then a smaller one
    ]
    for w, ltri, lam in cases:
        P = _system(O, w, ltri, seed=11)
        for eps, ncgmax in ((1e-30, 2000), (1e-4, 50)):
            cp = O.CGProblem(P["sens"], P["wmatd"], P["wdfak"], P["dat"], P["sa"], P["par"], P["smatm"], ltri,
                             w.grid.nx, lam, eps, ncgmax, nachbar=P["nachbar"], maxn=w.grid.maxn)
            cgfac, bvec = O.update_rhs(cp)
            ctx.set_grid(w.grid)
            ctx.set_sens(P["sens"])
            ctx.set_regularization(ltri, P["smatm"], nx=w.grid.nx, nachbar=P["nachbar"])
            print(f"[noalloc] manz={w.grid.elanz} ltri={ltri} eps={eps} ncgmax={ncgmax}", flush=True)
            d1, r1, n1 = ctx.cjg(lam, eps, ncgmax, P["wmatd"], P["wdfak"], cgfac, bvec, noalloc=True)
            d0, r0, n0 = ctx.cjg(lam, eps, ncgmax, P["wmatd"], P["wdfak"], cgfac, bvec)
            assert n1 == n0
            assert np.array_equal(d1, d0)
            assert np.array_equal(r1, r0)
            if eps == 1e-30:
                dpar_o, _, _ = O.cjggra(cp, cgfac, bvec)
                err = float(np.max(np.abs(d1 - dpar_o)) / np.max(np.abs(dpar_o)))
                print(f"  ncg {n1}, dpar vs oracle {err:.2e}", flush=True)
                assert err <= 1e-10


def test_cjg_noalloc_rejects_oversize(ctx, oracle_lib):
    w = configs.small_quad()
    P = _system(oracle_lib, w, 0, seed=3)
    ctx.set_grid(w.grid)
    ctx.set_sens(P["sens"])
    ctx.set_regularization(0, P["smatm"], nx=w.grid.nx)
    manz = w.grid.elanz
    # a residual history of 2^24 iterations alone needs 128 MiB
    with pytest.raises(CRTError) as ei:
        ctx.cjg(1.0, 1e-4, 1 << 24, P["wmatd"], P["wdfak"], np.ones(manz), np.ones(manz, np.complex128),
                noalloc=True)
    assert ei.value.code == 904
