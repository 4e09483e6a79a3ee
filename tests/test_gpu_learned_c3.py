"""The learned preconditioner at the 400 x 100 size (40000 cells, 96 electrodes, 4500
measurements).  The network is trained on the CG systems of Gauss-Newton steps on two
other models (seeded Gaussian random fields on the same grid and survey, none of them
the scored model), then scored on the first Gauss-Newton system of the workload's own
data:

* converged (eps 1e-30): the learned preconditioner's CG meets the stop test by itself
  (no fallback to Jacobi), and its update, like the Jacobi one (the reference's method),
  solves the scaled normal equations C (J^H W J + lam R) C y = bvec to rounding level,
  the residual evaluated in FP64 on the host from the device's J (oracle matvec) --
  the preconditioner changes the iteration count only, the two updates agree within
  1e-10;
* at the reference's stop test with eps 1e-12 (relative residual 1e-6) it takes fewer
  iterations than Jacobi, and the update agrees with the Jacobi one within the stop's
  accuracy (1e-4).
"""
import numpy as np
import pytest

from crtomo_stable_amd import configs, inversion as INV

NCG_CONV = 4000

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def test_learned_pc_c3(ctx, oracle_lib):
    w = configs.c3()
    g = w.grid
    ctx.set_grid(g)
    ctx.set_survey(w.survey)
    ctx.rwaven()
    ctx.precal()
    par0 = np.full(g.elanz, np.log((1.0 / 100.0) * np.exp(-1j * 5e-3)))
    ctx.set_cg_precision(32)
    ctx.set_cg_preconditioner(2)
    ctx.pc_reset()  # a new network trained on this test's inversions alone
    losses = []
    try:
        for m in range(2):
            sig_m = configs.grf_sigma(g, 4.0 + 6.0 * m, seed=1000 + m)
            dat_m, w_m = INV.synthetic_data(ctx, sig_m, seed=200 + m)
            INV.gauss_newton(ctx, g, dat_m, w_m, par0, 2, ltri=0, eps=1e-4, lam_factor=0.5,
                             on_update=lambda c, it: losses.append(c.pc_train(100, 3e-3)))
        print(f"[learned c3] trained on {len(losses)} systems, loss {losses[0][0]:.3e} -> {losses[-1][1]:.3e}",
              flush=True)
        ctx.set_cg_precision(64)
        ctx.set_cg_preconditioner(0)
        dat, wmatd = INV.synthetic_data(ctx, w.sigma)
        wdfak = np.ones(dat.size, np.int32)
        ctx.set_sigma(np.exp(par0))
        ctx.forward()
        ctx.voltages(1)
        ctx.sens(1)
        ctx.bsmatm(0, 1.0, 1.0, nx=g.nx, nz=g.nz)
        lam = ctx.blam0(wmatd, wdfak)
        jac_conv = ctx.update(lam, 1e-30, NCG_CONV, wmatd, wdfak, dat, par0)
        jac = ctx.update(lam, 1e-12, g.elanz // 10, wmatd, wdfak, dat, par0)
        ctx.set_cg_preconditioner(3)
        lrn_conv = ctx.update(lam, 1e-30, NCG_CONV, wmatd, wdfak, dat, par0)
        lrn = ctx.update(lam, 1e-12, g.elanz // 10, wmatd, wdfak, dat, par0)
        # the same stop test with the network's analytic prior alone (no network) and with
        # the exact FSAI rows of the same stencil pattern (the network's teacher)
        others = {}
        for name, kind in (("prior", 4), ("fsai_stencil", 1)):
            ctx.set_cg_preconditioner(kind)
            others[name] = ctx.update(lam, 1e-12, g.elanz // 10, wmatd, wdfak, dat, par0)
    finally:
        ctx.set_cg_precision(64)
        ctx.set_cg_preconditioner(0)
    e_conv = _rel(lrn_conv["dpar"], jac_conv["dpar"])
    e_stop = _rel(lrn["dpar"], jac["dpar"])
    print(f"[learned c3] converged: ncg {lrn_conv['ncg']} vs Jacobi {jac_conv['ncg']}, dpar diff {e_conv:.2e}; "
          f"eps 1e-12: ncg {lrn['ncg']} vs Jacobi {jac['ncg']}, dpar diff {e_stop:.2e}", flush=True)
    assert lrn_conv["ncg"] < NCG_CONV and jac_conv["ncg"] < NCG_CONV
    O = oracle_lib
    sa = ctx.voltages(1)[1]
    cp = O.CGProblem(ctx.get_sens(), wmatd, wdfak, dat, sa, par0, O.bsmatmreg(g, g.nx, g.nz, 1.0, 1.0), 0, g.nx,
                     lam, 1e-30, NCG_CONV)
    cgfac, bvec = O.update_rhs(cp)
    for name, r in (("jacobi", jac_conv), ("learned", lrn_conv)):
        res = O.cg_matvec(cp, cgfac, r["dpar"] / cgfac)[1] - bvec
        rel = float(np.linalg.norm(res) / np.linalg.norm(bvec))
        print(f"[learned c3] {name}: |B y - b| / |b| = {rel:.2e}", flush=True)
        assert rel <= 1e-12
    for name, r in others.items():
        e = _rel(r["dpar"], jac["dpar"])
        print(f"[learned c3] eps 1e-12, {name}: ncg {r['ncg']}, dpar diff vs Jacobi {e:.2e}", flush=True)
        assert e <= 1e-4
    assert e_conv <= 1e-10  # measured 5.6e-15
    assert e_stop <= 1e-4
    assert lrn["ncg"] < jac["ncg"]
