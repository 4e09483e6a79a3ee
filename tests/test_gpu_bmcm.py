"""Model uncertainty on the device (crt_buncert, src/bmcm_mod.f90:49-175) against the
oracle (orc_buncert, itself pinned to the diagonals the reference writes).

Tolerances: ata and ata_reg 1e-13 of max|ata| (the device sums Re(J^H W J) in
another order); cov relative to its largest entry within tol = cond(ata_reg) *
1e-15 (the inverse carries the conditioning), in [1e-12, 1e-8]; res = cov ata and
cov2 = res cov within the error cov's tolerance propagates through the other
factor (tol max|cov| |ata|_1, twice that times |cov|_1), and within 1e-13 of the
products of the device's own factors.  The regularisation fill itself (ata_reg -
ata, the reference's write order) is compared at 1e-13 of max|ata|.
"""
import os

import numpy as np
import pytest

from crtomo_stable_amd import configs

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _close(a, b, tol):
    return float(np.abs(a - b).max()) <= tol * float(np.abs(b).max())


def _check(d, m, gauss, what=("ata", "ata_reg", "cov", "res", "cov2")):
    scale = np.abs(m["ata"]).max()
    assert np.abs(d["ata"] - m["ata"]).max() <= 1e-13 * scale
    assert np.abs(d["ata_reg"] - m["ata_reg"]).max() <= 1e-13 * scale
    fill_d = d["ata_reg"] - d["ata"]; fill_o = m["ata_reg"] - m["ata"]
    assert np.abs(fill_d - fill_o).max() <= 1e-13 * scale
    cond = np.linalg.cond(m["ata_reg"])
    tol = min(1e-8, max(1e-12, cond * 1e-15))
    print(f"  cond(ata_reg) {cond:.2e} -> tol {tol:.1e}", flush=True)
    n1 = lambda a: float(np.abs(a).sum(axis=0).max())  # noqa: E731  (1-norm)
    err_cov = float(np.abs(d["cov"] - m["cov"]).max())
    assert err_cov <= tol * float(np.abs(m["cov"]).max()), err_cov
    # res = cov ata and cov2 = res cov carry cov's error times the other factor's norm
    bounds = {"res": tol * float(np.abs(m["cov"]).max()) * n1(m["ata"]),
              "cov2": 2 * tol * float(np.abs(m["cov"]).max()) * n1(m["ata"]) * n1(m["cov"])}
    for k in ("cov", "res", "cov2"):
        if k in what and m[k] is not None:
            if k != "cov":
                e = float(np.abs(d[k] - m[k]).max())
                print(f"  {k}: max error {e:.2e}, bound {bounds[k]:.2e}, max|{k}| {np.abs(m[k]).max():.2e}",
                      flush=True)
                assert e <= bounds[k], k
            assert np.array_equal(d["diag_" + k], np.diagonal(d[k]))
    # the products themselves, on the device's own factors
    assert np.abs(d["res"] - d["cov"] @ d["ata"]).max() <= 1e-13 * np.abs(d["cov"]).max() * n1(d["ata"])
    if d["cov2"] is not None:
        assert np.abs(d["cov2"] - d["res"] @ d["cov"]).max() <= 1e-13 * np.abs(d["res"]).max() * n1(d["cov"])


def test_buncert_random(ctx, oracle_lib):
    """Random complex J: ltri 0 (Gauss_dble branch; the reference's ltri-0 fill is
    what it is) and a neighbour chain for ltri 1 (Cholesky branch and LU branch)."""
    O = oracle_lib
    rng = np.random.default_rng(11)
    nanz, manz, nx = 45, 36, 6
    J = rng.standard_normal((nanz, manz)) + 1j * rng.standard_normal((nanz, manz))
    w = rng.random(nanz) + 0.5
    wd = rng.integers(1, 3, nanz).astype(np.int32)
    z = np.zeros(nanz, np.complex128)
    sm0 = np.stack([rng.random(manz) + 2.0, -0.01 * rng.random(manz), -0.01 * rng.random(manz)])
    ctx.set_sens(J.T)
    ctx.set_regularization(0, sm0, nx=nx)
    cp0 = O.CGProblem(J.T, w, wd, z, z, np.zeros(manz, np.complex128), sm0, 0, nx, 0.7, 1e-4, 1)
    err, m = O.buncert(cp0, gauss=True)
    ed, d = ctx.buncert(0.7, w, wd, gauss=True)
    print(f"[buncert random ltri 0] oracle {err} device {ed}", flush=True)
    assert err == 0 and ed == 0
    _check(d, m, True)
    # ltri 1: chain of neighbours j-1, j+1
    maxn = 4
    nb = np.zeros((maxn + 1, manz), np.int32)
    sm1 = np.zeros((maxn + 1, manz))
    cpl = -0.3 * (rng.random(manz) + 0.5)
    for j in range(manz):
        ks = [x for x in (j - 1, j + 1) if 0 <= x < manz]
        for k, x in enumerate(ks):
            nb[k, j] = x + 1
            sm1[k, j] = cpl[min(j, x)]
        nb[maxn, j] = len(ks)
        sm1[maxn, j] = 1.0 + 0.6 * len(ks)
    ctx.set_regularization(1, sm1, nx=nx, nachbar=nb)
    cp1 = O.CGProblem(J.T, w, wd, z, z, np.zeros(manz, np.complex128), sm1, 1, nx, 0.7, 1e-4, 1,
                      nachbar=nb, maxn=maxn)
    for gauss in (False, True):
        err, m = O.buncert(cp1, gauss=gauss)
        ed, d = ctx.buncert(0.7, w, wd, gauss=gauss)
        print(f"[buncert random ltri 1 gauss={gauss}] oracle {err} device {ed}", flush=True)
        assert err == 0 and ed == 0
        _check(d, m, gauss)
    # without res / cov2
    ed, d = ctx.buncert(0.7, w, wd, lres=False)
    assert ed == 0 and d["res"] is None and d["cov2"] is None and d["cov"] is not None


@pytest.mark.parametrize("ltri", [0, 1])
def test_buncert_golden_grids(ctx, oracle_lib, ltri):
    """The CG problems of the golden grids (ltri 0: small_quad, Gauss_dble; ltri 1:
    small triangles, chold + linvd), the reference's diagonals as the oracle test
    uses them, and the Cholesky rejection of the indefinite ltri-0 fill (108)."""
    O = oracle_lib
    name = "small_quad" if ltri == 0 else "small_tri"
    w = configs.small_quad(cutmck=O.cutmck) if ltri == 0 else \
        configs.c1(scale=0.2, nelec=12, nmeas=60, seed=7, cutmck=O.cutmck)
    ref = np.load(os.path.join(GOLD, f"ref_{name}.npz"))
    gold = np.load(os.path.join(GOLD, "ref_buncert.npz"))
    p = f"cg{ltri}_"
    g = w.grid
    if ltri == 0:
        smatm, nachbar = O.bsmatmreg(g, g.nx, g.nz, 1.0, 1.0), None
    else:
        nachbar = O.bnachbar(g)
        smatm = O.bsmatmtri(g, nachbar, 1.0, 1.0)
    wmatd = ref[p + "wmatd"]; wdfak = np.ones(ref["sigmaa"].size, np.int32)
    lam = float(ref[p + "lam"])
    cp = O.CGProblem(ref["sens"], wmatd, wdfak, ref[p + "dat"], ref["sigmaa"], ref[p + "par"], smatm, ltri, g.nx,
                     lam, float(ref[p + "eps"]), 2, nachbar=nachbar, maxn=g.maxn)
    gauss = bool(gold[f"q{ltri}_gauss"])
    err, m = O.buncert(cp, gauss=gauss)
    ctx.set_sens(ref["sens"])
    ctx.set_regularization(ltri, smatm, nx=g.nx, nachbar=nachbar)
    ed, d = ctx.buncert(lam, wmatd, wdfak, gauss=gauss)
    print(f"[buncert {name}] manz={g.elanz} gauss={gauss} oracle {err} device {ed}", flush=True)
    assert err == 0 and ed == 0
    _check(d, m, gauss)
    rel = lambda a, b: float(np.abs(a - b).max() / np.abs(b).max())  # noqa: E731
    assert rel(d["diag_ata"], gold[f"q{ltri}_ata"]) <= 1e-13
    assert rel(d["diag_ata_reg"], gold[f"q{ltri}_ata_reg"]) <= 1e-13
    if ltri == 0:
        assert ctx.buncert(lam, wmatd, wdfak, gauss=False, full=False)[0] == 108
