"""GPU parity of the Jacobian at the full sizes the bench and the Gauss-Newton step run,
against the oracle's bsensi (src/bsensi.f90:98-187) on sampled measurements.

The electrode-pair kernel is instantiated per tile count NT = ceil(eanz/16) and, when
the Phi/T staging of all wavenumbers does not fit its LDS budget, accumulates G over
wavenumber chunks of KC < K.  The small parity grids of the forward tests run NT 1-2
in one chunk; here:

* the 64-electrode workload (20k triangles, 2000 measurements): NT = 4, one chunk,
  and the same with the chunk forced to K//2 (two chunks);
* the 400 x 100 quad workload (96 electrodes, 4500 measurements): NT = 6, two chunks
  of 7 by default, and forced to one wavenumber per chunk and to chunks of 5 (a
  ragged last chunk).

Full-size oracle forwards are out of reach on the CPU (1344 band factorisations of
6.6 GFLOP at 400 x 100), so the Jacobian is checked on the device's own potentials:
sampled (electrode, wavenumber) systems against the oracle's forward, then the
voltages from those potentials (oracle rtrafo + bvolti) and J rows of 24 measurements
spread over the survey (oracle bsensi on the same potentials), each at 1e-10
relative (per measurement row against its largest entry).  The 400 x 100 case also
runs a converged update (eps 1e-30) and evaluates the residual of the scaled normal
equations C (J^H W J + lam R) C y = bvec in FP64 on the host from the returned J.
"""
import os

import numpy as np
import pytest

from crtomo_stable_amd import configs, grid as G

TOL = 1e-10
pytestmark = pytest.mark.gpu


def _kc(K, maxn, eanz):
    """The chunk length the launcher picks (crt_launch_sens)."""
    nt = (eanz + 15) // 16
    gbytes = nt * (nt + 1) // 2 * 256 * 16
    budget = max(gbytes, 80 * 1024)
    kc = K
    while kc > 1 and 2 * ((kc * maxn + 3) & ~3) * 16 * nt * 16 > budget:
        kc -= 1
    return kc


def _rows_rel(a, b):
    num = np.abs(a - b).max(axis=0)
    den = np.abs(b).max(axis=0)
    return float((num / den).max())


def _sub_survey(s, idx):
    return G.Survey(strnr=s.strnr[idx].copy(), vnr=s.vnr[idx].copy())


def _check(ctx, O, w, picks_sys, kc_forced):
    g, s = w.grid, w.survey
    ctx.set_grid(g)
    ctx.set_survey(s)
    kwn, kwi, _, _ = ctx.rwaven()
    ctx.precal()
    ctx.set_sigma(w.sigma)
    ctx.forward()
    volt, _ = ctx.voltages(1)
    K, E = kwn.size, g.eanz
    kc = _kc(K, g.maxn, E)
    print(f"[jacobian] {w.name} sanz={g.sanz} mb={g.mb} elanz={g.elanz} nanz={s.nanz} E={E} K={K} "
          f"NT={(E + 15) // 16} KC={kc}", flush=True)
    kpot = ctx.get_kpot()
    okwn, okwi, _, _ = O.rwaven(g, s)
    assert np.array_equal(okwn, kwn) and np.array_equal(okwi, kwi)
    elbg, relbg, kg, lb = O.precal(g, okwn)
    ref = O.forward(g, w.sigma, elbg, relbg, kg, lb, systems=picks_sys)
    for l, k in picks_sys:
        a = kpot[k - 1, l - 1]; b = ref[k - 1, l - 1]
        assert np.abs(a - b).max() / np.abs(b).max() <= TOL, (l, k)
    del ref
    v_ref, _ = O.bvolti(g, s, O.rtrafo(kpot, okwi))
    assert float(np.max(np.abs(volt - v_ref) / np.abs(v_ref))) <= TOL
    idx = np.unique(np.linspace(0, s.nanz - 1, 24).astype(int))
    s_ref = O.bsensi(g, _sub_survey(s, idx), okwi, kpot, elbg, w.sigma, v_ref[idx])
    for f in [None] + list(kc_forced):
        if f is not None:
            assert f < kc
            os.environ["CRT_SENS_KC"] = str(f)
        try:
            ctx.sens(1)
        finally:
            os.environ.pop("CRT_SENS_KC", None)
        sens = ctx.get_sens()
        err = _rows_rel(sens[:, idx], s_ref)
        print(f"  KC={kc if f is None else f}: J rows rel err {err:.2e}", flush=True)
        assert err <= TOL
        del sens
    return kc, K


def test_jacobian_64_electrodes(ctx, oracle_lib):
    w = configs.c1()
    K = 13
    kc, K = _check(ctx, oracle_lib, w, [(1, 1), (32, 7), (64, 13)], [K // 2])
    assert kc == K  # one chunk by default


def test_jacobian_96_electrodes_and_update(ctx, oracle_lib):
    O = oracle_lib
    w = configs.c3()
    kc, K = _check(ctx, O, w, [(1, 1), (96, 14)], [1, 5])
    assert kc < K  # the default runs the cross-chunk accumulation
    # converged normal-equation solve on the same J
    g = w.grid
    volt, sa = ctx.voltages(1)
    ctx.sens(1)
    ctx.bsmatm(0, 1.0, 1.0, nx=g.nx, nz=g.nz)
    rng = np.random.default_rng(4)
    dat = sa + 0.03 * np.abs(sa) * rng.standard_normal(sa.size)
    wmatd = 1.0 / (0.03 * np.abs(sa) + 1e-4) ** 2
    wmatd = wmatd / wmatd.max()
    wdfak = np.ones(sa.size, np.int32)
    par = np.log(w.sigma)
    lam = ctx.blam0(wmatd, wdfak)
    r = ctx.update(lam, 1e-30, 1500, wmatd, wdfak, dat, par)
    sens = ctx.get_sens()
    smatm = O.bsmatmreg(g, g.nx, g.nz, 1.0, 1.0)
    cp = O.CGProblem(sens, wmatd, wdfak, dat, sa, par, smatm, 0, g.nx, lam, 1e-30, 1500)
    cgfac, bvec = O.update_rhs(cp)
    assert np.max(np.abs(r["cgfac"] - cgfac) / cgfac) <= 1e-12
    assert float(np.max(np.abs(r["bvec"] - bvec)) / np.max(np.abs(bvec))) <= 1e-12
    y = r["dpar"] / cgfac
    res = O.cg_matvec(cp, cgfac, y)[1] - bvec
    rel = float(np.linalg.norm(res) / np.linalg.norm(bvec))
    print(f"  converged update: ncg {r['ncg']}, |B y - b| / |b| = {rel:.2e}", flush=True)
    assert rel <= 1e-12  # measured 1.5e-14 after 1139 iterations
