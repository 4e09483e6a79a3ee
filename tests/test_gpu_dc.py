"""GPU parity of the DC inversion (ldc) and the final phase improvement (lfpi)
against the oracle's DC twins, which match the reference build bit for bit
(tests/test_oracle_cpu.py):

* forward: kompadc/scaldc/choldc/vredc (src/inv.f90:355-398) -> kpotdc, rtrafo's DC
  branch -> hpotdc, bvolti's DC branch (volt, sigmaa, including -pi for negative V),
  bsendc -> sensdc: 1e-10 relative;
* update + cjggdc (src/update.f90, src/cg_mod.f90:106-196) on sensdc (DC) and on
  DBLE(sens) from DIMAG(bvec) (FPI): cgfac and bvec 1e-12, dpar 1e-10 after 3
  iterations and for the converged solve (eps 1e-30), where CG rounding does not
  decide the result (the complex CG tests explain the bound);
* blam0's ldc branch: 1e-12; buncert's DC branch (bata on sensdc) as the complex
  buncert tests hold it.
"""
import numpy as np
import pytest

from crtomo_stable_amd import configs, grid as G
from crtomo_stable_amd.lib import CRTError

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def _workload(ltri):
    return configs.small_quad() if ltri == 0 else configs.c1(scale=0.2, nelec=12, nmeas=60, seed=7)


def _flip_half(s):
    """Every second measurement with swapped potential electrodes: negative voltages."""
    vnr = s.vnr.copy()
    e3, e4 = vnr % 10000, vnr // 10000
    sw = (np.arange(vnr.size) % 2 == 0) & (e3 > 0)
    vnr[sw] = e3[sw] * 10000 + e4[sw]
    return G.Survey(strnr=s.strnr.copy(), vnr=vnr)


def _oracle_dc(O, w, survey):
    g = w.grid
    kwn, kwi, _, _ = O.rwaven(g, survey)
    elbg, relbg, kg, lb = O.precal(g, kwn)
    kp = O.forward_dc(g, w.sigma, elbg, relbg, kg, lb)
    hp = O.rtrafo_dc(kp, kwi)
    volt, sa = O.bvolti_dc(g, survey, hp)
    sd = O.bsendc(g, survey, kwi, kp, elbg, w.sigma, volt)
    return dict(kpot=kp, hpot=hp, volt=volt, sa=sa, sens=sd)


@pytest.mark.parametrize("ltri", [0, 1])
def test_dc_forward_sens(ctx, oracle_lib, ltri):
    O = oracle_lib
    w = _workload(ltri)
    s = _flip_half(w.survey)
    print(f"[dc forward] ltri={ltri} sanz={w.grid.sanz} nanz={s.nanz}", flush=True)
    o = _oracle_dc(O, w, s)
    assert (o["volt"].real < 0).any()
    try:
        ctx.set_inversion_kind(1)
        ctx.set_grid(w.grid)
        ctx.set_survey(s)
        ctx.rwaven()
        ctx.precal()
        ctx.set_sigma(w.sigma)
        ctx.forward()
        kp = ctx.get_kpot()
        assert kp.dtype == np.float64
        assert _rel(kp, o["kpot"]) <= 1e-10
        assert _rel(ctx.get_hpot(), o["hpot"]) <= 1e-10
        volt, sa = ctx.voltages(1)
        assert np.all(volt.imag == 0.0)
        assert float(np.max(np.abs(volt - o["volt"]) / np.abs(o["volt"]))) <= 1e-10
        assert _rel(sa, o["sa"]) <= 1e-10
        assert np.array_equal(np.sign(sa.imag), np.sign(o["sa"].imag))
        ctx.sens(1)
        sd = ctx.get_sens()
        assert sd.dtype == np.float64
        assert _rel(sd, o["sens"]) <= 1e-10
        # the DC twins exist only for the inversion's routines
        with pytest.raises(CRTError):
            ctx.voltages(0)
    finally:
        ctx.set_inversion_kind(0)


def _cg_problem(O, w, ltri, kind, seed=0):
    g, s = w.grid, w.survey
    kwn, kwi, _, _ = O.rwaven(g, s)
    elbg, relbg, kg, lb = O.precal(g, kwn)
    if kind == 1:
        o = _oracle_dc(O, w, s)
        sa, sensdc, sens = o["sa"], o["sens"], np.zeros(o["sens"].shape, np.complex128)
    else:
        kp = O.forward(g, w.sigma, elbg, relbg, kg, lb)
        volt, sa = O.bvolti(g, s, O.rtrafo(kp, kwi))
        sens, sensdc = O.bsensi(g, s, kwi, kp, elbg, w.sigma, volt), None
    rng = np.random.default_rng(seed)
    dat = sa + 0.02 * (rng.standard_normal(sa.size) + 1j * 0.1 * rng.standard_normal(sa.size))
    wmatd = 1.0 / (0.03 + 0.01 * rng.random(sa.size)) ** 2
    wdfak = np.ones(sa.size, np.int32)
    par = np.log(w.sigma) + 0.05 * (rng.standard_normal(w.sigma.size) + 1j * rng.standard_normal(w.sigma.size))
    if ltri == 0:
        smatm, nachbar = O.bsmatmreg(g, g.nx, g.nz, 1.0, 1.0), None
    else:
        nachbar = O.bnachbar(g)
        smatm = O.bsmatmtri(g, nachbar, 1.0, 1.0)
    return dict(sens=sens, sensdc=sensdc, sa=sa, dat=dat, wmatd=wmatd, wdfak=wdfak, par=par, smatm=smatm,
                nachbar=nachbar)


@pytest.mark.parametrize("kind", [1, 2])
@pytest.mark.parametrize("ltri", [0, 1])
@pytest.mark.parametrize("ncgmax,eps", [(3, 1e-30), (3000, 1e-30)])
def test_dc_fpi_update(ctx, oracle_lib, kind, ltri, ncgmax, eps):
    O = oracle_lib
    w = _workload(ltri)
    lam = 3.0
    P = _cg_problem(O, w, ltri, kind)
    cp = O.CGProblem(P["sens"], P["wmatd"], P["wdfak"], P["dat"], P["sa"], P["par"], P["smatm"], ltri, w.grid.nx,
                     lam, eps, ncgmax, nachbar=P["nachbar"], maxn=w.grid.maxn)
    cgfac_o, bvec_o, dpar_o, cgres_o, ncg_o = O.update_dc(cp, kind, P["sensdc"])
    print(f"[dc/fpi update] kind={kind} ltri={ltri} ncgmax={ncgmax} oracle ncg={ncg_o}", flush=True)
    try:
        ctx.set_inversion_kind(kind)
        ctx.set_grid(w.grid)
        ctx.set_sens(P["sensdc"] if kind == 1 else P["sens"])
        ctx.set_regularization(ltri, P["smatm"], nx=w.grid.nx, nachbar=P["nachbar"])
        r = ctx.update(lam, eps, ncgmax, P["wmatd"], P["wdfak"], P["dat"], P["par"], sigmaa=P["sa"])
        assert np.max(np.abs(r["cgfac"] - cgfac_o) / cgfac_o) <= 1e-12
        assert _rel(r["bvec"], bvec_o) <= 1e-12
        assert _rel(r["dpar"], dpar_o) <= 1e-10
        # DC moves the real part only, FPI the imaginary part only
        assert np.all((r["dpar"].imag if kind == 1 else r["dpar"].real) == 0.0)
        assert _rel(r["par"] - P["par"], dpar_o) <= 1e-10
        assert abs(r["ncg"] - ncg_o) <= max(3, 0.03 * ncg_o)
        assert r["cgres"][0] == float(r["ncg"])
        # crt_cjg on the same scaled system returns the raw real CG solution
        dpar_c, _, ncg_c = ctx.cjg(lam, eps, ncgmax, P["wmatd"], P["wdfak"], cgfac_o, bvec_o)
        raw = dpar_o / cgfac_o
        assert _rel(dpar_c, raw.imag if kind == 2 else raw.real) <= 1e-10
        assert np.all(dpar_c.imag == 0.0)
        if kind == 1:
            assert abs(ctx.blam0(P["wmatd"], P["wdfak"]) -
                       O.blam0(O.CGProblem(P["sensdc"].astype(np.complex128), P["wmatd"], P["wdfak"], P["dat"],
                                           P["sa"], P["par"], P["smatm"], ltri, w.grid.nx, lam, eps, ncgmax,
                                           nachbar=P["nachbar"], maxn=w.grid.maxn), True, 1.0, 1.0)) <= \
                1e-12 * ctx.blam0(P["wmatd"], P["wdfak"])
            # buncert of the DC inversion: bata on the real sensdc (src/bmcm_mod.f90:197-206),
            # the rest as the complex branches; the oracle's bata on sensdc with zero
            # imaginary parts forms the same products in the same order
            from test_gpu_bmcm import _check
            cpd = O.CGProblem(P["sensdc"].astype(np.complex128), P["wmatd"], P["wdfak"], P["dat"], P["sa"],
                              P["par"], P["smatm"], ltri, w.grid.nx, lam, eps, ncgmax, nachbar=P["nachbar"],
                              maxn=w.grid.maxn)
            err_o, m = O.buncert(cpd, gauss=True)
            err_d, d = ctx.buncert(lam, P["wmatd"], P["wdfak"], gauss=True)
            assert err_o == 0 and err_d == 0
            _check(d, m, True)
    finally:
        ctx.set_inversion_kind(0)


@pytest.mark.parametrize("kind", [1, 2])
@pytest.mark.parametrize("ltri,prec", [(0, 2), (1, 1)])
def test_dc_fpi_fast_cg_forms(ctx, oracle_lib, kind, ltri, prec):
    """The CG forms without a counterpart in the reference on the real J of the DC and
    final-phase-improvement branches: FSAI preconditioning built on the real J (pattern
    2 on the regular grid, 1 on the triangle mesh) alone and with the FP32 copy of the
    real J.  The converged update (eps = 1e-30) equals the oracle's cjggdc to 1e-10 with
    FP64 J and to 1e-6 with the FP32 copy (the FP32 J products, as for the complex
    inversion); FSAI needs fewer iterations than the Jacobi scaling at eps 1e-12; the
    learned preconditioner is refused (903)."""
    O = oracle_lib
    w = _workload(ltri)
    lam = 3.0
    P = _cg_problem(O, w, ltri, kind)
    cp = O.CGProblem(P["sens"], P["wmatd"], P["wdfak"], P["dat"], P["sa"], P["par"], P["smatm"], ltri, w.grid.nx,
                     lam, 1e-30, 3000, nachbar=P["nachbar"], maxn=w.grid.maxn)
    _, _, dpar_o, _, ncg_o = O.update_dc(cp, kind, P["sensdc"])
    try:
        ctx.set_inversion_kind(kind)
        ctx.set_grid(w.grid)
        ctx.set_sens(P["sensdc"] if kind == 1 else P["sens"])
        ctx.set_regularization(ltri, P["smatm"], nx=w.grid.nx, nachbar=P["nachbar"])
        args = (P["wmatd"], P["wdfak"], P["dat"], P["par"])
        jac = ctx.update(lam, 1e-12, 400, *args, sigmaa=P["sa"])
        ctx.set_cg_preconditioner(prec)
        conv = ctx.update(lam, 1e-30, 3000, *args, sigmaa=P["sa"])
        stop = ctx.update(lam, 1e-12, 400, *args, sigmaa=P["sa"])
        ctx.set_cg_precision(32)
        conv32 = ctx.update(lam, 1e-30, 3000, *args, sigmaa=P["sa"])
        ctx.set_cg_preconditioner(3)
        with pytest.raises(CRTError) as ei:
            ctx.update(lam, 1e-12, 400, *args, sigmaa=P["sa"])
        assert ei.value.code == 903
    finally:
        ctx.set_cg_preconditioner(0)
        ctx.set_cg_precision(64)
        ctx.set_inversion_kind(0)
    e64, e32 = _rel(conv["dpar"], dpar_o), _rel(conv32["dpar"], dpar_o)
    print(f"[dc/fpi fast cg] kind={kind} ltri={ltri} fsai {prec}: ncg {stop['ncg']} vs Jacobi {jac['ncg']}; "
          f"converged {conv['ncg']} (FP64 J {e64:.2e}), {conv32['ncg']} (FP32 J {e32:.2e})", flush=True)
    assert e64 <= 1e-10 and e32 <= 1e-6
    assert np.all((conv32["dpar"].imag if kind == 1 else conv32["dpar"].real) == 0.0)
    assert stop["ncg"] < jac["ncg"]


@pytest.mark.parametrize("ltri", [0, 1])
def test_fpi_zero_rhs(ctx, oracle_lib, ltri):
    """cjggdc tests its stop before the first step as well (src/cg_mod.f90:143-150):
    a zero right-hand side (the FPI case with Im(bvec) = 0: data equal to the forward
    values and a real model) ends the solve with ncg = 0 and dpar = 0, as the oracle."""
    O = oracle_lib
    w = _workload(ltri)
    P = _cg_problem(O, w, ltri, 2)
    dat = P["sa"].copy()
    par = P["par"].real.astype(np.complex128)
    lam, eps, ncgmax = 3.0, 1e-4, 200
    cp = O.CGProblem(P["sens"], P["wmatd"], P["wdfak"], dat, P["sa"], par, P["smatm"], ltri, w.grid.nx,
                     lam, eps, ncgmax, nachbar=P["nachbar"], maxn=w.grid.maxn)
    cgfac_o, bvec_o, dpar_o, cgres_o, ncg_o = O.update_dc(cp, 2)
    assert np.all(bvec_o.imag == 0.0)
    assert ncg_o == 0 and np.all(dpar_o == 0.0)
    try:
        ctx.set_inversion_kind(2)
        ctx.set_grid(w.grid)
        ctx.set_sens(P["sens"])
        ctx.set_regularization(ltri, P["smatm"], nx=w.grid.nx, nachbar=P["nachbar"])
        r = ctx.update(lam, eps, ncgmax, P["wmatd"], P["wdfak"], dat, par, sigmaa=P["sa"])
        print(f"[fpi zero rhs] ltri={ltri} ncg gpu {r['ncg']} oracle {ncg_o}", flush=True)
        assert r["ncg"] == 0 and r["cgres"][0] == 0.0
        assert np.all(r["dpar"] == 0.0)
        assert np.array_equal(r["par"], par)
    finally:
        ctx.set_inversion_kind(0)


@pytest.mark.parametrize("fac,code", [(-1.0, 28), (-4.0, 27)])
def test_dc_error_codes(ctx, oracle_lib, fac, code):
    """The DC twins stop where the complex routines do not: scaldc on a non-positive
    diagonal entry (errnr 27, src/scaldc.f90:54-58) and choldc on a non-positive pivot
    (errnr 28, src/choldc.f90:66-70).  One cell's conductivity is made negative: at
    -1 times its value every diagonal entry stays positive but the matrix is
    indefinite (28), at -4 times the diagonal of its corner nodes turns negative (27).
    The oracle's DC forward returns the same codes in the same (wavenumber, electrode)
    order; the complex forward of the same model does not stop."""
    O = oracle_lib
    w = configs.small_quad()
    g = w.grid
    cx, cy = g.centroids()
    sig = w.sigma.real.astype(np.complex128)
    e = int(np.argmin((cx - cx.mean()) ** 2 + (cy + 2.0) ** 2))
    sig[e] *= fac
    kwn, kwi, _, _ = O.rwaven(g, w.survey)
    elbg, relbg, kg, lb = O.precal(g, kwn)
    with pytest.raises(RuntimeError, match=f"errnr {code}"):
        O.forward_dc(g, sig, elbg, relbg, kg, lb)
    try:
        ctx.set_inversion_kind(1)
        ctx.set_grid(g)
        ctx.set_survey(w.survey)
        ctx.rwaven()
        ctx.precal()
        ctx.set_sigma(sig)
        with pytest.raises(CRTError) as ei:
            ctx.forward()
        print(f"[dc error] fac={fac} code {ei.value.code}", flush=True)
        assert ei.value.code == code
    finally:
        ctx.set_inversion_kind(0)
    # the complex inversion factorises the same matrices without a stop
    ctx.set_sigma(sig)
    ctx.forward()
