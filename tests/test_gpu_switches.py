"""crmod.cfg's forward switches on the device against the oracle (which is pinned to the
reference build's outputs in tests/test_oracle_cpu.py): singularity removal (lsr:
refsig, potana and kompab's right-hand side, src/fem.f90:350-389), the analytic
solution alone (lana), the fictitious sink (lsink), and bkfak's geometric factors
with bvolt (wkfak).  Potentials per (source, wavenumber) row at 1e-10 relative; the
no-flow grid with a sink up to its near-constant mode as in test_gpu_forward.py."""
import numpy as np
import pytest

from crtomo_stable_amd import configs, lib

TOL = 1e-10

pytestmark = pytest.mark.gpu


def _rel_rows(a, b):
    a2 = a.reshape(-1, a.shape[-1]); b2 = b.reshape(-1, b.shape[-1])
    num = np.abs(a2 - b2).max(axis=1)
    den = np.abs(b2).max(axis=1)
    return float((num / np.where(den > 0, den, 1.0)).max())


def _workload(name):
    if name == "quad":
        return configs.small_quad()
    if name == "noflow":
        return configs.small_quad(mixed=False)
    return configs.c1(scale=0.2, nelec=12, nmeas=60, seed=7)


def _setup(ctx, w):
    ctx.set_grid(w.grid)
    ctx.set_survey(w.survey)
    ctx.rwaven()
    lb = ctx.precal()
    ctx.set_sigma(w.sigma)
    return lb


SWITCHES = {"lsr": dict(lsr=True), "lana": dict(lana=True), "lsink": dict(lsink=True),
            "lsr_lsink": dict(lsr=True, lsink=True)}


@pytest.mark.parametrize("name", ["quad", "tri"])
@pytest.mark.parametrize("sw", list(SWITCHES))
def test_switch_forward(ctx, oracle_lib, name, sw):
    O = oracle_lib
    w = _workload(name)
    g = w.grid
    nsink = g.sanz // 2 + 3
    opts = SWITCHES[sw]
    print(f"[switches] {name} {sw}: sanz={g.sanz} mb={g.mb} electrodes={g.eanz}", flush=True)
    _setup(ctx, w)
    try:
        ctx.set_fem_switches(lsr=opts.get("lsr", False), lana=opts.get("lana", False),
                             nsink=nsink if opts.get("lsink") else 0)
        ctx.forward()
        kpot = ctx.get_kpot()
        volt, sa = ctx.voltages(0, kfak=ctx.bkfak())
    finally:
        ctx.set_fem_switches()
    kwn, kwi, _, _ = O.rwaven(g, w.survey)
    elbg, relbg, kg, lb = O.precal(g, kwn)
    kref = O.forward_sw(g, w.sigma, kwn, elbg, relbg, kg, lb, lsr=opts.get("lsr", False),
                        lana=opts.get("lana", False), lsink=opts.get("lsink", False), nsink=nsink)
    kfak = O.bkfak(g, w.survey)
    vref, saref = O.bvolt(g, w.survey, O.rtrafo(kref, kwi), kfak)
    ek = _rel_rows(kpot, kref)
    ev = float(np.max(np.abs(sa - saref)) / np.max(np.abs(saref)))
    print(f"[switches] {name} {sw}: kpot {ek:.2e}, kfak * volt {ev:.2e}", flush=True)
    assert ek <= TOL and ev <= TOL
    # the switches change the potentials (not a silent plain forward)
    kplain = O.forward(g, w.sigma, elbg, relbg, kg, lb)
    assert _rel_rows(kref, kplain) > 1e-3


def test_switch_noflow_sink(ctx, oracle_lib):
    O = oracle_lib
    w = _workload("noflow")
    g = w.grid
    nsink = g.sanz // 2 + 3
    print(f"[switches] noflow lsink: sanz={g.sanz} mb={g.mb}", flush=True)
    lb = _setup(ctx, w)
    assert not lb
    try:
        ctx.set_fem_switches(nsink=nsink)
        ctx.forward()
        kpot = ctx.get_kpot()
    finally:
        ctx.set_fem_switches()
    kwn, kwi, _, _ = O.rwaven(g, w.survey)
    elbg, relbg, kg, lb = O.precal(g, kwn)
    kref = O.forward_sw(g, w.sigma, kwn, elbg, relbg, kg, lb, lsink=True, nsink=nsink)
    centered = lambda a: a - a.mean(axis=-1, keepdims=True)  # noqa: E731
    e_raw, e_c = _rel_rows(kpot, kref), _rel_rows(centered(kpot), centered(kref))
    print(f"[switches] noflow lsink: kpot raw {e_raw:.2e}, centred {e_c:.2e}", flush=True)
    assert e_raw <= 1e-8 and e_c <= TOL


def test_switch_refusals(ctx):
    """What the switches do not run on: lsr / lana in the DC inversion, the FP32
    forward, the condensed method, the batch; bkfak's errnr 93."""
    w = _workload("quad")
    _setup(ctx, w)
    try:
        ctx.set_fem_switches(lsr=True)
        with pytest.raises(lib.CRTError, match="903"):
            ctx.set_forward_method(1)
            ctx.forward()
        ctx.set_forward_method(-1)
        ctx.set_inversion_kind(1)
        ctx.set_sigma(w.sigma)
        with pytest.raises(lib.CRTError, match="903"):
            ctx.forward()
        ctx.set_inversion_kind(0)
        ctx.set_sigma(w.sigma)
        with pytest.raises(lib.CRTError, match="903"):
            lib.Batch([ctx])
    finally:
        ctx.set_forward_method(-1)
        ctx.set_inversion_kind(0)
        ctx.set_fem_switches()
    # potential electrodes M = N: the denominator of src/bkfak.f90:128 vanishes
    from crtomo_stable_amd.grid import Survey
    ctx.set_survey(Survey(np.array([1 + 2 * 10000], np.int32), np.array([3 + 3 * 10000], np.int32)))
    with pytest.raises(lib.CRTError, match="errnr 93"):
        ctx.bkfak()
    ctx.set_survey(w.survey)
