"""Mixed-precision forward: FP32 band factorisation on the matrix cores won back by
FP64 iterative refinement (crt_set_forward_precision(ctx, 32, ...)).

The refined potentials, voltages and sensitivities are held to the same bar as
the FP64 path: relative error <= 1e-10 against the CPU oracle (potentials and
sensitivities per source / measurement row, voltages per datum), and the
refinement must converge without falling back to the FP64 factorisation.
"""
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


def _setup(ctx, w):
    ctx.set_grid(w.grid)
    ctx.set_survey(w.survey)
    ctx.rwaven()
    lb = ctx.precal()
    ctx.set_sigma(w.sigma)
    return lb


def _forward32(ctx, w, mode=1):
    ctx.set_forward_precision(32)
    try:
        ctx.forward()
        sweeps, ratio, fallbacks = ctx.refine_info()
        volt, sa = ctx.voltages(mode)
        ctx.sens(mode)
        return dict(volt=volt, sa=sa, kpot=ctx.get_kpot(), sens=ctx.get_sens(), sweeps=sweeps, ratio=ratio,
                    fallbacks=fallbacks, times=ctx.times())
    finally:
        ctx.set_forward_precision(64)


def test_fp32_refined_c0_parity(ctx, oracle_lib):
    """C0 (60x20 quads, mb 34: three tiles of 16): every output against the oracle."""
    O = oracle_lib
    w = configs.c0()
    print(f"[fp32 c0] sanz={w.grid.sanz} mb={w.grid.mb}", flush=True)
    _setup(ctx, w)
    g = _forward32(ctx, w)
    print(f"[fp32 c0] sweeps={g['sweeps']} ratio={g['ratio']:.2e} fallbacks={g['fallbacks']}", flush=True)
    assert g["fallbacks"] == 0 and 1 <= g["sweeps"] <= 12
    kwn, kwi, _, _ = O.rwaven(w.grid, w.survey)
    elbg, relbg, kg, lb = O.precal(w.grid, kwn)
    kpot = O.forward(w.grid, w.sigma, elbg, relbg, kg, lb)
    volt, sa = O.bvolti(w.grid, w.survey, O.rtrafo(kpot, kwi))
    sens = O.bsensi(w.grid, w.survey, kwi, kpot, elbg, w.sigma, volt)
    assert _rel_rows(g["kpot"], kpot) <= TOL
    assert float(np.max(np.abs(g["volt"] - volt) / np.abs(volt))) <= TOL
    assert float(np.max(np.abs(g["sa"] - sa) / np.abs(sa))) <= TOL
    assert _rel_rows(g["sens"].T, sens.T) <= TOL


@pytest.mark.parametrize("nx,nz,mixed", [(120, 100, True), (60, 50, False)])
def test_fp32_refined_sampled(ctx, oracle_lib, nx, nz, mixed):
    """A wide band (streaming kernel at 11+ tiles) and a no-flow grid (one factor per
    wavenumber, the other sources through the shared factor): sampled systems
    against the oracle (no-flow: mean-removed, as the FP64 test of that case)."""
    O = oracle_lib
    w = configs.small_quad(nx=nx, nz=nz, nelec=8, mixed=mixed)
    D = -(-w.grid.mb // 16)
    print(f"[fp32 sampled] nx={nx} nz={nz} mixed={mixed} sanz={w.grid.sanz} mb={w.grid.mb} D={D}", flush=True)
    _setup(ctx, w)
    g = _forward32(ctx, w)
    print(f"[fp32 sampled] sweeps={g['sweeps']} ratio={g['ratio']:.2e} fallbacks={g['fallbacks']}", flush=True)
    if mixed:
        assert g["fallbacks"] == 0
    kwn, kwi, _, _ = O.rwaven(w.grid, w.survey)
    elbg, relbg, kg, lb = O.precal(w.grid, kwn)
    K, E = kwn.size, w.grid.eanz
    picks = [(1, 1), (E // 2, K // 2), (E, K)]
    ref = O.forward(w.grid, w.sigma, elbg, relbg, kg, lb, systems=picks)
    for l, k in picks:
        a = g["kpot"][k - 1, l - 1]; b = ref[k - 1, l - 1]
        if not mixed:
            a = a - a.mean(); b = b - b.mean()
        assert np.abs(a - b).max() / np.abs(b).max() <= TOL, (l, k)


def test_fp32_refined_vs_fp64_c1(ctx):
    """Full C1 (20k triangles, 832 systems): the refined forward against the FP64
    forward on the device, per voltage, sensitivity row and potential row."""
    w = configs.c1()
    print(f"[fp32 c1] sanz={w.grid.sanz} mb={w.grid.mb}", flush=True)
    _setup(ctx, w)
    ctx.forward()
    v64, sa64 = ctx.voltages(1)
    ctx.sens(1)
    s64 = ctx.get_sens()
    k64 = ctx.get_kpot()
    t64 = ctx.times()
    g = _forward32(ctx, w)
    ev = float(np.max(np.abs(g["volt"] - v64) / np.abs(v64)))
    es = _rel_rows(g["sens"].T, s64.T)
    ek = _rel_rows(g["kpot"], k64)
    t = g["times"]
    print(f"[fp32 c1] sweeps={g['sweeps']} ratio={g['ratio']:.2e} fallbacks={g['fallbacks']} volt {ev:.2e} "
          f"sens {es:.2e} kpot {ek:.2e}; fp64 factor {t64['factor']:.2f} solve {t64['solve']:.2f} ms; fp32 factor "
          f"{t['factor']:.2f} solve {t['solve']:.2f} refine {t['refine']:.2f} ms", flush=True)
    assert g["fallbacks"] == 0
    assert ev <= TOL and es <= TOL and ek <= TOL


def test_fp32_rejects_narrow_band(ctx):
    """Bands of one or two tiles have no FP32 kernel: 903, and the FP64 path stays."""
    w = configs.small_quad()
    _setup(ctx, w)
    if -(-w.grid.mb // 16) >= 3:
        pytest.skip("small_quad band is wide enough for the FP32 kernel")
    with pytest.raises(lib.CRTError) as ei:
        ctx.set_forward_precision(32)
    assert ei.value.code == 903
    ctx.forward()
