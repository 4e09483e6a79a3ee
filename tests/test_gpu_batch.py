"""Batched forward (crt_batch_*): every context's systems assembled, factorised and
solved in chunks through one shared workspace, captured once as a hipGraph and
replayed.

The batch runs the same kernels on the same inputs as crt_forward, so kpot and the
voltages must equal each context's own crt_forward bit for bit (default kernel
variants); with the autotuner the factor kernel may differ per band width, which
changes rounding only (relative 1e-10, the FP64 tolerance of the forward tests).
One system per model is also checked against the oracle.
"""
import numpy as np
import pytest

from crtomo_stable_amd import configs, lib

TOL = 1e-10

pytestmark = pytest.mark.gpu


def _models():
    return [
        configs.small_quad(nx=24, nz=8, nelec=10),                  # band of 1 tile
        configs.c2_model(0, min_cells=1500, max_cells=1500),        # 3 tiles
        configs.small_quad(nx=30, nz=10, nelec=8, mixed=False),     # no-flow: one factor per wavenumber
        configs.c2_model(5, min_cells=6000, max_cells=6000),        # 5 tiles
        configs.c2_model(3, min_cells=2000, max_cells=2000),        # 3 tiles again (shares the tuned variant)
        configs.small_quad(nx=100, nz=100, nelec=6),                # 11 tiles (streaming kernel)
    ]


def _ctx(w):
    c = lib.Context(0)
    c.set_grid(w.grid)
    c.set_survey(w.survey)
    c.rwaven()
    c.lbeta = c.precal()
    c.set_sigma(w.sigma)
    c.set_forward_method(0)  # the batch runs the band factorisation of every system
    return c


def _rel_rows(a, b):
    a2 = a.reshape(-1, a.shape[-1]); b2 = b.reshape(-1, b.shape[-1])
    num = np.abs(a2 - b2).max(axis=1)
    den = np.abs(b2).max(axis=1)
    return float((num / np.where(den > 0, den, 1.0)).max())


@pytest.fixture(scope="module")
def models():
    ws = _models()
    cs = [_ctx(w) for w in ws]
    for w, c in zip(ws, cs):
        print(f"[batch] {w.name} sanz={w.grid.sanz} mb={w.grid.mb} D={-(-w.grid.mb // 16)} "
              f"eanz={w.grid.eanz} K={c.kwnanz}", flush=True)
    yield ws, cs
    for c in cs:
        c.close()


def _single(cs):
    out = []
    for c in cs:
        c.forward()
        out.append((c.get_kpot(), c.voltages(1)[0]))
    return out


def test_batch_equals_single(models):
    ws, cs = models
    ref = _single(cs)
    with lib.Batch(cs) as b:
        for rep in range(3):  # capture, then two replays
            b.forward()
            for (kp, v), c, w in zip(ref, cs, ws):
                got = c.get_kpot()
                assert np.array_equal(got, kp), (w.name, rep)
                assert np.array_equal(c.voltages(1)[0], v), (w.name, rep)
        ms, n, variant, _ = b.info()
        assert n == sum(c.kwnanz * (w.grid.eanz if c.lbeta else 1) for c, w in zip(cs, ws))
        assert ms > 0.0
        print(f"[batch] {n} systems, {ms:.3f} ms per graph launch", flush=True)


def test_batch_chunks_and_new_sigma(models, oracle_lib):
    """A workspace that holds a few systems only (many chunks per model), and a
    second sigma set between replays of the captured graph."""
    ws, cs = models
    per_sys = max(16 * w.grid.sanz * (-(-w.grid.mb // 16) + 1) * 16 + 64 * w.grid.sanz for w in ws)
    with lib.Batch(cs, workspace_bytes=3 * per_sys) as b:
        b.forward()
        ref = _single(cs)
        b.forward()
        for (kp, _), c, w in zip(ref, cs, ws):
            assert np.array_equal(c.get_kpot(), kp), w.name
        sig2 = [w.sigma * (1.5 + 0.25j * (i % 3)) for i, w in enumerate(ws)]
        for c, s in zip(cs, sig2):
            c.set_sigma(s)
        b.forward()
        got = [c.get_kpot() for c in cs]
    ref2 = _single(cs)
    for g, (kp, _), w in zip(got, ref2, ws):
        assert np.array_equal(g, kp), w.name
    # anchor: the last electrode at the middle wavenumber of each model against the oracle
    O = oracle_lib
    for g, w, c, s in zip(got, ws, cs, sig2):
        kwn, _, _, _ = O.rwaven(w.grid, w.survey)
        elbg, relbg, kg, lb = O.precal(w.grid, kwn)
        l, k = w.grid.eanz, (kwn.size + 1) // 2
        r = O.forward(w.grid, s, elbg, relbg, kg, lb, systems=[(l, k)])[k - 1, l - 1]
        a = g[k - 1, l - 1]
        if not c.lbeta:
            a = a - a.mean(); r = r - r.mean()
        assert np.abs(a - r).max() / np.abs(r).max() <= TOL, w.name
    for c, w in zip(cs, ws):
        c.set_sigma(w.sigma)


def test_batch_autotune(models):
    ws, cs = models
    ref = _single(cs)
    with lib.Batch(cs, autotune=True) as b:
        b.forward()
        _, _, variant, tune = b.info()
        for D in range(17):
            if variant[D]:
                print(f"[batch] D={D}: variant {variant[D]}, ms {tune[D]}", flush=True)
        for (kp, v), c, w in zip(ref, cs, ws):
            assert _rel_rows(c.get_kpot(), kp) <= TOL, w.name
            got_v = c.voltages(1)[0]
            assert float(np.max(np.abs(got_v - v) / np.abs(v))) <= TOL, w.name


def test_batch_mixed_precision(models):
    """The mixed-precision variant (FP32 factorisation on the matrix cores, FP64
    refinement captured in the graph, a sweep skipped once the device check accepts the
    chunk): the refined voltages within 1e-10 of each context's FP64 forward, every
    model refined rather than redone in FP64, the same again on a replay."""
    ws, cs = models
    pick = [i for i, w in enumerate(ws) if 3 <= -(-w.grid.mb // 16) <= 16]
    assert len(pick) >= 3
    ref = []
    for i in pick:
        cs[i].forward()
        ref.append(cs[i].voltages(1)[0])
    with lib.Batch([cs[i] for i in pick], mixed=True) as b:
        for rep in range(2):
            b.forward()
            sweeps, ratio, fallbacks = b.refine_info()
            print(f"[batch mixed] replay {rep}: {b.info()[0]:.2f} ms, sweeps {sweeps}, ratio {ratio:.2e}, "
                  f"fallbacks {fallbacks}", flush=True)
            assert fallbacks == 0 and 1 <= sweeps <= 12 and ratio <= 1e-12
            for i, r in zip(pick, ref):
                v = cs[i].voltages(1)[0]
                err = float(np.abs(v / r - 1).max())
                print(f"[batch mixed]   {ws[i].name}: max relative voltage error {err:.2e}", flush=True)
                assert err < TOL
