"""The batched forward on the full 256-model workload (seeded quad meshes of 5k..40k
cells at aspect 4:1, 48..96 electrodes, Gaussian-random-field models; the workload the
bench's c2 line runs): one graph launch for every (electrode, wavenumber) system of
all models.

* Every model: the voltages of the batch equal those of a fresh context's own
  crt_forward bit for bit (same kernels on the same inputs; the batch's default kernel
  variants);
* a stratified sample (the smallest, the largest and every 32nd model): kpot of the
  batch equal to the fresh context's bit for bit, and one (electrode, wavenumber)
  system of four of them against the oracle at 1e-10.
The fresh contexts run with a 16 GB panel budget (CRT_PANEL_BUDGET_GB) so that they fit
beside the batch; the factor and solve kernels run one workgroup per system, so the
chunking does not change the results.
"""
import os

import numpy as np
import pytest

from crtomo_stable_amd import configs, lib

pytestmark = pytest.mark.gpu
NMODELS = 256


def _ctx(w):
    c = lib.Context(0)
    c.set_grid(w.grid)
    c.set_survey(w.survey)
    c.rwaven()
    c.lbeta = c.precal()
    c.set_sigma(w.sigma)
    c.set_forward_method(0)  # the batch runs the band factorisation of every system
    return c


def test_c2_batch_equals_single(oracle_lib):
    O = oracle_lib
    ws, cs = [], []
    try:
        for i in range(NMODELS):
            ws.append(configs.c2_model(i))
            cs.append(_ctx(ws[-1]))
            if (i + 1) % 32 == 0:
                print(f"[c2] {i + 1} models set up", flush=True)
        cells = np.array([w.grid.elanz for w in ws])
        sample = sorted({int(np.argmin(cells)), int(np.argmax(cells))} | set(range(0, NMODELS, 32)))
        with lib.Batch(cs) as b:
            b.forward()
            ms, nsys, _, _ = b.info()
            print(f"[c2] {nsys} systems in one graph launch, {ms:.0f} ms", flush=True)
            volt_b = [c.voltages(0)[0] for c in cs]
            kpot_b = {i: cs[i].get_kpot() for i in sample}
        for c in cs:
            c.close()
        cs = []
        os.environ["CRT_PANEL_BUDGET_GB"] = "16"
        try:
            for i, w in enumerate(ws):
                c = _ctx(w)
                c.forward()
                assert np.array_equal(c.voltages(0)[0], volt_b[i]), (i, w.name)
                if i in kpot_b:
                    assert np.array_equal(c.get_kpot(), kpot_b[i]), (i, w.name)
                c.close()
                if (i + 1) % 32 == 0:
                    print(f"[c2] {i + 1} models compared with their own forward", flush=True)
        finally:
            os.environ.pop("CRT_PANEL_BUDGET_GB", None)
        for i in sample[:4]:
            w = ws[i]
            kwn, _, _, _ = O.rwaven(w.grid, w.survey)
            elbg, relbg, kg, lb = O.precal(w.grid, kwn)
            l, k = w.grid.eanz // 2, (kwn.size + 1) // 2
            r = O.forward(w.grid, w.sigma, elbg, relbg, kg, lb, systems=[(l, k)])[k - 1, l - 1]
            a = kpot_b[i][k - 1, l - 1]
            err = float(np.abs(a - r).max() / np.abs(r).max())
            print(f"[c2] model {i} ({w.grid.elanz} cells) system (l={l}, k={k}) vs oracle {err:.2e}", flush=True)
            assert err <= 1e-10
    finally:
        for c in cs:
            c.close()
