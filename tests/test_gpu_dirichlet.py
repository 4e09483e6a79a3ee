"""GPU parity of the Dirichlet boundaries against the oracle (itself pinned to the
reference's randb / randdc / randb2 / randbdc2 in tests/test_oracle_cpu.py):

* type-13 elements (src/randb.f90:40-84): identity rows and columns of their nodes in
  every system; complex forward, voltages and Jacobian (bsensi) and the DC forward,
  1e-10 as the other forward tests;
* boundary.dat values of zero (src/randb2.f90 with rwd = 0) on a no-flow grid: the
  same Dirichlet nodes through crt_set_boundary_values;
* what the device refuses rather than run differently: non-zero boundary values
  (903), a source electrode on a Dirichlet node (903), a grid with mixed and
  Dirichlet boundaries together (101, as the reference's precal).
"""
import numpy as np
import pytest

from crtomo_stable_amd import configs, grid as G
from crtomo_stable_amd.lib import CRTError

TOL = 1e-10
pytestmark = pytest.mark.gpu


def _rel_rows(a, b):
    a2 = a.reshape(-1, a.shape[-1]); b2 = b.reshape(-1, b.shape[-1])
    den = np.abs(b2).max(axis=1)
    return float((np.abs(a2 - b2).max(axis=1) / np.where(den > 0, den, 1.0)).max())


def _setup(ctx, w):
    ctx.set_grid(w.grid)
    ctx.set_survey(w.survey)
    ctx.rwaven()
    lb = ctx.precal()
    ctx.set_sigma(w.sigma)
    return lb


def _oracle(O, w, with_volt=True):
    kwn, kwi, _, _ = O.rwaven(w.grid, w.survey)
    elbg, relbg, kg, lb = O.precal(w.grid, kwn)
    kp = O.forward(w.grid, w.sigma, elbg, relbg, kg, lb)
    volt = sens = None
    if with_volt:
        volt, sa = O.bvolti(w.grid, w.survey, O.rtrafo(kp, kwi))
        sens = O.bsensi(w.grid, w.survey, kwi, kp, elbg, w.sigma, volt)
    kd = O.forward_dc(w.grid, w.sigma, elbg, relbg, kg, lb)
    return kp, volt, sens, kd


@pytest.mark.parametrize("how", ["type13", "zero_values", "sources_on_dirichlet"])
def test_dirichlet_forward(ctx, oracle_lib, how):
    """``sources_on_dirichlet``: the first two source electrodes on Dirichlet nodes (the
    first source's system is assembled by kompab, whose right-hand side randb2 clears;
    the second reuses the factor with kompb's unit source, as the reference does)."""
    O = oracle_lib
    if how == "type13":
        w = configs.small_quad(dirichlet=True)
        nodes = None
    else:
        w = configs.small_quad(mixed=False)
        nodes = np.arange(w.grid.sanz - 30, w.grid.sanz, dtype=np.int32) + 1
        if how == "sources_on_dirichlet":
            nodes = np.concatenate([w.grid.enr[:2], nodes]).astype(np.int32)
    print(f"[dirichlet] {how} sanz={w.grid.sanz} mb={w.grid.mb}", flush=True)
    try:
        if nodes is not None:
            O.set_boundary(nodes, np.zeros(nodes.size, np.complex128))
        # a source on a Dirichlet node has zero voltages: bvolti's logarithm stops (82)
        kp, volt, sens, kd = _oracle(O, w, with_volt=how != "sources_on_dirichlet")
    finally:
        O.set_boundary()
    lb = _setup(ctx, w)
    assert not lb
    if nodes is not None:
        ctx.set_boundary_values(nodes, np.zeros(nodes.size, np.complex128))
    ctx.forward()
    assert _rel_rows(ctx.get_kpot(), kp) <= TOL
    if volt is not None:
        v, _ = ctx.voltages(1)
        assert float(np.max(np.abs(v - volt) / np.abs(volt))) <= TOL
        ctx.sens(1)
        assert _rel_rows(ctx.get_sens().T, sens.T) <= TOL
    try:
        ctx.set_inversion_kind(1)
        _setup(ctx, w)
        if nodes is not None:
            ctx.set_boundary_values(nodes, np.zeros(nodes.size, np.complex128))
        ctx.forward()
        assert _rel_rows(ctx.get_kpot(), kd) <= TOL
    finally:
        ctx.set_inversion_kind(0)
        ctx.set_boundary_values([], [])


def test_dirichlet_refusals(ctx):
    w = configs.small_quad(mixed=False)
    _setup(ctx, w)
    nodes = np.arange(w.grid.sanz - 30, w.grid.sanz, dtype=np.int32) + 1
    ctx.set_boundary_values(nodes, np.full(nodes.size, 1.0 + 0.5j))
    with pytest.raises(CRTError) as e1:
        ctx.forward()
    assert e1.value.code == 903
    with pytest.raises(CRTError) as e3:
        ctx.set_boundary_values([w.grid.sanz + 1], [0.0])
    assert e3.value.code == 30
    ctx.set_boundary_values([], [])
    ctx.forward()
    # mixed and Dirichlet boundary elements in one grid
    g = configs.small_quad().grid
    n11 = int(g.nelanz[-1])
    g2 = G.replace(g, typ=np.array(list(g.typ) + [13]),
                   nelanz=np.array(list(g.nelanz[:-1]) + [n11 // 2, n11 - n11 // 2]),
                   selanz=np.array(list(g.selanz) + [2]))
    ctx.set_grid(g2)
    ctx.set_survey(w.survey)
    ctx.rwaven()
    with pytest.raises(CRTError) as e4:
        ctx.precal()
    assert e4.value.code == 101
