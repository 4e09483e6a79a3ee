"""The condensed forward (static condensation onto the mixed-boundary nodes) against
the band factorisation of every system (the reference's algorithm, src/chol.f90 +
src/vre.f90 per (electrode, wavenumber)) and against the oracle.

Both are FP64 direct solves of the same systems, so kpot, the voltages and the
Jacobian agree to rounding: relative 1e-10 here (the FP64 tolerance of the forward
tests; measured ~1e-13).  Workloads: the 60x20 quad grid (30 electrodes), a
triangle mesh of a quarter the cells, the 20k-triangle mesh with 64 electrodes, and the 400x100 quad grid
with 96 electrodes (sampled systems against the oracle there), the DC kind on the
60x20 and 400x100 quad grids and the triangle mesh, and a source electrode on a
mixed-boundary node.
"""
import numpy as np
import pytest

from crtomo_stable_amd import configs, lib

TOL = 1e-10

pytestmark = pytest.mark.gpu


def _rel_rows(a, b):
    a2 = a.reshape(-1, a.shape[-1])
    b2 = b.reshape(-1, b.shape[-1])
    num = np.abs(a2 - b2).max(axis=1)
    den = np.abs(b2).max(axis=1)
    return float((num / np.where(den > 0, den, 1.0)).max())


def _setup(w, method, kind=0):
    c = lib.Context(0)
    c.set_grid(w.grid)
    c.set_survey(w.survey)
    c.rwaven()
    c.precal()
    c.set_inversion_kind(kind)
    c.set_sigma(w.sigma)
    c.set_forward_method(method)
    return c


def _pair(w, kind=0, sens=True):
    out = []
    for method in (0, 1):
        c = _setup(w, method, kind)
        try:
            c.forward()
            assert c.forward_method() == method
            kp = c.get_kpot()
            v = c.voltages(1)[0]
            J = None
            if sens:
                c.sens(1)
                J = c.get_sens()
            out.append(dict(kpot=kp, volt=v, sens=J, times=c.times()))
        finally:
            c.close()
    return out


@pytest.mark.parametrize("name", ["c0", "mid_tri", "c1"])
def test_condensed_equals_band(name):
    if name == "c0":
        w = configs.c0()
    elif name == "mid_tri":
        w = configs.c1(scale=0.5, nelec=24, nmeas=200, seed=7)
    else:
        w = configs.c1()
    print(f"[condensed] {name}: sanz={w.grid.sanz} mb={w.grid.mb} electrodes={w.grid.eanz}", flush=True)
    b, c = _pair(w)
    ek = _rel_rows(c["kpot"], b["kpot"])
    ev = float(np.max(np.abs(c["volt"] - b["volt"]) / np.abs(b["volt"])))
    es = _rel_rows(c["sens"].T, b["sens"].T)
    tb = sum(b["times"][p] for p in ("diag", "factor", "solve"))
    tc = sum(c["times"][p] for p in ("diag", "factor", "solve", "prefix", "schur", "product"))
    print(f"[condensed] {name}: kpot {ek:.2e} volt {ev:.2e} sens {es:.2e}; forward band {tb:.2f} ms, "
          f"condensed {tc:.2f} ms ({c['times']})", flush=True)
    assert ek <= TOL and ev <= TOL and es <= TOL


def test_condensed_oracle_c0(oracle_lib):
    O = oracle_lib
    w = configs.c0()
    c = _setup(w, 1)
    try:
        c.forward()
        kp = c.get_kpot()
    finally:
        c.close()
    kwn, kwi, _, _ = O.rwaven(w.grid, w.survey)
    elbg, relbg, kg, lb = O.precal(w.grid, kwn)
    r = O.forward(w.grid, w.sigma, elbg, relbg, kg, lb)
    e = _rel_rows(kp, r)
    print(f"[condensed] c0 vs oracle {e:.2e}", flush=True)
    assert e <= TOL


@pytest.mark.parametrize("name", ["c0", "c1", "c3"])
def test_condensed_dc(name):
    """The DC kind: the condensed kernels run real arithmetic on the real parts (one MFMA
    chain per product), the band method the complex kernels on (Re sigma, 0); c1 has
    separators, so the prefix kernels run too; c3 (38 boundary tiles) runs the dense
    kernel on eight waves per system.  The imaginary parts stay exactly 0."""
    w = {"c0": configs.c0, "c1": configs.c1, "c3": configs.c3}[name]()
    b, c = _pair(w, kind=1, sens=False)
    ek = _rel_rows(c["kpot"], b["kpot"])
    ev = float(np.max(np.abs(c["volt"] - b["volt"]) / np.abs(b["volt"])))
    tb = sum(b["times"][p] for p in ("diag", "factor", "solve"))
    tc = sum(c["times"][p] for p in ("diag", "factor", "solve", "prefix", "schur", "product"))
    print(f"[condensed] dc {name}: kpot {ek:.2e} volt {ev:.2e}; forward band {tb:.2f} ms, condensed {tc:.2f} ms "
          f"({c['times']})", flush=True)
    assert ek <= TOL and ev <= TOL
    assert not np.any(c["kpot"].imag)


def test_condensed_boundary_electrode(oracle_lib):
    """Sources on mixed-boundary nodes (their right-hand side sits in the boundary
    block): three nodes of type-11 elements added to the electrodes."""
    O = oracle_lib
    from crtomo_stable_amd import grid as G
    w = configs.c0()
    g = w.grid
    off, bnodes = 0, []
    for t, n in zip(g.typ, g.nelanz):
        if t == 11:
            bnodes = list(dict.fromkeys(int(v) for v in g.nrel[off:off + n, :2].ravel()))
        off += int(n)
    extra = [v for v in (bnodes[0], bnodes[len(bnodes) // 2], bnodes[-1]) if v not in set(g.enr.tolist())]
    enr = np.concatenate([g.enr, np.array(extra, np.int32)])
    g2 = G.Grid(**{**g.__dict__, "enr": enr})
    ne = enr.size
    sv = G.Survey(strnr=np.array([1 * 10000 + ne, (ne - 1) * 10000 + 2, (ne - 2) * 10000 + ne], np.int32),
                  vnr=np.array([3 * 10000 + 4, 5 * 10000 + 6, 7 * 10000 + 8], np.int32))
    w2 = configs.Workload(name="boundary_electrodes", grid=g2, survey=sv, sigma=w.sigma)
    b, c = _pair(w2, sens=False)
    ek = _rel_rows(c["kpot"], b["kpot"])
    kwn, kwi, _, _ = O.rwaven(g2, sv)
    elbg, relbg, kg, lb = O.precal(g2, kwn)
    r = O.forward(g2, w.sigma, elbg, relbg, kg, lb)
    eo = _rel_rows(c["kpot"], r)
    print(f"[condensed] boundary electrodes {extra}: vs band {ek:.2e}, vs oracle {eo:.2e}", flush=True)
    assert len(extra) == 3
    assert ek <= TOL and eo <= TOL


def test_condensed_c3_sampled(oracle_lib):
    """400 x 100 quads, 96 electrodes: sampled systems against the oracle, the
    voltages against the band path."""
    O = oracle_lib
    w = configs.c3()
    b, c = _pair(w, sens=False)
    ek = _rel_rows(c["kpot"], b["kpot"])
    ev = float(np.max(np.abs(c["volt"] - b["volt"]) / np.abs(b["volt"])))
    tb = sum(b["times"][p] for p in ("diag", "factor", "solve"))
    tc = sum(c["times"][p] for p in ("diag", "factor", "solve", "prefix", "schur", "product"))
    print(f"[condensed] c3: kpot {ek:.2e} volt {ev:.2e}; forward band {tb:.1f} ms, condensed {tc:.1f} ms "
          f"({c['times']})", flush=True)
    kwn, kwi, _, _ = O.rwaven(w.grid, w.survey)
    elbg, relbg, kg, lb = O.precal(w.grid, kwn)
    K, E = kwn.size, w.grid.eanz
    picks = [(1, 1), (E // 2, K // 2), (E, K)]
    ref = O.forward(w.grid, w.sigma, elbg, relbg, kg, lb, systems=picks)
    for l, k in picks:
        a = c["kpot"][k - 1, l - 1]
        r = ref[k - 1, l - 1]
        e = float(np.abs(a - r).max() / np.abs(r).max())
        print(f"[condensed] c3 system (l={l}, k={k}) vs oracle {e:.2e}", flush=True)
        assert e <= TOL
    assert ek <= TOL and ev <= TOL


def test_condensed_not_applicable():
    """No type-11 elements: method 1 is refused (903), the default runs the band path."""
    w = configs.small_quad(mixed=False)
    c = _setup(w, 1)
    try:
        with pytest.raises(lib.CRTError) as ei:
            c.forward()
        assert ei.value.code == 903
        c.set_forward_method(-1)
        c.forward()
        assert c.forward_method() == 0
    finally:
        c.close()
