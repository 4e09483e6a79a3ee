"""GPU parity of the single-system drop-ins chol(a), vre(a,b,pot) and cholz(a,p,n)
against the oracle restatements of src/chol.f90, src/vre.f90 and src/cholz.f90, and of
the DC twins choldc(a), vredc(a,b,pot) (src/choldc.f90, src/vredc.f90).
Tolerance 1e-10 relative (FP64); error code 28 on a vanishing pivot."""
import numpy as np
import pytest

from crtomo_stable_amd import configs, grid as G

pytestmark = pytest.mark.gpu


def _band_system(O):
    w = configs.small_quad()
    kwn, kwi, _, _ = O.rwaven(w.grid, w.survey)
    elbg, relbg, kg, lb = O.precal(w.grid, kwn)
    a, b = O.kompab(w.grid, w.sigma, elbg, relbg, kg, 3, 2)
    O.scalab(w.grid.sanz, w.grid.mb, a, b)
    return w.grid.sanz, w.grid.mb, a, b


def test_chol_vre_dropin(ctx, oracle_lib):
    O = oracle_lib
    sanz, mb, a, b = _band_system(O)
    print(f"[dropin] chol/vre sanz={sanz} mb={mb}", flush=True)
    ag = a.copy()
    assert ctx.chol(sanz, mb, ag) == 0
    ao = a.copy()
    assert O.chol(sanz, mb, ao) == 0
    mask = np.zeros((sanz, mb + 1), bool)
    for i in range(sanz):
        mask[i, max(0, mb - i):] = True
    m = mask.ravel()
    assert np.max(np.abs(ag[m] - ao[m])) <= 1e-10 * np.max(np.abs(ao[m]))
    rng = np.random.default_rng(5)
    bb = rng.standard_normal(sanz) + 1j * rng.standard_normal(sanz)
    pg = ctx.vre(sanz, mb, ao, bb)
    po = O.vre(sanz, mb, ao, bb)
    assert np.max(np.abs(pg - po)) <= 1e-10 * np.max(np.abs(po))


def test_chol_pivot_error(ctx):
    """a zero row gives a vanishing pivot: errnr 28 (src/chol.f90:67-71)."""
    sanz, mb = 40, 5
    a = np.zeros(sanz * (mb + 1), np.complex128)
    for i in range(sanz):
        a[i * (mb + 1) + mb] = 4.0 + 0.1j
        if i > 0:
            a[i * (mb + 1) + mb - 1] = -1.0
    a[17 * (mb + 1): 18 * (mb + 1)] = 0.0
    assert ctx.chol(sanz, mb, a) == 28


def test_choldc_vredc_dropin(ctx, oracle_lib):
    """DC twins: the real band of the DC path (kompab with a real conductivity, scaled),
    choldc against the complex oracle chol on the same numbers and against L L^T = A,
    vredc against the oracle vre; a negative pivot gives errnr 28 (src/choldc.f90)."""
    O = oracle_lib
    w = configs.small_quad()
    kwn, kwi, _, _ = O.rwaven(w.grid, w.survey)
    elbg, relbg, kg, lb = O.precal(w.grid, kwn)
    sig = w.sigma.real.astype(np.complex128)
    a, b = O.kompab(w.grid, sig, elbg, relbg, kg, 3, 2)
    O.scalab(w.grid.sanz, w.grid.mb, a, b)
    sanz, mb = w.grid.sanz, w.grid.mb
    assert np.abs(a.imag).max() == 0.0
    ar = a.real.copy()
    assert ctx.choldc(sanz, mb, ar) == 0
    ao = a.copy()
    assert O.chol(sanz, mb, ao) == 0
    mask = np.zeros((sanz, mb + 1), bool)
    for i in range(sanz):
        mask[i, max(0, mb - i):] = True
    m = mask.ravel()
    assert np.max(np.abs(ar[m] - ao.real[m])) <= 1e-10 * np.max(np.abs(ao[m]))
    # L L^T reproduces the scaled matrix
    L = np.zeros((sanz, sanz)); A = np.zeros((sanz, sanz))
    for i in range(sanz):
        for j in range(max(0, i - mb), i + 1):
            L[i, j] = ar[i * (mb + 1) + mb - (i - j)]
            A[i, j] = A[j, i] = a.real[i * (mb + 1) + mb - (i - j)]
    assert np.max(np.abs(L @ L.T - A)) <= 1e-12 * np.max(np.abs(A))
    rng = np.random.default_rng(6)
    bb = rng.standard_normal(sanz)
    pg = ctx.vredc(sanz, mb, ar, bb)
    po = O.vre(sanz, mb, ao, bb.astype(np.complex128))
    assert np.max(np.abs(pg - po)) <= 1e-10 * np.max(np.abs(po))
    assert np.max(np.abs(pg - (-np.linalg.solve(A, bb)))) <= 1e-10 * np.max(np.abs(pg))
    # indefinite: one negative diagonal
    neg = np.zeros(40 * 6)
    for i in range(40):
        neg[i * 6 + 5] = 4.0 if i != 17 else -4.0
        if i > 0:
            neg[i * 6 + 4] = -1.0
    assert ctx.choldc(40, 5, neg) == 28


def test_cholz_dropin(ctx, oracle_lib):
    O = oracle_lib
    rng = np.random.default_rng(11)
    n = 60
    m = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    A = (m + m.T) / 2 + n * np.eye(n)  # complex symmetric, diagonally dominant
    err_g, ag, pg = ctx.cholz(np.asfortranarray(A))
    err_o, ao, po = O.cholz(A)
    assert err_g == 0 and err_o == 0
    low = np.tril(np.ones((n, n), bool), -1)
    assert np.max(np.abs(ag[low] - ao[low])) <= 1e-10 * np.max(np.abs(ao[low]))
    assert np.max(np.abs(pg - po)) <= 1e-10 * np.max(np.abs(po))
    # L L^T reproduces A (no conjugation)
    L = np.where(low, ag, 0) + np.diag(pg)
    assert np.max(np.abs(L @ L.T - A)) <= 1e-10 * np.max(np.abs(A))


@pytest.mark.parametrize("name", ["minimal", "mixed_surf"])
def test_tomodir_crmod_volt_dat(ctx, tmp_path, name):
    """A CRMod tomodir run by the library alone (crmod.cfg, elem.dat, elec.dat, rho.dat,
    config.dat read by its host code; forward on the device; volt.dat written as
    src/wdatm.f90 writes it) against the reference's committed volt.dat: the same
    configurations and polarity, magnitudes and phases within the file's REAL*4
    precision."""
    import os
    from crtomo_stable_amd import grid as G
    from test_lib_cpu import _tomodir
    cfg, src = _tomodir(str(tmp_path), name)
    print(f"[tomodir] {name}", flush=True)
    ctx.tomodir_crmod(cfg)
    s_out, m_out, p_out = G.read_volt(os.path.join(str(tmp_path), "mod", "volt.dat"))
    s_ref, m_ref, p_ref = G.read_volt(os.path.join(src, "volt.dat"))
    assert np.array_equal(s_out.strnr, s_ref.strnr) and np.array_equal(s_out.vnr, s_ref.vnr)
    assert np.max(np.abs(m_out - m_ref) / m_ref) < 1e-6
    assert np.max(np.abs(p_out - p_ref)) < 1e-3


@pytest.mark.parametrize("name", ["minimal", "mixed_surf"])
def test_tomodir_crmod_pot_sens_files(ctx, tmp_path, name):
    """The CRMod outputs of the potentials and sensitivities switches (src/bpot.f90 +
    src/wpot.f90, src/bsens.f90 + src/wsens.f90): one file per measurement, named with
    its number zero-filled to int(log10(nanz)) + 1 digits; the potential of each
    measurement's current dipole at every node, and per element the centroid and
    dV/dsigma, against the library's own hpot and J within the files' REAL*4."""
    import os
    from crtomo_stable_amd import grid as G, lib
    from test_lib_cpu import _tomodir
    cfg, src = _tomodir(str(tmp_path), name)
    txt = open(cfg).read().replace("F ! potentials ?", "T ! potentials ?").replace(
        "F ! sensitivities ?", "T ! sensitivities ?")
    open(cfg, "w").write(txt)
    for d in ("pot", "sens"):
        os.makedirs(os.path.join(str(tmp_path), "mod", d), exist_ok=True)
    print(f"[tomodir pot/sens] {name}", flush=True)
    ctx.tomodir_crmod(cfg)
    g = G.read_elem(os.path.join(src, "elem.dat"))
    # the electrodes too: the buffers of get_hpot are sized by them
    g.enr = np.ascontiguousarray(G.read_elec(os.path.join(src, "elec.dat")), np.int32)
    assert g.eanz > 0
    survey, _, _ = G.read_volt(os.path.join(src, "volt.dat"))
    nanz = survey.nanz
    nd = int(np.log10(np.float32(nanz))) + 1
    pots = sorted(os.listdir(os.path.join(str(tmp_path), "mod", "pot")))
    sens = sorted(os.listdir(os.path.join(str(tmp_path), "mod", "sens")))
    assert pots == [f"pot{i:0{nd}d}.dat" for i in range(1, nanz + 1)]
    assert sens == [f"sens{i:0{nd}d}.dat" for i in range(1, nanz + 1)]
    # the library's own arrays after the run (its context holds grid, survey, kpot, J)
    ctx.grid, ctx.survey = g, survey
    hpot = ctx.get_hpot()          # (eanz, sanz)
    J = ctx.get_sens()             # (elanz, nanz): dV/dsigma of the sensitivities switch
    cx, cy = g.centroids()
    for i in sorted({0, nanz // 2, nanz - 1}):
        e1, e2 = survey.strnr[i] % 10000, survey.strnr[i] // 10000
        p = np.zeros(g.sanz, np.complex128)
        if e2 > 0:
            p += hpot[e2 - 1]
        if e1 > 0:
            p -= hpot[e1 - 1]
        a = np.loadtxt(os.path.join(str(tmp_path), "mod", "pot", pots[i]))
        assert a.shape == (g.sanz, 4)
        assert np.allclose(a[:, 0], g.x.astype(np.float32), rtol=0, atol=1e-6 * np.abs(g.x).max())
        assert np.allclose(a[:, 1], g.y.astype(np.float32), rtol=0, atol=1e-6 * max(np.abs(g.y).max(), 1.0))
        assert np.max(np.abs(a[:, 2] + 1j * a[:, 3] - p)) <= 1e-6 * np.max(np.abs(p))
        b = np.loadtxt(os.path.join(str(tmp_path), "mod", "sens", sens[i]), skiprows=1)
        head = np.loadtxt(os.path.join(str(tmp_path), "mod", "sens", sens[i]), max_rows=1)
        assert b.shape == (g.elanz, 4)
        assert np.allclose(b[:, 0], cx, rtol=1e-6, atol=1e-6 * np.abs(cx).max())
        assert np.allclose(b[:, 1], cy, rtol=1e-6, atol=1e-6 * max(np.abs(cy).max(), 1.0))
        jr = J[:, i]
        assert np.max(np.abs(b[:, 2] + 1j * b[:, 3] - jr)) <= 1e-6 * np.max(np.abs(jr))
        s = jr.sum()
        assert abs(head[0] - abs(s)) <= 1e-6 * abs(s)
        assert abs(head[1] - 1e3 * np.angle(s)) <= 1e-3 * max(1.0, abs(1e3 * np.angle(s)))


def _dirichlet_rb_tomodir(tmp):
    """The reference's examples/DirichletRB/Modeling tomodir (3 x 1 quads, type-13
    sides and bottom, 2D, boundary.dat of zero values), measurements switched on."""
    import os
    import shutil
    src = os.path.join(os.path.dirname(__file__), "golden", "reference_fixtures", "dirichlet_rb")
    for d, files in (("grid", ("elem.dat", "elec.dat")), ("rho", ("rho.dat",)), ("config", ("config.dat",)),
                     ("exe", ("boundary.dat", "crmod.cfg")), ("mod", ())):
        os.makedirs(os.path.join(tmp, d), exist_ok=True)
        for f in files:
            shutil.copyfile(os.path.join(src, f), os.path.join(tmp, d, f))
    cfg = os.path.join(tmp, "exe", "crmod.cfg")
    lines = open(cfg).read().split("\n")
    lines[7] = "T                ! measurements ?"
    open(cfg, "w").write("\n".join(lines))
    return cfg, src


def test_tomodir_crmod_2d_dirichlet(ctx, oracle_lib, tmp_path):
    """crmod.cfg lines 12-17 (src/fem.f90:177-193): 2D (swrtr = 0: one wavenumber k = 0,
    kpot is the potential, src/fem.f90:241-250,393) with type-13 Dirichlet boundaries and
    boundary.dat values on the reference's DirichletRB example; volt.dat against the
    oracle's bvolt of the same 2D forward (REAL*4 of the file)."""
    import os
    O = oracle_lib
    cfg, src = _dirichlet_rb_tomodir(str(tmp_path))
    ctx.tomodir_crmod(cfg)
    survey, mag, pha = G.read_volt(os.path.join(str(tmp_path), "mod", "volt.dat"))
    g = G.read_elem(os.path.join(src, "elem.dat"))
    g = G.fix_rnr(G.replace(g, enr=G.read_elec(os.path.join(src, "elec.dat"))))
    sig = G.read_rho(os.path.join(src, "rho.dat"))
    kwn, kwi = np.zeros(1), np.full(1, np.pi)
    elbg, relbg, kg, lb = O.precal(g, kwn)
    bn = np.array([5, 6, 7, 8], np.int32)
    try:
        O.set_boundary(bn, np.zeros(4, np.complex128))
        kp = O.forward(g, sig, elbg, relbg, kg, lb)
    finally:
        O.set_boundary()
    hp = O.rtrafo(kp, kwi)
    assert np.array_equal(hp, kp[0])  # the weight pi makes the back-transform kpot itself
    _, sa = O.bvolt(g, survey, hp)
    assert np.allclose(mag, np.abs(sa), rtol=2e-7, atol=0)
    print(f"[tomodir 2D dirichlet] |V| {np.abs(sa)}", flush=True)


@pytest.mark.parametrize("line,text", [(11, "T ! another dataset ?")])
def test_tomodir_crmod_refuses_unsupported_switches(ctx, tmp_path, line, text):
    """The switch this call does not run (lagain: a second dataset in the same run)
    returns 903 instead of being ignored."""
    from crtomo_stable_amd.lib import CRTError
    from test_lib_cpu import _tomodir
    cfg, _ = _tomodir(str(tmp_path), "minimal")
    lines = open(cfg).read().rstrip("\n").split("\n")
    while len(lines) <= line:
        lines.append("")
    lines[line] = text
    open(cfg, "w").write("\n".join(lines) + "\n")
    with pytest.raises(CRTError) as ei:
        ctx.tomodir_crmod(cfg)
    assert ei.value.code == 903


@pytest.mark.parametrize("mswitch,lsink", [(6, True), (1, False), (2, True)])
def test_tomodir_crmod_switches(ctx, oracle_lib, tmp_path, mswitch, lsink):
    """crmod.cfg's mswitch (bit 0 lana, 1 wkfak, 2 lsr) and the fictitious sink run
    through the tomodir: volt.dat against the oracle's forward with the same switches
    (src/fem.f90:302-423: forward_sw, rtrafo, bkfak, bvolt), written by the same
    writer, within the file's REAL*4."""
    import os
    from crtomo_stable_amd import grid as G, lib
    from test_lib_cpu import _tomodir
    O = oracle_lib
    cfg, src = _tomodir(str(tmp_path), "minimal")
    g = G.read_elem(os.path.join(src, "elem.dat"))
    g = G.replace(g, enr=G.read_elec(os.path.join(src, "elec.dat")))
    g = G.fix_rnr(g)
    survey, _, _ = G.read_volt(os.path.join(src, "volt.dat"))
    sigma = G.read_rho(os.path.join(src, "rho.dat"))
    nsink = g.sanz // 2
    lines = open(cfg).read().rstrip("\n").split("\n")
    lines[13] = ("T" if lsink else "F") + " ! fictitious sink ?"
    lines[14] = f"{nsink} ! fictitious sink node number"
    lines.append(f"{mswitch:4d}")
    open(cfg, "w").write("\n".join(lines) + "\n")
    print(f"[tomodir switches] mswitch={mswitch} lsink={lsink} nsink={nsink}", flush=True)
    ctx.tomodir_crmod(cfg)
    s_out, m_out, p_out = G.read_volt(os.path.join(str(tmp_path), "mod", "volt.dat"))
    kwn, kwi, _, _ = O.rwaven(g, survey)
    elbg, relbg, kg, lb = O.precal(g, kwn)
    lsr, lana = bool(mswitch & 4), bool(mswitch & 1) and not (mswitch & 4)
    kp = O.forward_sw(g, sigma, kwn, elbg, relbg, kg, lb, lsr=lsr, lana=lana, lsink=lsink, nsink=nsink)
    kfak = O.bkfak(g, survey) if (mswitch & 2) and lb else None
    _, sa = O.bvolt(g, survey, O.rtrafo(kp, kwi), kfak)
    ref = os.path.join(str(tmp_path), "ref_volt.dat")
    lib.write_volt(ref, survey.strnr, survey.vnr, sa)
    s_ref, m_ref, p_ref = G.read_volt(ref)
    assert np.array_equal(s_out.strnr, s_ref.strnr) and np.array_equal(s_out.vnr, s_ref.vnr)
    em = float(np.max(np.abs(m_out - m_ref) / m_ref))
    ep = float(np.max(np.abs(p_out - p_ref)))
    print(f"[tomodir switches] |V| {em:.2e}, phase {ep:.2e} mrad", flush=True)
    assert em < 1e-6 and ep < 1e-3
