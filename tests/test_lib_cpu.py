"""CPU checks of the HIP library's boundary: the shared object loads without a
GPU, exports every function declared in include/crtomo_hip.h, and its host-side
bookkeeping (CutMcK, argument validation) matches the oracle."""
import os
import re

import numpy as np
import pytest

from crtomo_stable_amd import configs, grid as G, lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(ROOT, "include", "crtomo_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(crt_[a-z_]+)\s*\(", src)))


def test_library_exports_header():
    L = lib.load()
    names = header_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    assert set(lib.EXPORTS) <= set(names)
    assert b"gfx950" in L.crt_version()


def test_library_cutmck_matches_oracle(oracle_lib):
    for w in (configs.c0(cutmck=oracle_lib.cutmck),
              configs.c1(scale=0.2, nelec=12, nmeas=60, seed=7, cutmck=oracle_lib.cutmck)):
        # undo nothing: renumber the already ordered grid once more with both implementations
        p1, mb1 = lib.cutmck(w.grid)
        p2, mb2, _ = oracle_lib.cutmck(w.grid)
        assert mb1 == mb2 and np.array_equal(p1, p2)


@pytest.mark.parametrize("name", ["minimal", "mixed_surf"])
def test_library_cutmck_fixtures(name):
    d = os.path.join(ROOT, "tests", "golden", "reference_fixtures", name)
    g = G.read_elem(os.path.join(d, "elem.dat.orig"))
    perm, mb = lib.cutmck(g)
    assert mb == G.read_elem(os.path.join(d, "elem.dat")).mb
    e_old = G.read_elec(os.path.join(d, "elec.dat.orig"))
    assert np.array_equal(perm[e_old - 1], G.read_elec(os.path.join(d, "elec.dat")))


def test_no_cpu_fallback_in_product():
    """The product package never imports the oracle."""
    pkg = os.path.join(ROOT, "crtomo_stable_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith(".py"):
                txt = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in re.sub(r"#.*", "", txt).replace("oracle restatement", ""), f


@pytest.mark.parametrize("name", ["small_quad", "small_tri"])
def test_bnachbar_vs_reference_golden(name):
    """Edge-sort neighbour table (host code of the library) against the table the
    reference's bnachbar (src/bnachbar.f90) wrote for the same grid, and against the
    oracle's pairwise scan; bit-exact."""
    import os
    from crtomo_stable_amd import configs, lib
    from oracle import oracle as O
    O.build()
    w = configs.small_quad(cutmck=O.cutmck) if name == "small_quad" else \
        configs.c1(scale=0.2, nelec=12, nmeas=60, seed=7, cutmck=O.cutmck)
    nb = lib.bnachbar(w.grid)
    assert np.array_equal(nb, O.bnachbar(w.grid))
    if name == "small_tri":
        ref = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_small_tri.npz"))
        assert np.array_equal(nb, ref["cg1_nachbar"])


def _tomodir(tmp, name):
    """A CRMod tomodir from the reference's committed fixture files; config.dat takes
    the configurations of the committed volt.dat."""
    import shutil
    src = os.path.join(os.path.dirname(__file__), "golden", "reference_fixtures", name)
    for d in ("exe", "grid", "rho", "config", "mod"):
        os.makedirs(os.path.join(tmp, d), exist_ok=True)
    shutil.copy(os.path.join(src, "elem.dat"), os.path.join(tmp, "grid", "elem.dat"))
    shutil.copy(os.path.join(src, "elec.dat"), os.path.join(tmp, "grid", "elec.dat"))
    shutil.copy(os.path.join(src, "rho.dat"), os.path.join(tmp, "rho", "rho.dat"))
    survey, _, _ = G.read_volt(os.path.join(src, "volt.dat"))
    with open(os.path.join(tmp, "config", "config.dat"), "w") as f:
        f.write(f"{survey.nanz}\n")
        for a, b in zip(survey.strnr, survey.vnr):
            f.write(f"{a} {b}\n")
    cfg = os.path.join(tmp, "exe", "crmod.cfg")
    with open(cfg, "w") as f:
        f.write("***FILES***\n../grid/elem.dat\n../grid/elec.dat\n../rho/rho.dat\n../config/config.dat\n"
                "F ! potentials ?\n../mod/pot/pot.dat\nT ! measurements ?\n../mod/volt.dat\n"
                "F ! sensitivities ?\n../mod/sens/sens.dat\nF ! another dataset ?\n1 ! 2D (=0) or 2.5D (=1)\n"
                "F ! fictitious sink ?\n1660 ! fictitious sink node number\nF ! boundary values ?\nboundary.dat\n")
    return cfg, src


@pytest.mark.parametrize("name", ["minimal", "mixed_surf"])
def test_tomodir_sizes(tmp_path, name):
    """The library's readers of crmod.cfg, elem.dat, elec.dat and config.dat against
    the Python readers on the reference's committed fixtures."""
    cfg, src = _tomodir(str(tmp_path), name)
    g = G.read_elem(os.path.join(src, "elem.dat"))
    enr = G.read_elec(os.path.join(src, "elec.dat"))
    survey, _, _ = G.read_volt(os.path.join(src, "volt.dat"))
    assert lib.tomodir_sizes(cfg) == (g.sanz, g.mb, g.elanz, enr.size, survey.nanz)


def test_write_volt_wdatm_format(tmp_path):
    """volt.dat writer: REAL*4 magnitude and 1000*phase, and the polarity rule of
    src/wdatm.f90:62-72 (a phase within rounding of +-pi is moved by pi and the
    potential electrodes swap), checked against an independent restatement."""
    rng = np.random.default_rng(4)
    n = 40
    strnr = rng.integers(1, 60, n) * 10000 + rng.integers(1, 60, n)
    vnr = rng.integers(1, 60, n) * 10000 + rng.integers(1, 60, n)
    mag = 10.0 ** rng.uniform(-4, 1, n)
    pha = rng.uniform(-0.05, 0.05, n)
    pha[::3] += np.pi * np.sign(rng.standard_normal(n))[::3]
    sa = mag * np.exp(1j * pha)
    path = str(tmp_path / "volt.dat")
    lib.write_volt(path, strnr, vnr, sa)
    s2, m2, p2 = G.read_volt(path)
    ph = np.arctan2(sa.imag, sa.real)
    npi = np.round(ph / np.pi) * np.pi
    flip = np.abs(npi) > 1e-12
    v_exp = np.where(flip, (vnr % 10000) * 10000 + vnr // 10000, vnr)
    assert np.array_equal(s2.strnr, strnr) and np.array_equal(s2.vnr, v_exp)
    assert np.array_equal(m2.astype(np.float32), np.abs(sa).astype(np.float32))
    assert np.array_equal(p2.astype(np.float32), (1e3 * (ph - npi)).astype(np.float32))
