"""Writes the FP64 golden vectors of tests/golden/*.npz from the reference's own
routines (the amdflang build of /root/reference/src made by oracle/Makefile, run
in this container through oracle/ref_harness.f90).  The committed files are data:
inputs are regenerated by crtomo_stable_amd.configs / the fixture grids, outputs
are what the reference computed.

    make -C oracle ref && python tests/golden/make_golden.py [name ...]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from crtomo_stable_amd import configs, grid as G  # noqa: E402
from oracle import oracle as O  # noqa: E402


def fixture_workload(name):
    d = os.path.join(HERE, "reference_fixtures", name)
    g = G.read_elem(os.path.join(d, "elem.dat"))
    g = G.replace(g, enr=G.read_elec(os.path.join(d, "elec.dat")))
    g = G.fix_rnr(g)
    survey, _, _ = G.read_volt(os.path.join(d, "volt.dat"))
    sigma = G.read_rho(os.path.join(d, "rho.dat"))
    return configs.Workload(name, g, survey, sigma)


def forward_golden(w, sens=True, bsens=True):
    R = O.RefSession(w.grid, w.survey)
    kwn, kwi, amin, amax = R.rwaven()
    elbg, relbg, kg, lb = R.precal()
    kpot = R.forward(w.sigma, nthreads=4)
    hpot, volt, sa = R.voltages(True)
    out = dict(kwn=kwn, kwnwi=kwi, amin=amin, amax=amax, lbeta=lb, volt=volt, sigmaa=sa,
               hpot_elec=hpot[:, w.grid.enr - 1], kpot_src1=kpot[:, 0, :])
    if sens:
        out["sens"] = R.sens(True)
    if sens and bsens:
        _, volt0, sa0 = R.voltages(False)
        out["volt_bvolt"] = volt0
        out["sens_bsens"] = R.sens(False)
    return out, R


def cg_golden(w, R, out, ltri, lam, eps, ncgmax, seed):
    rng = np.random.default_rng(seed)
    sa = out["sigmaa"]
    dat = sa + 0.02 * (rng.standard_normal(sa.size) + 0.1j * rng.standard_normal(sa.size))
    wmatd = 1.0 / (0.03 + 0.01 * rng.random(sa.size)) ** 2
    wdfak = np.ones(sa.size, np.int32)
    par = np.log(w.sigma) + 0.05 * rng.standard_normal(w.sigma.size)
    smatm, nachbar = R.bsmatm(ltri, w.grid.nx, w.grid.nz, 1.0, 1.0)
    cp = O.CGProblem(out["sens"], wmatd, wdfak, dat, sa, par, smatm, ltri, w.grid.nx, lam, eps, ncgmax,
                     nachbar=nachbar if ltri else None, maxn=w.grid.maxn)
    cgfac, bvec, dpar, cgres, ncg = R.update(cp)
    cgres[ncg + 1:] = 0.0  # the reference leaves the tail unset
    # starting lambda (src/blam0.f90) on the same Jacobian and weights
    lam0 = R.blam0(fpi=False, alfx=1.0, alfz=1.0)
    lam0_fpi = R.blam0(fpi=True, alfx=2.0, alfz=0.5)
    p = f"cg{ltri}_"
    return {p + "dat": dat, p + "wmatd": wmatd, p + "par": par, p + "smatm": smatm, p + "nachbar": nachbar,
            p + "cgfac": cgfac, p + "bvec": bvec, p + "dpar": dpar, p + "cgres": cgres, p + "ncg": ncg,
            p + "lam": lam, p + "eps": eps, p + "ncgmax": ncgmax, p + "lam0": lam0, p + "lam0_fpi": lam0_fpi}


def main(names):
    O.build()
    if not O.ref_available():
        raise SystemExit("reference build missing: make -C oracle ref")
    want = lambda n: not names or n in names  # noqa: E731
    # fixture grids of the reference's tests (CRMod voltages)
    for name in ("minimal", "mixed_surf"):
        if want(name):
            w = fixture_workload(name)
            out, _ = forward_golden(w, sens=(name == "minimal"))
            np.savez_compressed(os.path.join(HERE, f"ref_{name}.npz"), **out)
    # seeded synthetic grids: forward, sensitivities and one CG update
    if want("small_quad"):
        w = configs.small_quad(cutmck=O.cutmck)
        out, R = forward_golden(w)
        out.update(cg_golden(w, R, out, 0, 10.0, 1e-4, w.grid.elanz // 10, 1))
        np.savez_compressed(os.path.join(HERE, "ref_small_quad.npz"), **out)
    if want("small_quad_noflow"):
        # no mixed boundary: one factorisation per wavenumber serves every source
        w = configs.small_quad(mixed=False, cutmck=O.cutmck)
        out, R = forward_golden(w)
        np.savez_compressed(os.path.join(HERE, "ref_small_quad_noflow.npz"), **out)
    if want("switches"):
        # crmod.cfg's forward switches (src/fem.f90:177-203, 302-397): singularity
        # removal (refsig, potana, kompab's lsr terms), the analytic solution alone, the
        # fictitious sink; bkfak's geometric factors.  kpot of the first three sources.
        import tempfile
        out = {}
        for tag, w in (("quad", configs.small_quad(cutmck=O.cutmck)),
                       ("noflow", configs.small_quad(mixed=False, cutmck=O.cutmck)),
                       ("tri", configs.c1(scale=0.2, nelec=12, nmeas=60, seed=7, cutmck=O.cutmck))):
            R = O.RefSession(w.grid, w.survey)
            R.rwaven()
            R.precal()
            nsink = w.grid.sanz // 2 + 3
            out[tag + "_sigma0"] = R.refsig(w.sigma)
            out[tag + "_pota_2_3"] = R.potana(2, 3)
            out[tag + "_nsink"] = nsink
            runs = {"lsink": dict(lsink=True, nsink=nsink)}
            if tag != "noflow":
                runs.update(lsr=dict(lsr=True), lana=dict(lana=True), lsr_lsink=dict(lsr=True, lsink=True, nsink=nsink))
            for name, sw in runs.items():
                out[f"{tag}_kpot_{name}"] = R.forward_sw(w.sigma, **sw)[:, :3, :]
            cwd = os.getcwd()
            with tempfile.TemporaryDirectory() as d:
                os.chdir(d)
                try:
                    out[tag + "_kfak"] = R.bkfak()
                finally:
                    os.chdir(cwd)
        np.savez_compressed(os.path.join(HERE, "ref_switches.npz"), **out)
    if want("small_tri_sytop"):
        # boundary factors with the topography-aware singularity position (lsytop)
        w = configs.c1(scale=0.2, nelec=12, nmeas=60, seed=7, cutmck=O.cutmck)
        R = O.RefSession(w.grid, w.survey)
        kwn, kwi, _, _ = R.rwaven()
        elbg, relbg, kg, lb = R.precal(lsytop=True)
        R.forward(w.sigma, nthreads=4)
        hpot, volt, sa = R.voltages(True)
        np.savez_compressed(os.path.join(HERE, "ref_small_tri_sytop.npz"), kg=kg, volt=volt, sigmaa=sa,
                            lbeta=lb)
    if want("small_tri"):
        w = configs.c1(scale=0.2, nelec=12, nmeas=60, seed=7, cutmck=O.cutmck)
        out, R = forward_golden(w, bsens=False)
        out.pop("kpot_src1")
        out.update(cg_golden(w, R, out, 1, 3.0, 1e-12, 400, 2))
        np.savez_compressed(os.path.join(HERE, "ref_small_tri.npz"), **out)
    if want("buncert"):
        # model uncertainty (src/bmcm_mod.f90 buncert) on the CG problems of the two
        # grids above: ltri 1 through chold/linvd, ltri 0 through Gauss_dble (its
        # regularisation fill leaves ata_reg indefinite, where chold stops)
        out = {}
        for ltri, name, gauss in ((0, "small_quad", True), (1, "small_tri", False)):
            w = configs.small_quad(cutmck=O.cutmck) if ltri == 0 else \
                configs.c1(scale=0.2, nelec=12, nmeas=60, seed=7, cutmck=O.cutmck)
            ref = np.load(os.path.join(HERE, f"ref_{name}.npz"))
            p = f"cg{ltri}_"
            R = O.RefSession(w.grid, w.survey)
            smatm, nachbar = R.bsmatm(ltri, w.grid.nx, w.grid.nz, 1.0, 1.0)
            cp = O.CGProblem(ref["sens"], ref[p + "wmatd"], np.ones(ref["sigmaa"].size, np.int32), ref[p + "dat"],
                             ref["sigmaa"], ref[p + "par"], smatm, ltri, w.grid.nx, float(ref[p + "lam"]),
                             float(ref[p + "eps"]), 2, nachbar=nachbar if ltri else None, maxn=w.grid.maxn)
            R.update(cp)
            bu = R.buncert(float(ref[p + "lam"]), w.sigma, lres=True, lcov2=True, gauss=gauss)
            for k, v in bu.items():
                out[f"q{ltri}_{k}"] = v
            out[f"q{ltri}_gauss"] = gauss
        np.savez_compressed(os.path.join(HERE, "ref_buncert.npz"), **out)
    if want("dc"):
        # DC inversion (ldc) and final phase improvement (lfpi): kompadc..vredc forward,
        # rtrafo/bvolti DC branches, bsendc, update + cjggdc, blam0's ldc branch; every
        # second measurement with swapped potential electrodes (negative voltages); the
        # FPI update runs on the complex Jacobian of ref_small_quad / ref_small_tri
        out = {}
        for ltri in (0, 1):
            w = configs.small_quad(cutmck=O.cutmck) if ltri == 0 else \
                configs.c1(scale=0.2, nelec=12, nmeas=60, seed=7, cutmck=O.cutmck)
            s = w.survey
            vnr = s.vnr.copy(); e3, e4 = vnr % 10000, vnr // 10000
            sw = (np.arange(vnr.size) % 2 == 0) & (e3 > 0)
            vnr[sw] = e3[sw] * 10000 + e4[sw]
            s = G.Survey(strnr=s.strnr.copy(), vnr=vnr)
            R = O.RefSession(w.grid, s)
            R.rwaven(); R.precal()
            kp = R.forward_dc(w.sigma)
            hp, volt, sa = R.voltages_dc()
            sd = R.sens_dc()
            p = f"dc{ltri}_"
            out.update({p + "vnr": vnr, p + "kpot_src1": kp[:, 0, :], p + "hpot_elec": hp[:, w.grid.enr - 1],
                        p + "volt": volt, p + "sigmaa": sa, p + "sens": sd})
            Rc = O.RefSession(w.grid, w.survey)
            Rc.rwaven(); Rc.precal()
            Rc.forward(w.sigma, nthreads=4)
            _, vc, sac = Rc.voltages(True)
            sc = Rc.sens(True)
            smatm, nachbar = Rc.bsmatm(ltri, w.grid.nx, w.grid.nz, 1.0, 1.0)
            rng = np.random.default_rng(20 + ltri)
            wmatd = 1.0 / (0.03 + 0.01 * rng.random(sac.size)) ** 2
            par = np.log(w.sigma) + 0.05 * (rng.standard_normal(w.sigma.size) + 1j * rng.standard_normal(w.sigma.size))
            out[p + "wmatd"], out[p + "par"] = wmatd, par
            for kind in (1, 2):
                sa_k = sa if kind == 1 else sac
                dat = sa_k + 0.02 * (rng.standard_normal(sa_k.size) + 0.1j * rng.standard_normal(sa_k.size))
                cp = O.CGProblem(np.zeros_like(sc) if kind == 1 else sc, wmatd, np.ones(sa_k.size, np.int32), dat,
                                 sa_k, par, smatm, ltri, w.grid.nx, 3.0, 1e-12, 400,
                                 nachbar=nachbar if ltri else None, maxn=w.grid.maxn)
                cgfac, bvec, dpar, cgres, ncg = Rc.update(cp, kind=kind, sensdc=sd if kind == 1 else None)
                cgres[ncg + 1:] = 0.0
                q = f"{p}k{kind}_"
                out.update({q + "dat": dat, q + "cgfac": cgfac, q + "bvec": bvec, q + "dpar": dpar,
                            q + "cgres": cgres, q + "ncg": ncg})
                if kind == 1:
                    out[q + "lam0"] = Rc.blam0(dc=True)
        np.savez_compressed(os.path.join(HERE, "ref_dc.npz"), **out)
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main(sys.argv[1:])
