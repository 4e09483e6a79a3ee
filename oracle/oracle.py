"""ctypes front end of the CPU oracle (``_build/liboracle.so``) and of the
reference build (``_ref/libcrtomo_ref.so``).

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of bench.py import this module; the product never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libcrtomo_ref.so")
# the reference routines and harness with chol, vre, choldc and vredc replaced by the
# library's Fortran shims (crtomo_stable_amd/fortran, make -C oracle dropin)
DROPIN_PATH = os.path.join(HERE, "_ref", "libcrtomo_dropin.so")

_i = C.c_int
_d = C.c_double
_p = C.c_void_p


class _Grid(C.Structure):
    _fields_ = [("sanz", _i), ("mb", _i), ("typanz", _i), ("typ", _p), ("nelanz", _p),
                ("selanz", _p), ("maxn", _i), ("nrel", _p), ("x", _p), ("y", _p), ("rnr", _p),
                ("elanz", _i), ("relanz", _i), ("eanz", _i), ("enr", _p)]


class _CG(C.Structure):
    _fields_ = [("nanz", _i), ("manz", _i), ("ltri", _i), ("nx", _i), ("maxn", _i), ("ncgmax", _i),
                ("lam", _d), ("eps", _d), ("sens", _p), ("wmatd", _p), ("wdfak", _p), ("dat", _p),
                ("sigmaa", _p), ("par", _p), ("smatm", _p), ("nachbar", _p)]


def build(force: bool = False) -> None:
    if force or not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", HERE])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        for name in ("orc_bessk0", "orc_bessk1", "orc_bessi0", "orc_bessi1", "orc_gammln"):
            getattr(_lib, name).restype = _d
            getattr(_lib, name).argtypes = [_d]
        _lib.orc_gauleg.argtypes = [_d, _d, _p, _p, _i]
        _lib.orc_gaulag.argtypes = [_p, _p, _i, _d]
        _lib.orc_bsytop.restype = _d
    return _lib


def ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(_p)


class OracleGrid:
    """Keeps the numpy arrays alive for the C struct."""

    def __init__(self, g):
        self.g = g
        self._keep = [g.typ, g.nelanz, g.selanz, g.nrel, g.x, g.y, g.rnr, g.enr]
        self.s = _Grid(g.sanz, g.mb, g.typanz, ptr(g.typ), ptr(g.nelanz), ptr(g.selanz), g.maxn,
                       ptr(g.nrel), ptr(g.x), ptr(g.y), ptr(g.rnr), g.elanz, g.relanz, g.eanz,
                       ptr(g.enr))

    @property
    def ref(self):
        return C.byref(self.s)


def _chk(err: int, what: str) -> None:
    if err != 0:
        raise RuntimeError(f"oracle {what}: errnr {err}")


# ----------------------------------------------------------- scalar helpers
def bessel(name: str, x: float) -> float:
    return getattr(lib(), "orc_" + name)(float(x))


def gauleg(n: int, x1: float = 0.0, x2: float = 1.0):
    x = np.zeros(n); w = np.zeros(n)
    lib().orc_gauleg(x1, x2, ptr(x), ptr(w), n)
    return x, w


def gaulag(n: int, alf: float = 0.0):
    x = np.zeros(n); w = np.zeros(n)
    lib().orc_gaulag(ptr(x), ptr(w), n, alf)
    return x, w


# ----------------------------------------------------------- path functions
def rwaven(grid, survey):
    og = OracleGrid(grid)
    kwn = np.zeros(30); kwi = np.zeros(30)
    n = _i(); amin = _d(); amax = _d()
    err = lib().orc_rwaven(og.ref, survey.nanz, ptr(survey.strnr), ptr(survey.vnr), C.byref(n),
                           ptr(kwn), ptr(kwi), C.byref(amin), C.byref(amax))
    _chk(err, "rwaven")
    return kwn[: n.value].copy(), kwi[: n.value].copy(), amin.value, amax.value


def precal(grid, kwn, lsytop: bool = False):
    og = OracleGrid(grid)
    kwn = np.ascontiguousarray(kwn, np.float64)
    K = kwn.size
    elbg = np.zeros((K, grid.npk, grid.elanz))
    relbg = np.zeros((grid.npk, max(grid.relanz, 1)))
    kg = np.zeros((K, grid.eanz, max(grid.relanz, 1)))
    lbeta = _i(); sytop = _d()
    err = lib().orc_precal(og.ref, K, ptr(kwn), int(lsytop), ptr(elbg), ptr(relbg), ptr(kg),
                           C.byref(lbeta), C.byref(sytop))
    _chk(err, "precal")
    return elbg, relbg, kg, bool(lbeta.value)


def forward(grid, sigma, elbg, relbg, kg, lbeta, systems=None):
    """kpot (kwnanz, eanz, sanz) = the reference's kpot(sanz, eanz, kwnanz)."""
    og = OracleGrid(grid)
    sigma = np.ascontiguousarray(sigma, np.complex128)
    K = elbg.shape[0]
    kpot = np.zeros((K, grid.eanz, grid.sanz), np.complex128)
    if systems is None:
        n, sl, sk = 0, None, None
    else:
        sl = np.ascontiguousarray([s[0] for s in systems], np.int32)
        sk = np.ascontiguousarray([s[1] for s in systems], np.int32)
        n = sl.size
    err = lib().orc_forward(og.ref, ptr(sigma), K, ptr(elbg), ptr(relbg), ptr(kg), int(lbeta),
                            ptr(kpot), n, ptr(sl), ptr(sk))
    _chk(err, "forward")
    return kpot


def refsig(grid, sigma) -> complex:
    """sigma0 of src/refsig.f90."""
    og = OracleGrid(grid)
    s0 = np.zeros(1, np.complex128)
    lib().orc_refsig(og.ref, ptr(np.ascontiguousarray(sigma, np.complex128)), ptr(s0))
    return complex(s0[0])


def potana(grid, kwn, l, k, sigma0):
    """pota (sanz) of src/potana.f90 for electrode l, wavenumber k (1-based)."""
    og = OracleGrid(grid)
    pota = np.zeros(grid.sanz, np.complex128)
    s0 = np.array([sigma0], np.complex128)
    lib().orc_potana(og.ref, ptr(np.ascontiguousarray(kwn, np.float64)), int(l), int(k), ptr(s0), ptr(pota))
    return pota


def forward_sw(grid, sigma, kwn, elbg, relbg, kg, lbeta, lsr=False, lana=False, lsink=False, nsink=0):
    """kpot (kwnanz, eanz, sanz) of the forward loop of src/fem.f90:302-397 with the
    singularity removal (lsr), the analytic solution alone (lana) and the fictitious
    sink (lsink at node nsink, 1-based)."""
    og = OracleGrid(grid)
    sigma = np.ascontiguousarray(sigma, np.complex128)
    K = elbg.shape[0]
    kpot = np.zeros((K, grid.eanz, grid.sanz), np.complex128)
    err = lib().orc_forward_sw(og.ref, ptr(sigma), K, ptr(np.ascontiguousarray(kwn, np.float64)), ptr(elbg),
                               ptr(relbg), ptr(kg), int(lbeta), int(lsr), int(lana), int(lsink), int(nsink),
                               ptr(kpot))
    _chk(err, "forward_sw")
    return kpot


def bkfak(grid, survey):
    """kfak (nanz) of src/bkfak.f90."""
    og = OracleGrid(grid)
    kfak = np.zeros(survey.nanz)
    err = lib().orc_bkfak(og.ref, survey.nanz, ptr(survey.strnr), ptr(survey.vnr), ptr(kfak))
    _chk(err, "bkfak")
    return kfak


_boundary_keep = None


def set_boundary(nodes=None, values=None):
    """Prescribed boundary values for the oracle's forward loops (randb2 / randbdc2);
    None or empty removes them."""
    global _boundary_keep
    if nodes is None or len(nodes) == 0:
        _boundary_keep = None
        lib().orc_set_boundary(0, None, None)
        return
    nd = np.ascontiguousarray(nodes, np.int32)
    vals = np.ascontiguousarray(values, np.complex128)
    _boundary_keep = (nd, vals)
    lib().orc_set_boundary(nd.size, ptr(nd), ptr(vals))


def rtrafo(kpot, kwnwi):
    K, E, S = kpot.shape
    hpot = np.zeros((E, S), np.complex128)
    kwnwi = np.ascontiguousarray(kwnwi, np.float64)
    lib().orc_rtrafo(S, E, K, ptr(kwnwi), ptr(np.ascontiguousarray(kpot)), ptr(hpot))
    return hpot


def bvolti(grid, survey, hpot):
    volt = np.zeros(survey.nanz, np.complex128); sa = np.zeros_like(volt)
    err = lib().orc_bvolti(survey.nanz, ptr(survey.strnr), ptr(survey.vnr), ptr(grid.enr), grid.sanz,
                           ptr(np.ascontiguousarray(hpot)), ptr(volt), ptr(sa))
    _chk(err, "bvolti")
    return volt, sa


def bvolt(grid, survey, hpot, kfak=None):
    volt = np.zeros(survey.nanz, np.complex128); sa = np.zeros_like(volt)
    kf = None if kfak is None else np.ascontiguousarray(kfak, np.float64)
    lib().orc_bvolt(survey.nanz, ptr(survey.strnr), ptr(survey.vnr), ptr(grid.enr), grid.sanz,
                    ptr(np.ascontiguousarray(hpot)), ptr(kf), ptr(volt), ptr(sa))
    return volt, sa


def bsensi(grid, survey, kwnwi, kpot, elbg, sigma, volt):
    """sens (manz, nanz) = the reference's sens(nanz, manz)."""
    og = OracleGrid(grid)
    sens = np.zeros((grid.elanz, survey.nanz), np.complex128)
    lib().orc_bsensi(og.ref, survey.nanz, ptr(survey.strnr), ptr(survey.vnr), kpot.shape[0],
                     ptr(np.ascontiguousarray(kwnwi, np.float64)), ptr(np.ascontiguousarray(kpot)),
                     ptr(elbg), ptr(np.ascontiguousarray(sigma, np.complex128)),
                     ptr(np.ascontiguousarray(volt, np.complex128)), ptr(sens))
    return sens


def bsens(grid, survey, kwnwi, kpot, elbg, strom=None):
    og = OracleGrid(grid)
    st = np.ones(survey.nanz) if strom is None else np.ascontiguousarray(strom, np.float64)
    sens = np.zeros((grid.elanz, survey.nanz), np.complex128)
    lib().orc_bsens(og.ref, survey.nanz, ptr(survey.strnr), ptr(survey.vnr), ptr(st), kpot.shape[0],
                    ptr(np.ascontiguousarray(kwnwi, np.float64)), ptr(np.ascontiguousarray(kpot)),
                    ptr(elbg), ptr(sens))
    return sens


def kompab(grid, sigma, elbg, relbg, kg, l, k):
    og = OracleGrid(grid)
    a = np.zeros((grid.mb + 1) * grid.sanz, np.complex128)
    b = np.zeros(grid.sanz, np.complex128)
    err = lib().orc_kompab(og.ref, ptr(np.ascontiguousarray(sigma, np.complex128)), ptr(elbg), ptr(relbg),
                           ptr(kg), elbg.shape[0], l, k, ptr(a), ptr(b))
    _chk(err, "kompab")
    return a, b


def scalab(sanz, mb, a, b):
    fak = np.zeros(sanz)
    err = lib().orc_scalab(sanz, mb, ptr(a), ptr(b), ptr(fak))
    _chk(err, "scalab")
    return fak


def chol(sanz, mb, a):
    return lib().orc_chol(sanz, mb, ptr(a))


def vre(sanz, mb, a, b):
    pot = np.zeros(sanz, np.complex128)
    lib().orc_vre(sanz, mb, ptr(a), ptr(b), ptr(pot))
    return pot


def cholz(a_colmajor: np.ndarray):
    n = a_colmajor.shape[0]
    a = np.asfortranarray(a_colmajor.astype(np.complex128)).copy(order="F")
    p = np.zeros(n, np.complex128)
    err = lib().orc_cholz(a.ctypes.data_as(_p), ptr(p), n)
    return err, a, p


def bsmatmreg(grid, nx, nz, alfx, alfz):
    og = OracleGrid(grid)
    sm = np.zeros((3, grid.elanz))
    lib().orc_bsmatmreg(og.ref, nx, nz, _d(alfx), _d(alfz), ptr(sm))
    return sm


def bnachbar(grid):
    og = OracleGrid(grid)
    nb = np.zeros((grid.maxn + 1, grid.elanz), np.int32)
    lib().orc_bnachbar(og.ref, ptr(nb))
    return nb


def bsmatmtri(grid, nachbar, alfx, alfz):
    og = OracleGrid(grid)
    sm = np.zeros((grid.maxn + 1, grid.elanz))
    lib().orc_bsmatmtri(og.ref, ptr(np.ascontiguousarray(nachbar, np.int32)), _d(alfx), _d(alfz), ptr(sm))
    return sm


class CGProblem:
    """Arrays in the reference layouts: sens (manz, nanz) row-major = sens(nanz, manz)."""

    def __init__(self, sens, wmatd, wdfak, dat, sigmaa, par, smatm, ltri, nx, lam, eps, ncgmax,
                 nachbar=None, maxn=4):
        self.arrs = dict(
            sens=np.ascontiguousarray(sens, np.complex128), wmatd=np.ascontiguousarray(wmatd, np.float64),
            wdfak=np.ascontiguousarray(wdfak, np.int32), dat=np.ascontiguousarray(dat, np.complex128),
            sigmaa=np.ascontiguousarray(sigmaa, np.complex128), par=np.ascontiguousarray(par, np.complex128),
            smatm=np.ascontiguousarray(smatm, np.float64),
            nachbar=None if nachbar is None else np.ascontiguousarray(nachbar, np.int32))
        manz, nanz = self.arrs["sens"].shape
        a = self.arrs
        self.s = _CG(nanz, manz, ltri, nx, maxn, ncgmax, lam, eps, ptr(a["sens"]), ptr(a["wmatd"]),
                     ptr(a["wdfak"]), ptr(a["dat"]), ptr(a["sigmaa"]), ptr(a["par"]), ptr(a["smatm"]),
                     ptr(a["nachbar"]))
        self.manz, self.nanz = manz, nanz


def update_rhs(p: CGProblem):
    cgfac = np.zeros(p.manz); bvec = np.zeros(p.manz, np.complex128)
    lib().orc_update_rhs(C.byref(p.s), ptr(cgfac), ptr(bvec))
    return cgfac, bvec


def blam0(p: CGProblem, fpi=False, alfx=1.0, alfz=1.0):
    f = lib().orc_blam0
    f.restype = C.c_double
    return f(C.byref(p.s), int(bool(fpi)), _d(alfx), _d(alfz))


def buncert(p: CGProblem, lres=True, lcov2=True, gauss=False):
    """(err, dict of manz x manz matrices ata, ata_reg, cov, res, cov2) of
    src/bmcm_mod.f90:49-175 with lam = p.s.lam (res/cov2 None when not asked)."""
    n = p.manz
    m = [np.zeros((n, n)) for _ in range(5)]
    err = lib().orc_buncert(C.byref(p.s), int(bool(lres)), int(bool(lcov2)), int(bool(gauss)),
                            *[ptr(x) for x in m])
    names = ["ata", "ata_reg", "cov", "res", "cov2"]
    out = {k: x.T for k, x in zip(names, m)}  # column-major storage
    if not lres:
        out["res"] = None
    if not (lres and lcov2):
        out["cov2"] = None
    return err, out


def cjggra(p: CGProblem, cgfac, bvec):
    dpar = np.zeros(p.manz, np.complex128); cgres = np.zeros(p.s.ncgmax + 1)
    ncg = _i()
    lib().orc_cjggra(C.byref(p.s), ptr(cgfac), ptr(np.ascontiguousarray(bvec)), ptr(dpar), ptr(cgres),
                     C.byref(ncg))
    return dpar, cgres, ncg.value


def forward_dc(grid, sigma, elbg, relbg, kg, lbeta):
    """kpotdc (kwnanz, eanz, sanz) of the DC forward loop (src/inv.f90:355-398)."""
    og = OracleGrid(grid)
    sigma = np.ascontiguousarray(sigma, np.complex128)
    K = elbg.shape[0]
    kpot = np.zeros((K, grid.eanz, grid.sanz))
    err = lib().orc_forward_dc(og.ref, ptr(sigma), K, ptr(elbg), ptr(relbg), ptr(kg), int(lbeta), ptr(kpot))
    _chk(err, "forward_dc")
    return kpot


def rtrafo_dc(kpotdc, kwnwi):
    K, E, S = kpotdc.shape
    hpot = np.zeros((E, S))
    lib().orc_rtrafo_dc(S, E, K, ptr(np.ascontiguousarray(kwnwi, np.float64)), ptr(np.ascontiguousarray(kpotdc)),
                        ptr(hpot))
    return hpot


def bvolti_dc(grid, survey, hpotdc):
    volt = np.zeros(survey.nanz, np.complex128); sa = np.zeros_like(volt)
    err = lib().orc_bvolti_dc(survey.nanz, ptr(survey.strnr), ptr(survey.vnr), ptr(grid.enr), grid.sanz,
                              ptr(np.ascontiguousarray(hpotdc)), ptr(volt), ptr(sa))
    _chk(err, "bvolti_dc")
    return volt, sa


def bsendc(grid, survey, kwnwi, kpotdc, elbg, sigma, volt):
    """sensdc (manz, nanz) = the reference's sensdc(nanz, manz)."""
    og = OracleGrid(grid)
    sens = np.zeros((grid.elanz, survey.nanz))
    lib().orc_bsendc(og.ref, survey.nanz, ptr(survey.strnr), ptr(survey.vnr), kpotdc.shape[0],
                     ptr(np.ascontiguousarray(kwnwi, np.float64)), ptr(np.ascontiguousarray(kpotdc)),
                     ptr(elbg), ptr(np.ascontiguousarray(sigma, np.complex128)),
                     ptr(np.ascontiguousarray(volt, np.complex128)), ptr(sens))
    return sens


def update_dc(p: CGProblem, kind, sensdc=None):
    """The DC (kind 1, sensdc (manz, nanz) real) or FPI (kind 2, DBLE of p's sens)
    branches of update + cjggdc: (cgfac, bvec, dpar, cgres, ncg) with dpar scaled
    as src/update.f90:302-307 (FPI: dpar = i * CG solution)."""
    sd = None if sensdc is None else np.ascontiguousarray(sensdc, np.float64)
    cgfac = np.zeros(p.manz); bvec = np.zeros(p.manz, np.complex128)
    lib().orc_update_rhs_dc(C.byref(p.s), int(kind), ptr(sd), ptr(cgfac), ptr(bvec))
    dpar = np.zeros(p.manz, np.complex128); cgres = np.zeros(p.s.ncgmax + 1); ncg = _i()
    lib().orc_cjggdc(C.byref(p.s), int(kind), ptr(sd), ptr(cgfac), ptr(bvec), ptr(dpar), ptr(cgres),
                     C.byref(ncg))
    if kind == 2:
        dpar = 1j * dpar.real
    return cgfac, bvec, dpar * cgfac, cgres, ncg.value


def cg_matvec(p: CGProblem, cgfac, pvec):
    ap = np.zeros(p.nanz, np.complex128); out = np.zeros(p.manz, np.complex128)
    lib().orc_cg_matvec(C.byref(p.s), ptr(cgfac), ptr(np.ascontiguousarray(pvec, np.complex128)), ptr(ap),
                        ptr(out))
    return ap, out


def besp(grid):
    """Restatement of besp_elem (src/besp_elem.f90:21-72) in numpy: centroids as
    src/relem.f90:122-143 (sequential sums over the nodes), the distance to the last
    edge neighbour per cell, and the pair extremes with the reference's running-value
    rules (the minima start at 1e5; grid_minx / grid_miny never fall below the running
    grid_min).  Pairs in blocks of rows."""
    g = grid
    n, maxn = g.elanz, g.maxn
    cx = np.zeros(n); cy = np.zeros(n)
    iel = 0
    for t, ne, sel in zip(g.typ, g.nelanz, g.selanz):
        if t < 10:
            for m in range(sel):
                idx = g.nrel[iel:iel + ne, m] - 1
                cx[iel:iel + ne] = cx[iel:iel + ne] + g.x[idx]
                cy[iel:iel + ne] = cy[iel:iel + ne] + g.y[idx]
            cx[iel:iel + ne] /= sel
            cy[iel:iel + ne] /= sel
        iel += ne
    cx, cy = cx[:n], cy[:n]
    nb = bnachbar(g)  # (maxn + 1, elanz), last row the count
    abst = np.zeros(n)
    for j in range(maxn):
        k = nb[j]
        has = k > 0
        d = np.sqrt((cx[has] - cx[k[has] - 1]) ** 2 + (cy[has] - cy[k[has] - 1]) ** 2)
        abst[has] = d
    gmin, gmax, mnx, mxx, mny, mxy = 1e5, 0.0, 1e5, 0.0, 1e5, 0.0
    for i0 in range(0, n, 512):
        ax = np.abs(cx[i0:i0 + 512, None] - cx[None, :])
        ay = np.abs(cy[i0:i0 + 512, None] - cy[None, :])
        ar = np.sqrt(ax ** 2 + ay ** 2)
        off = np.ones_like(ar, bool)
        rows = np.arange(i0, min(i0 + 512, n))
        off[rows - i0, rows] = False
        gmin = min(gmin, ar[off].min()); gmax = max(gmax, ar[off].max())
        mnx = min(mnx, ax[off].min()); mxx = max(mxx, ax[off].max())
        mny = min(mny, ay[off].min()); mxy = max(mxy, ay[off].max())
    mit = abst.sum() / n
    std = np.sum(np.sqrt((abst - mit) ** 2)) / max(1, n - 1)
    srt = np.sort(abst)
    med = 0.5 * (srt[n // 2 - 1] + srt[n // 2]) if n % 2 == 0 else srt[n // 2]
    return np.array([abst.min(), abst.max(), mit, med, std, gmin, gmax, max(mnx, gmin), mxx, max(mny, gmin), mxy])


def cutmck(grid):
    perm = np.zeros(grid.sanz, np.int32); st = _i()
    mb = lib().orc_cutmck(grid.sanz, grid.typanz, ptr(grid.nelanz), ptr(grid.selanz), grid.maxn,
                          ptr(grid.nrel), grid.elanz, ptr(perm), C.byref(st))
    if mb < 0:
        raise RuntimeError("cutmck: disconnected graph")
    return perm, mb, st.value


# ----------------------------------------------------------- reference build
_ref = None


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    """The reference's own routines (amdflang build); raises if not built."""
    global _ref
    if _ref is None:
        if not ref_available():
            raise FileNotFoundError(REF_PATH)
        _ref = C.CDLL(REF_PATH)
    return _ref


class RefSession:
    """Fills the reference modules and drives its routines (oracle/ref_harness.f90)."""

    def __init__(self, grid, survey, lib=None):
        """``lib``: another build of the harness (the drop-in library, whose chol / vre /
        choldc / vredc are the GPU shims); the reference build by default."""
        self.grid, self.survey = grid, survey
        r = self.r = ref() if lib is None else lib
        self._keep = [grid.typ, grid.nelanz, grid.selanz, grid.nrel, grid.x, grid.y, grid.rnr, grid.enr]
        rnr = grid.rnr if grid.rnr.size else np.zeros(1, np.int32)
        self._keep.append(rnr)
        _chk(r.ref_set_grid(grid.sanz, grid.mb, grid.typanz, ptr(grid.typ), ptr(grid.nelanz),
                            ptr(grid.selanz), grid.maxn, ptr(grid.nrel), ptr(grid.x), ptr(grid.y),
                            ptr(rnr), grid.eanz, ptr(grid.enr)), "ref_set_grid")
        _chk(r.ref_set_data(survey.nanz, ptr(survey.strnr), ptr(survey.vnr)), "ref_set_data")

    def set_boundary(self, nodes=None, values=None):
        """boundary.dat values (randb2 / randbdc2 in the reference's forward loops)."""
        nd = np.ascontiguousarray([] if nodes is None else nodes, np.int32)
        vals = np.ascontiguousarray([] if values is None else values, np.complex128)
        self._bnd = (nd, vals)
        _chk(self.r.ref_set_boundary(nd.size, ptr(nd if nd.size else np.zeros(1, np.int32)),
                                     ptr(vals if vals.size else np.zeros(1, np.complex128))), "ref_set_boundary")

    def besp(self):
        """besp_elem's statistics (src/besp_elem.f90) in the order esp_min, esp_max,
        esp_mit, esp_med, esp_std, grid_min, grid_max, grid_minx, grid_maxx, grid_miny,
        grid_maxy (writes inv.gstat into the working directory)."""
        st = np.zeros(11)
        _chk(self.r.ref_besp(ptr(st)), "ref_besp")
        return st

    def rwaven(self):
        kwn = np.zeros(30); kwi = np.zeros(30); n = _i(); amin = _d(); amax = _d()
        _chk(self.r.ref_rwaven(C.byref(n), ptr(kwn), ptr(kwi), C.byref(amin), C.byref(amax)), "ref_rwaven")
        self.K = n.value
        return kwn[: n.value].copy(), kwi[: n.value].copy(), amin.value, amax.value

    def precal(self, lsytop=False):
        g = self.grid
        elbg = np.zeros((self.K, g.npk, g.elanz))
        relbg = np.zeros((g.npk, max(g.relanz, 1)))
        kg = np.zeros((self.K, g.eanz, max(g.relanz, 1)))
        lb = _i()
        _chk(self.r.ref_precal(int(lsytop), ptr(elbg), ptr(relbg), ptr(kg), C.byref(lb)), "ref_precal")
        return elbg, relbg, kg, bool(lb.value)

    def forward(self, sigma, nthreads=1):
        g = self.grid
        kpot = np.zeros((self.K, g.eanz, g.sanz), np.complex128)
        s = np.ascontiguousarray(sigma, np.complex128)
        _chk(self.r.ref_forward(ptr(s), int(nthreads), ptr(kpot)), "ref_forward")
        return kpot

    def refsig(self, sigma):
        s0 = np.zeros(1, np.complex128)
        _chk(self.r.ref_refsig(ptr(np.ascontiguousarray(sigma, np.complex128)), ptr(s0)), "ref_refsig")
        return complex(s0[0])

    def potana(self, l, k):
        """pota of electrode l, wavenumber k (1-based), with the sigma0 of refsig()."""
        pota = np.zeros(self.grid.sanz, np.complex128)
        _chk(self.r.ref_potana(int(l), int(k), ptr(pota)), "ref_potana")
        return pota

    def forward_sw(self, sigma, lsr=False, lana=False, lsink=False, nsink=0):
        g = self.grid
        kpot = np.zeros((self.K, g.eanz, g.sanz), np.complex128)
        s = np.ascontiguousarray(sigma, np.complex128)
        _chk(self.r.ref_forward_sw(ptr(s), int(lsr), int(lana), int(lsink), int(nsink), ptr(kpot)), "ref_forward_sw")
        return kpot

    def bkfak(self):
        """kfak of src/bkfak.f90 (writes tmp.kfak into the working directory)."""
        kfak = np.zeros(self.survey.nanz)
        _chk(self.r.ref_bkfak(ptr(kfak)), "ref_bkfak")
        return kfak

    def voltages(self, complex_inv=True):
        g = self.grid
        hpot = np.zeros((g.eanz, g.sanz), np.complex128)
        volt = np.zeros(self.survey.nanz, np.complex128); sa = np.zeros_like(volt)
        err = self.r.ref_voltages(int(complex_inv), ptr(hpot), ptr(volt), ptr(sa))
        _chk(err, "ref_voltages")
        return hpot, volt, sa

    def sens(self, inversion=True):
        sens = np.zeros((self.grid.elanz, self.survey.nanz), np.complex128)
        self.r.ref_sens(int(inversion), ptr(sens))
        return sens

    def bsmatm(self, ltri, nx=0, nz=0, alfx=1.0, alfz=1.0):
        g = self.grid
        ncol = 3 if ltri == 0 else g.maxn + 1
        sm = np.zeros((ncol, g.elanz))
        nb = np.zeros((g.maxn + 1, g.elanz), np.int32)
        self.r.ref_bsmatm(ltri, nx, nz, _d(alfx), _d(alfz), ptr(sm), ptr(nb))
        return sm, nb

    def blam0(self, fpi=False, alfx=1.0, alfz=1.0, dc=False):
        """lammax of src/blam0.f90 on the state of the last update() call (dc: the
        ldc branch over sensdc)."""
        f = self.r.ref_blam0
        f.restype = C.c_double
        return f(int(bool(fpi)), _d(alfx), _d(alfz), int(bool(dc)))

    def forward_dc(self, sigma):
        g = self.grid
        kpot = np.zeros((self.K, g.eanz, g.sanz))
        s = np.ascontiguousarray(sigma, np.complex128)
        _chk(self.r.ref_forward_dc(ptr(s), ptr(kpot)), "ref_forward_dc")
        return kpot

    def voltages_dc(self):
        g = self.grid
        hpot = np.zeros((g.eanz, g.sanz))
        volt = np.zeros(self.survey.nanz, np.complex128); sa = np.zeros_like(volt)
        _chk(self.r.ref_voltages_dc(ptr(hpot), ptr(volt), ptr(sa)), "ref_voltages_dc")
        return hpot, volt, sa

    def sens_dc(self):
        sens = np.zeros((self.grid.elanz, self.survey.nanz))
        self.r.ref_sens_dc(ptr(sens))
        return sens

    def buncert(self, lam, sigma, lres=True, lcov2=True, gauss=False):
        """src/bmcm_mod.f90:49 buncert on the state of the last update() call (J, W)
        and bsmatm() call (R): the main diagonals the reference writes to ata.diag,
        ata_reg.diag, cov1_m.diag, res_m.diag, cov2_m.diag, parsed back (first data
        column of each, cov1_m/cov2_m: all three).  The system must be positive
        definite: the reference's chold stops the process otherwise."""
        import shutil
        import tempfile
        d = tempfile.mkdtemp(prefix="crt_buncert_")
        try:
            s = np.ascontiguousarray(sigma, np.complex128)
            db = d.encode()
            err = self.r.ref_buncert(_d(lam), int(bool(lres)), int(bool(lcov2)), int(bool(gauss)), ptr(s),
                                   C.c_char_p(db), len(db))
            _chk(err, "ref_buncert")
            out = {}
            for key, fname, ncol in (("ata", "ata.diag", 1), ("ata_reg", "ata_reg.diag", 1),
                                     ("cov1", "cov1_m.diag", 3), ("res", "res_m.diag", 1),
                                     ("cov2", "cov2_m.diag", 3)):
                path = os.path.join(d, fname)
                if not os.path.exists(path):
                    continue
                lines = open(path).read().splitlines()
                n = int(float(lines[0].split()[0]))
                rows = [[float(t.replace("D", "E")) for t in ln.replace(",", " ").split()] for ln in lines[1:n + 1]]
                arr = np.array(rows)
                out[key] = arr[:, 0] if ncol == 1 else arr[:, :3]
            return out
        finally:
            shutil.rmtree(d, ignore_errors=True)

    def update(self, p: CGProblem, kind=0, sensdc=None):
        """src/update.f90 -> cjg; kind 1: DC (ldc, sensdc (manz, nanz)), kind 2: FPI (lfpi)."""
        a = p.arrs
        sd = np.zeros(1) if sensdc is None else np.ascontiguousarray(sensdc, np.float64)
        cgfac = np.zeros(p.manz); bvec = np.zeros(p.manz, np.complex128)
        dpar = np.zeros(p.manz, np.complex128); cgres = np.zeros(p.s.ncgmax + 1); ncg = _i()
        _chk(self.r.ref_update(p.s.ncgmax, _d(p.s.eps), _d(p.s.lam), ptr(a["sens"]), ptr(a["wmatd"]),
                              ptr(a["wdfak"]), ptr(a["dat"]), ptr(a["sigmaa"]), ptr(a["par"]), ptr(cgfac),
                              ptr(bvec), ptr(dpar), ptr(cgres), C.byref(ncg), int(kind), ptr(sd)),
             "ref_update")
        return cgfac, bvec, dpar, cgres, ncg.value


def ref_chol(sanz, mb, a):
    return ref().ref_chol(sanz, mb, ptr(a))


def ref_vre(sanz, mb, a, b):
    pot = np.zeros(sanz, np.complex128)
    ref().ref_vre(sanz, mb, ptr(a), ptr(b), ptr(pot))
    return pot
