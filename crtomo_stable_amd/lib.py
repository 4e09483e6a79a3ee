"""ctypes binding of the HIP library (``_lib/libcrtomo_hip.so``, C-ABI in
``include/crtomo_hip.h``).

The library is the product: there is no CPU fallback.  Loading fails loudly if
the shared object has not been built (``python -m crtomo_stable_amd.build``).
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libcrtomo_hip.so")

EXPORTS = (
    "crt_create", "crt_destroy", "crt_version", "crt_set_grid", "crt_set_survey", "crt_rwaven",
    "crt_set_wavenumbers", "crt_precal", "crt_get_precal", "crt_set_sigma", "crt_forward", "crt_get_kpot",
    "crt_voltages", "crt_get_hpot", "crt_sens", "crt_get_sens", "crt_cutmck", "crt_chol", "crt_vre",
    "crt_cholz", "crt_get_times", "crt_set_sens", "crt_set_regularization", "crt_cjg", "crt_update",
)

ERRORS = {
    18: "unknown element type", 19: "band width exceeded", 26: "element geometry (det <= 0)",
    27: "zero diagonal", 28: "non-positive pivot in Cholesky", 35: "boundary geometry in beta",
    46: "electrode number out of range", 50: "model cells differ from grid cells",
    61: "max. iterations out of range", 73: "duplicate electrodes", 82: "zero voltage", 88: "non-positive data error",
    89: "regular grid needs nx, nz >= 2", 91: "invalid background model", 95: "phase beyond 1000 pi",
    99: "no valid data", 100: "no starting model", 102: "no mixed boundary and no sink", 104: "invalid error model",
    105: "invalid phase error model", 106: "mixed boundary in 2D", 108: "factorisation rejected ata_reg",
    110: "mixed boundary elements in 2D", 93: "vanishing geometric factor", 97: "allocation failed",
    900: "HIP runtime error",
    901: "call order (missing set-up call)", 902: "invalid argument", 903: "unsupported configuration",
}

_lib = None
_p = C.c_void_p


class CRTError(RuntimeError):
    def __init__(self, code: int, where: str):
        super().__init__(f"{where}: errnr {code} ({ERRORS.get(code, 'unknown')})")
        self.code = code


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP library not built: {LIB_PATH} (run python -m crtomo_stable_amd.build)")
        lib = C.CDLL(LIB_PATH)
        lib.crt_version.restype = C.c_char_p
        lib.crt_destroy.restype = None
        lib.crt_destroy.argtypes = [_p]
        _lib = lib
    return _lib


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(_p)


def _chk(code: int, where: str) -> None:
    if code != 0:
        raise CRTError(code, where)


def cutmck(grid):
    """Cuthill-McKee renumbering on the host (``cutmckee/cutmck.f:126-337``)."""
    lib = load()
    perm = np.zeros(grid.sanz, np.int32)
    mb = C.c_int()
    _chk(lib.crt_cutmck(grid.sanz, grid.typanz, _ptr(grid.nelanz), _ptr(grid.selanz), grid.maxn,
                        _ptr(grid.nrel), grid.elanz, _ptr(perm), C.byref(mb)), "crt_cutmck")
    return perm, mb.value


def bnachbar(grid):
    """Edge neighbours of the area elements (``src/bnachbar.f90:44-90``), host code:
    (maxn+1, elanz) = reference nachbar(elanz, maxn+1); last row = neighbour count."""
    lib = load()
    nrel = np.ascontiguousarray(np.asarray(grid.nrel, np.int32).reshape(-1, grid.maxn)[:grid.elanz])
    nb = np.zeros((grid.maxn + 1, grid.elanz), np.int32)
    _chk(lib.crt_bnachbar(grid.elanz, grid.maxn, _ptr(nrel), _ptr(nb)), "crt_bnachbar")
    return nb


def tomodir_sizes(cfg_path: str):
    """(sanz, mb, elanz, eanz, nanz) of a CRMod tomodir, read by the library's host code."""
    lib = load()
    v = [C.c_int() for _ in range(5)]
    _chk(lib.crt_tomodir_sizes(cfg_path.encode(), *[C.byref(x) for x in v]), "crt_tomodir_sizes")
    return tuple(x.value for x in v)


def write_volt(path: str, strnr, vnr, sigmaa):
    """volt.dat as ``src/wdatm.f90:58-76`` writes it (library host code)."""
    lib = load()
    st = np.ascontiguousarray(strnr, np.int32); vn = np.ascontiguousarray(vnr, np.int32)
    sa = np.ascontiguousarray(sigmaa, np.complex128)
    _chk(lib.crt_write_volt(path.encode(), int(st.size), _ptr(st), _ptr(vn), _ptr(sa)), "crt_write_volt")


class CrtomoCfg(C.Structure):
    """crt_crtomo_cfg (include/crtomo_hip.h): crtomo.cfg as src/rall.f90:141-593 reads it."""
    _P = C.c_char * 81
    _fields_ = [("mswitch", C.c_int), ("elem", _P), ("elec", _P), ("volt", _P), ("ramd", _P), ("ldiff", C.c_int),
                ("dd0", _P), ("dm0", _P), ("dfm0", _P), ("dstart", _P), ("iseedpri", C.c_int),
                ("modl_stdn", C.c_double), ("nx", C.c_int), ("nz", C.c_int), ("alfx", C.c_double),
                ("alfz", C.c_double), ("itmax", C.c_int), ("ldc", C.c_int), ("lrobust", C.c_int),
                ("lfphai", C.c_int), ("stabw0", C.c_double), ("stabm0", C.c_double), ("stabpA1", C.c_double),
                ("stabpB", C.c_double), ("stabpA2", C.c_double), ("stabp0", C.c_double), ("lrho0", C.c_int),
                ("bet0", C.c_double), ("pha0", C.c_double), ("lagain", C.c_int), ("swrtr", C.c_int),
                ("lsink", C.c_int), ("nsink", C.c_int), ("lrandb2", C.c_int), ("drandb", _P), ("ltri", C.c_int),
                ("llamf", C.c_int), ("lamfix", C.c_double), ("betamgs", C.c_double), ("iseed", C.c_int),
                ("lffhom", C.c_int), ("lnse", C.c_int), ("lnse2", C.c_int), ("lnsepri", C.c_int),
                ("lstart", C.c_int), ("lprior", C.c_int), ("lsens", C.c_int), ("lcov1", C.c_int), ("lres", C.c_int),
                ("lcov2", C.c_int), ("lgauss", C.c_int), ("lelerr", C.c_int), ("lphi0", C.c_int),
                ("lsytop", C.c_int), ("lvario", C.c_int), ("lverb", C.c_int)]


def read_crtomo_cfg(cfg_path: str) -> dict:
    """crtomo.cfg read by the library's host code (``src/rall.f90:141-593``); raises
    CRTError with the reference's errnr for a file the reference rejects."""
    lib = load()
    c = CrtomoCfg()
    _chk(lib.crt_read_crtomo_cfg(cfg_path.encode(), C.byref(c)), "crt_read_crtomo_cfg")
    out = {}
    for name, _ in CrtomoCfg._fields_:
        v = getattr(c, name)
        out[name] = v.decode() if isinstance(v, bytes) else v
    return out


BESP_NAMES = ("esp_min", "esp_max", "esp_mit", "esp_med", "esp_std", "grid_min", "grid_max", "grid_minx",
              "grid_maxx", "grid_miny", "grid_maxy")


class Context:
    """One device context holding grid, survey, element matrices, kpot and J."""

    def __init__(self, device: int = 0):
        lib = load()
        h = _p()
        _chk(lib.crt_create(C.byref(h), device), "crt_create")
        self._h = h
        self.lib = lib
        self.grid = None
        self.survey = None
        self.kwnanz = 0

    def close(self):
        if getattr(self, "_h", None):
            self.lib.crt_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---------------------------------------------------------------- set-up
    def set_grid(self, g):
        g.validate()
        rnr = g.rnr if g.rnr.size else np.zeros(1, np.int32)
        _chk(self.lib.crt_set_grid(self._h, g.sanz, g.mb, g.typanz, _ptr(g.typ), _ptr(g.nelanz), _ptr(g.selanz),
                                   g.maxn, _ptr(g.nrel), _ptr(g.x), _ptr(g.y), _ptr(rnr), g.eanz, _ptr(g.enr)),
             "crt_set_grid")
        self.grid = g

    def set_survey(self, s):
        _chk(self.lib.crt_set_survey(self._h, s.nanz, _ptr(s.strnr), _ptr(s.vnr)), "crt_set_survey")
        self.survey = s

    def rwaven(self):
        kwn = np.zeros(30); kwi = np.zeros(30); n = C.c_int(); amin = C.c_double(); amax = C.c_double()
        _chk(self.lib.crt_rwaven(self._h, C.byref(n), _ptr(kwn), _ptr(kwi), C.byref(amin), C.byref(amax)),
             "crt_rwaven")
        self.kwnanz = n.value
        return kwn[: n.value].copy(), kwi[: n.value].copy(), amin.value, amax.value

    def besp_elem(self, gstat_path=None):
        """Grid statistics of besp_elem (src/besp_elem.f90): dict of esp_min, esp_max,
        esp_mit, esp_med, esp_std, grid_min, grid_max, grid_minx, grid_maxx, grid_miny,
        grid_maxy; ``gstat_path`` also writes inv.gstat."""
        st = np.zeros(11)
        p = None if gstat_path is None else gstat_path.encode()
        _chk(self.lib.crt_besp_elem(self._h, _ptr(st), p), "crt_besp_elem")
        return dict(zip(BESP_NAMES, st))

    def set_boundary_values(self, nodes, values):
        """boundary.dat (src/rrandb.f90): 1-based nodes and complex values; zero values
        make them Dirichlet nodes, non-zero values are refused by forward() (903)."""
        nd = np.ascontiguousarray(nodes, np.int32)
        vals = np.ascontiguousarray(values, np.complex128)
        _chk(self.lib.crt_set_boundary_values(self._h, nd.size, _ptr(nd), _ptr(vals)), "crt_set_boundary_values")

    def set_wavenumbers(self, kwn, kwnwi):
        kwn = np.ascontiguousarray(kwn, np.float64); kwi = np.ascontiguousarray(kwnwi, np.float64)
        _chk(self.lib.crt_set_wavenumbers(self._h, kwn.size, _ptr(kwn), _ptr(kwi)), "crt_set_wavenumbers")
        self.kwnanz = kwn.size

    def precal(self, lsytop: bool = False) -> bool:
        lb = C.c_int()
        _chk(self.lib.crt_precal(self._h, int(lsytop), C.byref(lb)), "crt_precal")
        return bool(lb.value)

    def get_precal(self):
        g = self.grid
        K = self.kwnanz
        elbg = np.zeros((K, g.npk, g.elanz)); relbg = np.zeros((g.npk, max(g.relanz, 1)))
        kg = np.zeros((K, g.eanz, max(g.relanz, 1)))
        _chk(self.lib.crt_get_precal(self._h, _ptr(elbg), _ptr(relbg), _ptr(kg)), "crt_get_precal")
        return elbg, relbg, kg

    def set_inversion_kind(self, kind: int):
        """0 complex inversion, 1 DC inversion (ldc: REAL*8 kpot/hpot/sens, bsendc,
        cjggdc), 2 final phase improvement (lfpi: cjggdc on DBLE(sens)).  Discards
        sigma and its results: set_sigma again afterwards."""
        _chk(self.lib.crt_set_inversion_kind(self._h, int(kind)), "crt_set_inversion_kind")
        self.inv_kind = int(kind)

    def _pot_dtype(self):
        return np.float64 if getattr(self, "inv_kind", 0) == 1 else np.complex128

    def set_sigma(self, sigma):
        s = np.ascontiguousarray(sigma, np.complex128)
        assert s.size == self.grid.elanz
        _chk(self.lib.crt_set_sigma(self._h, _ptr(s)), "crt_set_sigma")

    # --------------------------------------------------------------- compute
    def forward(self):
        _chk(self.lib.crt_forward(self._h), "crt_forward")

    def get_kpot(self):
        """kpot as (kwnanz, eanz, sanz) = reference kpot(sanz, eanz, kwnanz)."""
        g = self.grid
        if g is None or g.eanz == 0:
            raise ValueError("get_kpot needs the grid with its electrodes (set_grid)")
        out = np.zeros((self.kwnanz, g.eanz, g.sanz), self._pot_dtype())
        _chk(self.lib.crt_get_kpot(self._h, _ptr(out)), "crt_get_kpot")
        return out

    def voltages(self, mode: int = 1, kfak=None):
        n = self.survey.nanz
        volt = np.zeros(n, np.complex128); sa = np.zeros(n, np.complex128)
        kf = None if kfak is None else np.ascontiguousarray(kfak, np.float64)
        _chk(self.lib.crt_voltages(self._h, mode, _ptr(kf), _ptr(volt), _ptr(sa)), "crt_voltages")
        return volt, sa

    def get_hpot(self):
        g = self.grid
        if g is None or g.eanz == 0:
            raise ValueError("get_hpot needs the grid with its electrodes (set_grid)")
        out = np.zeros((g.eanz, g.sanz), self._pot_dtype())
        _chk(self.lib.crt_get_hpot(self._h, _ptr(out)), "crt_get_hpot")
        return out

    def sens(self, mode: int = 1, strom=None):
        st = None if strom is None else np.ascontiguousarray(strom, np.float64)
        _chk(self.lib.crt_sens(self._h, mode, _ptr(st)), "crt_sens")

    def get_sens(self):
        """J as (manz, nanz) = reference sens(nanz, manz)."""
        out = np.zeros((self.grid.elanz, self.survey.nanz), self._pot_dtype())
        _chk(self.lib.crt_get_sens(self._h, _ptr(out)), "crt_get_sens")
        return out

    # ------------------------------------------------------- normal equations
    def set_sens(self, sens):
        """sens as (manz, nanz) = reference sens(nanz, manz)."""
        s = np.ascontiguousarray(sens, self._pot_dtype())
        manz, nanz = s.shape
        _chk(self.lib.crt_set_sens(self._h, nanz, manz, _ptr(s)), "crt_set_sens")

    def set_regularization(self, ltri, smatm, nx=0, nachbar=None):
        """smatm as (ncol, manz) = reference smatm(manz, ncol); nachbar (maxn+1, manz)."""
        sm = np.ascontiguousarray(smatm, np.float64)
        manz = sm.shape[1]
        nb = None if nachbar is None else np.ascontiguousarray(nachbar, np.int32)
        maxn = 0 if nb is None else nb.shape[0] - 1
        _chk(self.lib.crt_set_regularization(self._h, int(ltri), int(nx), manz, maxn, _ptr(sm), _ptr(nb)),
             "crt_set_regularization")
        self.manz = manz

    def buncert(self, lam, wmatd, wdfak, lres=True, lcov2=True, gauss=False, full=True):
        """Model uncertainty (src/bmcm_mod.f90:49-175) on the device: (errnr, dict of
        ata, ata_reg, cov, res, cov2 as manz x manz arrays when ``full``, and their main
        diagonals under "diag_<name>").  errnr 108: the factorisation rejected ata_reg."""
        manz = self.manz
        w = np.ascontiguousarray(wmatd, np.float64); wd = np.ascontiguousarray(wdfak, np.int32)
        names = ["ata", "ata_reg", "cov", "res", "cov2"]
        want = [True, True, True, bool(lres), bool(lres and lcov2)]
        mats = [np.zeros((manz, manz)) if (full and wq) else None for wq in want]
        diag = np.full((5, manz), np.nan)
        err = self.lib.crt_buncert(self._h, C.c_double(lam), _ptr(w), _ptr(wd), int(bool(lres)), int(bool(lcov2)),
                                   int(bool(gauss)), *[_ptr(m) for m in mats], _ptr(diag))
        if err not in (0, 108):
            raise CRTError(err, "crt_buncert")
        out = {}
        for q, (k, m) in enumerate(zip(names, mats)):
            out[k] = None if m is None else m.T  # column-major storage
            out["diag_" + k] = diag[q] if want[q] else None
        return err, out

    def bsmatm(self, ltri, alfx=1.0, alfz=1.0, nx=0, nz=0):
        """Regularisation built on the device (bsmatmreg for ltri 0 on the regular
        nx x nz grid, bnachbar + bsmatmtri for ltri 1/2) and installed for the CG;
        returns smatm as (ncol, manz) = reference smatm(manz, ncol)."""
        g = self.grid
        ncol = 3 if ltri == 0 else g.maxn + 1
        sm = np.zeros((ncol, g.elanz))
        _chk(self.lib.crt_bsmatm(self._h, int(ltri), int(nx), int(nz), C.c_double(alfx), C.c_double(alfz), _ptr(sm)),
             "crt_bsmatm")
        self.manz = g.elanz
        return sm

    def blam0(self, wmatd, wdfak, alfx=1.0, alfz=1.0, fpi=False):
        """Starting lambda (``src/blam0.f90:120-175``) from the resident Jacobian."""
        w = np.ascontiguousarray(wmatd, np.float64); wd = np.ascontiguousarray(wdfak, np.int32)
        lam = C.c_double()
        _chk(self.lib.crt_blam0(self._h, int(bool(fpi)), _ptr(w), _ptr(wd), C.c_double(alfx), C.c_double(alfz),
                                C.byref(lam)), "crt_blam0")
        return lam.value

    def set_cg_preconditioner(self, kind: int):
        """0: Jacobi scaling (the reference); 1/2: FSAI on the stencil / wider pattern;
        3: the learned preconditioner; 4: its analytic prior alone."""
        _chk(self.lib.crt_set_cg_preconditioner(self._h, int(kind)), "crt_set_cg_preconditioner")

    def pc_train(self, steps: int, lr: float = 1e-3):
        """Train the learned preconditioner on the last CG system; returns the loss
        before the first and after the last Adam step."""
        l0, l1 = C.c_double(), C.c_double()
        _chk(self.lib.crt_pc_train(self._h, int(steps), C.c_double(lr), C.byref(l0), C.byref(l1)), "crt_pc_train")
        return l0.value, l1.value

    def pc_weights(self, set_to=None):
        """Model vector (16 standardisation values, then the parameters); optionally set it first."""
        n = self.lib.crt_pc_size()
        if set_to is not None:
            w = np.ascontiguousarray(set_to, np.float32)
            assert w.size == n
            _chk(self.lib.crt_pc_weights(self._h, None, _ptr(w)), "crt_pc_weights")
        out = np.zeros(n, np.float32)
        _chk(self.lib.crt_pc_weights(self._h, _ptr(out), None), "crt_pc_weights")
        return out

    def pc_grad(self):
        n = self.lib.crt_pc_size() - 16
        g = np.zeros(n, np.float32)
        loss = C.c_double()
        _chk(self.lib.crt_pc_grad(self._h, C.byref(loss), _ptr(g)), "crt_pc_grad")
        return loss.value, g

    def pc_data(self):
        """The training set: raw features (n, 8), targets and mask (n, 5)."""
        n = C.c_int(0)
        _chk(self.lib.crt_pc_data(self._h, C.byref(n), None, None, None), "crt_pc_data")
        f = np.zeros((n.value, 8), np.float32)
        t = np.zeros((n.value, 5), np.float32)
        m = np.zeros((n.value, 5), np.float32)
        _chk(self.lib.crt_pc_data(self._h, C.byref(n), _ptr(f), _ptr(t), _ptr(m)), "crt_pc_data")
        return f, t, m

    def set_fem_switches(self, lsr: bool = False, lana: bool = False, nsink: int = 0):
        """crmod.cfg's forward switches: singularity removal, the analytic solution
        alone, the fictitious sink node (1-based, 0 none)."""
        _chk(self.lib.crt_set_fem_switches(self._h, int(bool(lsr)), int(bool(lana)), int(nsink)),
             "crt_set_fem_switches")

    def bkfak(self):
        """Half-space geometric factors of the survey (bkfak)."""
        k = np.zeros(self.survey.nanz)
        _chk(self.lib.crt_bkfak(self._h, _ptr(k)), "crt_bkfak")
        return k

    def pc_reset(self):
        """Drop the learned preconditioner's training set and model."""
        _chk(self.lib.crt_pc_reset(self._h), "crt_pc_reset")

    def tomodir_crmod(self, cfg_path: str):
        """Run a CRMod tomodir from its crmod.cfg (grid, electrodes, rho, config read by
        the library) and write volt.dat when the measurements switch is set."""
        _chk(self.lib.crt_tomodir_crmod(self._h, cfg_path.encode()), "crt_tomodir_crmod")

    def tomodir_crtomo(self, cfg_path: str):
        """Run a CRTomo tomodir (the inversion) from its crtomo.cfg; returns the last
        iteration index and its data RMS."""
        it = C.c_int()
        rms = C.c_double()
        self.lib.crt_tomodir_crtomo.argtypes = [_p, C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_double)]
        _chk(self.lib.crt_tomodir_crtomo(self._h, cfg_path.encode(), C.byref(it), C.byref(rms)), "crt_tomodir_crtomo")
        return it.value, rms.value

    def set_forward_precision(self, bits: int, tol: float = 1e-14, max_sweeps: int = 12):
        """64: FP64 band factorisation (default); 32: FP32 factorisation on the matrix
        cores with FP64 iterative refinement of every system to ``tol``."""
        _chk(self.lib.crt_set_forward_precision(self._h, int(bits), C.c_double(tol), int(max_sweeps)),
             "crt_set_forward_precision")

    def set_forward_method(self, method: int):
        """-1 (default): the condensed solve where the grid allows it, else the band
        factorisation of every system; 0: always the band factorisation (the
        reference's algorithm); 1: always condensed (903 where it does not apply)."""
        _chk(self.lib.crt_set_forward_method(self._h, int(method)), "crt_set_forward_method")

    def forward_method(self) -> int:
        """Method the last forward used: 0 band per system, 1 condensed."""
        m = C.c_int()
        _chk(self.lib.crt_get_forward_method(self._h, C.byref(m)), "crt_get_forward_method")
        return m.value

    def condense_info(self) -> dict:
        """The grid's condensation plan (after a condensed forward)."""
        v = np.zeros(8, np.int32)
        _chk(self.lib.crt_get_condense_info(self._h, v.ctypes.data_as(C.POINTER(C.c_int))), "crt_get_condense_info")
        return dict(zip(("ok", "interior", "boundary", "band", "tiles", "order", "pieces", "separator"),
                        (int(x) for x in v)))

    def refine_info(self):
        """(sweeps, final correction ratio, batches redone in FP64) of the last forward."""
        sw, fb = C.c_int(), C.c_int()
        ratio = C.c_double()
        _chk(self.lib.crt_get_refine_info(self._h, C.byref(sw), C.byref(ratio), C.byref(fb)), "crt_get_refine_info")
        return sw.value, ratio.value, fb.value

    def set_cg_precision(self, jbits: int):
        """64: FP64 J in the CG iterations (default); 32: complex FP32 copy of J."""
        _chk(self.lib.crt_set_cg_precision(self._h, int(jbits)), "crt_set_cg_precision")

    def cjg(self, lam, eps, ncgmax, wmatd, wdfak, cgfac, bvec, noalloc=False):
        """cjg (src/cg_mod.f90:84) on the resident J and regularisation: (raw dpar,
        cgres, ncg).  ``noalloc`` calls crt_cjg_noalloc, which keeps its buffers in the
        library's static device scratch instead of allocating."""
        manz = np.asarray(cgfac).size
        w = np.ascontiguousarray(wmatd, np.float64); wd = np.ascontiguousarray(wdfak, np.int32)
        cf = np.ascontiguousarray(cgfac, np.float64); bv = np.ascontiguousarray(bvec, np.complex128)
        dpar = np.zeros(manz, np.complex128); cgres = np.zeros(ncgmax + 1); ncg = C.c_int()
        fn = self.lib.crt_cjg_noalloc if noalloc else self.lib.crt_cjg
        _chk(fn(self._h, C.c_double(lam), C.c_double(eps), int(ncgmax), _ptr(w), _ptr(wd), _ptr(cf), _ptr(bv),
                _ptr(dpar), _ptr(cgres), C.byref(ncg)), "crt_cjg_noalloc" if noalloc else "crt_cjg")
        return dpar, cgres, ncg.value

    def update(self, lam, eps, ncgmax, wmatd, wdfak, dat, par, sigmaa=None, step=1.0):
        par = np.array(par, np.complex128)
        manz = par.size
        w = np.ascontiguousarray(wmatd, np.float64); wd = np.ascontiguousarray(wdfak, np.int32)
        d = np.ascontiguousarray(dat, np.complex128)
        sa = None if sigmaa is None else np.ascontiguousarray(sigmaa, np.complex128)
        dpar = np.zeros(manz, np.complex128); cgres = np.zeros(ncgmax + 1); ncg = C.c_int()
        cgfac = np.zeros(manz); bvec = np.zeros(manz, np.complex128)
        _chk(self.lib.crt_update(self._h, C.c_double(lam), C.c_double(eps), int(ncgmax), C.c_double(step), _ptr(w),
                                 _ptr(wd), _ptr(d), _ptr(sa), _ptr(par), _ptr(dpar), _ptr(cgres), C.byref(ncg),
                                 _ptr(cgfac), _ptr(bvec)), "crt_update")
        return dict(par=par, dpar=dpar, cgres=cgres, ncg=ncg.value, cgfac=cgfac, bvec=bvec)

    # ------------------------------------------------- single-system drop-ins
    def chol(self, sanz, mb, a):
        """In-place band Cholesky of a (reference band layout); returns errnr."""
        a2 = np.ascontiguousarray(a, np.complex128)
        err = self.lib.crt_chol(self._h, int(sanz), int(mb), _ptr(a2))
        if a2 is not a:
            a[...] = a2
        return err

    def vre(self, sanz, mb, a, b):
        pot = np.zeros(sanz, np.complex128)
        _chk(self.lib.crt_vre(self._h, int(sanz), int(mb), _ptr(np.ascontiguousarray(a, np.complex128)),
                              _ptr(np.ascontiguousarray(b, np.complex128)), _ptr(pot)), "crt_vre")
        return pot

    def choldc(self, sanz, mb, a):
        """In-place real band Cholesky (src/choldc.f90); returns errnr."""
        a2 = np.ascontiguousarray(a, np.float64)
        err = self.lib.crt_choldc(self._h, int(sanz), int(mb), _ptr(a2))
        if a2 is not a:
            a[...] = a2
        return err

    def vredc(self, sanz, mb, a, b):
        pot = np.zeros(sanz, np.complex128)
        _chk(self.lib.crt_vredc(self._h, int(sanz), int(mb), _ptr(np.ascontiguousarray(a, np.float64)),
                                _ptr(np.ascontiguousarray(b, np.float64)), _ptr(pot)), "crt_vredc")
        return pot

    def cholz(self, a_colmajor):
        """a (n, n) Fortran-ordered complex; returns (errnr, a_out, p)."""
        n = a_colmajor.shape[0]
        a = np.array(a_colmajor, dtype=np.complex128, order="F")
        p = np.zeros(n, np.complex128)
        err = self.lib.crt_cholz(self._h, n, a.ctypes.data_as(_p), _ptr(p))
        return err, a, p

    def times(self):
        ms = np.zeros(12)
        self.lib.crt_get_times(self._h, _ptr(ms), 12)
        return dict(zip(("diag", "factor", "solve", "volt", "sens", "precal", "cg", "cg_matvec", "refine", "schur",
                         "product", "prefix"), ms))


class Batch:
    """Forward of many contexts as one graph launch (``crt_batch_*``).

    Every context needs its grid, survey, wavenumbers, precal and sigma; after
    ``forward()`` each holds kpot as after ``Context.forward()``.  ``autotune``
    times the factor kernel variants per band width on the first chunk and keeps
    the fastest.  ``mixed`` runs the mixed-precision variant: FP32 factorisation and
    FP64 refinement inside the graph, models that do not converge redone in FP64."""

    def __init__(self, contexts, workspace_bytes: int = 0, autotune: bool = False, mixed: bool = False):
        lib = load()
        lib.crt_batch_create.argtypes = [C.POINTER(_p), C.c_int, C.c_size_t, C.c_int, C.POINTER(_p)]
        arr = (_p * len(contexts))(*[c._h.value for c in contexts])
        h = _p()
        flags = (1 if autotune else 0) | (2 if mixed else 0)
        _chk(lib.crt_batch_create(arr, len(contexts), workspace_bytes, flags, C.byref(h)), "crt_batch_create")
        self._h = h
        self.lib = lib
        self.contexts = list(contexts)

    def forward(self):
        _chk(self.lib.crt_batch_forward(self._h), "crt_batch_forward")

    def info(self):
        """(device ms of the last graph launch, systems per launch, variant per band
        width in tiles, tuning times [D, variant] in ms)."""
        ms = C.c_double()
        n = C.c_longlong()
        var = np.zeros(17, np.int32)
        tune = np.zeros(68, np.float32)
        _chk(self.lib.crt_batch_info(self._h, C.byref(ms), C.byref(n), _ptr(var), _ptr(tune)), "crt_batch_info")
        return ms.value, n.value, var, tune.reshape(17, 4)

    def refine_info(self):
        """Mixed variant, last launch: (most sweeps of a chunk, largest accepted
        correction ratio, models redone in FP64)."""
        sw = C.c_int()
        r = C.c_double()
        fb = C.c_int()
        _chk(self.lib.crt_batch_refine_info(self._h, C.byref(sw), C.byref(r), C.byref(fb)), "crt_batch_refine_info")
        return sw.value, r.value, fb.value

    def close(self):
        if getattr(self, "_h", None):
            self.lib.crt_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
