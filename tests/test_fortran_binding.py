"""The Fortran binding of the library (crtomo_stable_amd/fortran): the interface module
crt_hip_mod and the shims that keep the reference's routine names and arguments
(chol, vre, choldc, vredc, cholz; bvolt, bvolti, bsens, bsensi, bsendc, bnachbar,
blam0; crt_hip_setup, crt_hip_forward, crt_hip_cjg, crt_hip_bsmatm).

CPU: every file compiles with amdflang against the reference's own modules
(oracle/_ref/mod, built from /root/reference/src), and a shared library of the shims
links against libcrtomo_hip.so with --no-undefined; every C entry a shim calls is one
that include/crtomo_hip.h declares and the library exports; the drop-in library that
oracle/Makefile links (reference routines with chol/vre/choldc/vredc replaced) defines
those names itself and leaves crt_chol / crt_vre to libcrtomo_hip.so.
"""
import os
import re
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "crtomo_stable_amd", "fortran")
REF = os.path.join(ROOT, "oracle", "_ref")
HIPLIB = os.path.join(ROOT, "crtomo_stable_amd", "_lib", "libcrtomo_hip.so")
FC = "/opt/rocm/bin/amdflang"

need_ref = pytest.mark.skipif(not (os.path.exists(FC) and os.path.isdir(os.path.join(REF, "mod"))
                                   and os.path.exists(HIPLIB)),
                              reason="amdflang, the reference modules (make -C oracle ref) or the library missing")


def _nm(path, flags="-D"):
    out = subprocess.run(["nm", flags, path], capture_output=True, text=True, check=True).stdout
    defined, undefined = set(), set()
    for line in out.splitlines():
        parts = line.split()
        if len(parts) == 2 and parts[0] == "U":
            undefined.add(parts[1])
        elif len(parts) == 3 and parts[1] in "TDBRW":
            defined.add(parts[2])
    return defined, undefined


def _header_functions():
    txt = open(os.path.join(ROOT, "include", "crtomo_hip.h")).read()
    return set(re.findall(r"\b(crt_[a-z0-9_]+)\s*\(", txt))


@need_ref
def test_binding_compiles_and_links():
    tmp = tempfile.mkdtemp(prefix="crt_fbind_")
    try:
        objs = []
        for f in ("crt_hip_mod.f90", "crt_shims_forward.f90", "crt_shims_inversion.f90"):
            o = os.path.join(tmp, f[:-4] + ".o")
            subprocess.run([FC, "-O2", "-fopenmp", "-fPIC", "-I", os.path.join(REF, "mod"), "-J", tmp, "-c",
                            os.path.join(FDIR, f), "-o", o], check=True, capture_output=True, text=True)
            objs.append(o)
        so = os.path.join(tmp, "libshims.so")
        # the shims alone: everything they need from the reference is module data, which
        # the reference's module objects provide
        mods = [os.path.join(REF, "obj", m + ".o") for m in
                ("tic_toc", "errmod", "elemmod", "electrmod", "wavenmod", "sigmamod", "datmod", "alloci", "femmod",
                 "invmod", "konvmod", "modelmod", "ompmod", "pathmod", "randbmod", "cjgmod", "get_unit")]
        r = subprocess.run([FC, "-shared", "-fopenmp", "-o", so] + objs + mods +
                           ["-L" + os.path.dirname(HIPLIB), "-lcrtomo_hip", "-Wl,-rpath,/opt/rocm/llvm/lib",
                            "-Wl,--no-undefined"], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        defined, undefined = _nm(so)
        shims = {"chol_", "vre_", "choldc_", "vredc_", "cholz_", "bvolt_", "bvolti_", "bsens_", "bsensi_", "bsendc_", "bkfak_",
                 "bnachbar_", "blam0_", "crt_hip_setup_", "crt_hip_forward_", "crt_hip_cjg_", "crt_hip_bsmatm_"}
        assert shims <= defined, shims - defined
        calls = {u for u in undefined if u.startswith("crt_")}
        lib_defined, _ = _nm(HIPLIB)
        assert calls and calls <= _header_functions() and calls <= lib_defined, calls
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


@need_ref
def test_dropin_library_uses_the_shims():
    so = os.path.join(REF, "libcrtomo_dropin.so")
    if not os.path.exists(so):
        pytest.skip("make -C oracle dropin not run")
    defined, undefined = _nm(so)
    assert {"chol_", "vre_", "choldc_", "vredc_", "ref_forward", "ref_forward_dc"} <= defined
    assert {"crt_chol", "crt_vre", "crt_choldc", "crt_vredc", "crt_create"} <= undefined
