"""The reference's own forward loop driving the GPU through the Fortran shims.

oracle/_ref/libcrtomo_dropin.so is the reference's routines (compiled unmodified from
its sources) with chol, vre, choldc and vredc replaced by the library's Fortran shims
(crtomo_stable_amd/fortran/crt_shims_forward.f90).  Its harness runs the forward loop
of src/fem.f90:328-397 as the reference writes it -- an OpenMP loop over wavenumbers
with thread-private band, right-hand side and potential, kompab and scalab of the
reference on the host, then chol and vre -- here on 4 OpenMP threads, so the shims are
called from several threads at once on one library context.  kpot, the voltages
(the reference's rtrafo + bvolti on that kpot) and the DC loop (kompadc, scaldc, then
the choldc / vredc shims) are compared with the oracle at 1e-10.
"""
import ctypes as C
import os

import numpy as np
import pytest

from crtomo_stable_amd import configs

pytestmark = pytest.mark.gpu


def _rel_rows(a, b):
    a2 = a.reshape(-1, a.shape[-1]); b2 = b.reshape(-1, b.shape[-1])
    return float((np.abs(a2 - b2).max(axis=1) / np.abs(b2).max(axis=1)).max())


@pytest.mark.parametrize("name", ["c0", "small_tri"])
def test_reference_forward_loop_through_shims(oracle_lib, name):
    O = oracle_lib
    if not os.path.exists(O.DROPIN_PATH):
        pytest.skip("oracle/_ref/libcrtomo_dropin.so not built (make -C oracle dropin)")
    w = configs.c0() if name == "c0" else configs.c1(scale=0.2, nelec=12, nmeas=60, seed=7)
    g, s = w.grid, w.survey
    lib = C.CDLL(O.DROPIN_PATH)
    R = O.RefSession(g, s, lib=lib)
    kwn, kwi, _, _ = R.rwaven()
    R.precal()
    print(f"[fortran dropin] {name} sanz={g.sanz} mb={g.mb} systems={kwn.size * g.eanz}, 4 OpenMP threads",
          flush=True)
    kpot = R.forward(w.sigma, nthreads=4)
    hpot, volt, sa = R.voltages(complex_inv=True)
    okwn, okwi, _, _ = O.rwaven(g, s)
    elbg, relbg, kg, lb = O.precal(g, okwn)
    kref = O.forward(g, w.sigma, elbg, relbg, kg, lb)
    vref, saref = O.bvolti(g, s, O.rtrafo(kref, okwi))
    assert _rel_rows(kpot, kref) <= 1e-10
    assert float(np.max(np.abs(volt - vref) / np.abs(vref))) <= 1e-10
    # the DC loop: kompadc, scaldc, then the choldc / vredc shims
    kdc = R.forward_dc(w.sigma)
    kdc_ref = O.forward_dc(g, w.sigma, elbg, relbg, kg, lb)
    assert _rel_rows(kdc, kdc_ref) <= 1e-10
