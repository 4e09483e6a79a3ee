#!/usr/bin/env python3
"""Throughput of the CRTomo forward + Jacobian path on one MI355X.

One step = the complex forward solve for every (electrode, wavenumber) system,
the inverse Fourier back-transform with the voltages (bvolti) and the full
Jacobian (bsensi), on the seeded C1 workload: ~20k-triangle Delaunay mesh over
100 x 30 m with 2 m sinusoidal topography, 64 electrodes at 1.5 m, 2000
dipole-dipole + Wenner quadrupoles, complex GRF model.  The forward runs the
library's default method: static condensation onto the mixed-boundary nodes
(interior band factorised once per wavenumber, dense boundary Schur complement
per system); the band factorisation of every system (the reference's
algorithm) is timed beside it (band_per_system).  Grid, element matrices and
conductivities are resident in HBM before the timed region.

Prints one JSON line (see the "roofline" and "cpu_baseline" objects).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6   # MI355X dense FP64 (vector and matrix)
FP32_PEAK_TFLOPS = 157.3  # MI355X dense FP32 matrix
HBM_PEAK_GBS = 8000.0


KERNEL_SOURCES = {  # the source file each benched kernel is compiled from
    "crt_cond_trsm_kernel": "crt_condense.hip", "crt_cond_dense_kernel": "crt_condense.hip",
    "crt_cond_gemm_kernel": "crt_condense.hip",
    "crt_factor16la_kernel": "crt_band16.hip", "crt_factor16st_kernel": "crt_band16.hip",
    "crt_factor16_kernel": "crt_band16.hip", "crt_solve16_kernel": "crt_band16.hip",
    "crt_factor32st_kernel": "crt_band32.hip", "crt_factor_kernel": "crt_band.hip",
    "crt_sens_gm_kernel": "crt_volt_sens.hip", "crt_sens_gp_kernel": "crt_volt_sens.hip",
    "crt_cg_jq_kernel": "crt_cg.hip", "crt_cg_jhw_kernel": "crt_cg.hip",
}


def source_hash(kernel: str) -> str:
    """First 12 hex digits of the SHA-1 of the source file of ``kernel``."""
    import hashlib
    name = KERNEL_SOURCES[kernel.split("<")[0]]
    with open(os.path.join(ROOT, "crtomo_stable_amd", "csrc", name), "rb") as f:
        return hashlib.sha1(f.read()).hexdigest()[:12]


def measured_traffic(kernel: str, workload: str, method: str):
    """HBM bytes per launch of `kernel` on `workload` in the forward `method` ("condensed"
    or "band", since the interior factorisation of the condensed forward and the band
    factorisation of every system share a kernel) from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (profiles/<workload>_pmc_traffic.md, a line
    `traffic <kernel> <method> <bytes> <source sha1[:12]>`).  None when no measurement of
    this kernel is committed, or when its source changed since it was measured."""
    import re
    path = os.path.join(ROOT, "profiles", f"{workload}_pmc_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        m = re.match(r"^traffic\s+(\S+(?:,\s\S+)*)\s+(condensed|band)\s+(\d+)\s+([0-9a-f]{12})\s*$", line.strip())
        if m and m.group(1) == kernel and m.group(2) == method:
            return int(m.group(3)) if m.group(4) == source_hash(kernel) else None
    return None


def device_sync():
    lib = ctypes.CDLL("libamdhip64.so")
    lib.hipDeviceSynchronize()


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_cores() -> int:
    """Cores this process may run on (its CPU affinity set)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def _ref_sample(w, threads: int, sens_meas: int, fwd_electrodes: int):
    """One timing of the reference build (oracle/_ref) on `threads` OpenMP threads:
    the forward of min(K, threads) wavenumbers x `fwd_electrodes` electrodes and bsensi
    of `sens_meas` measurements, each scaled to the full step."""
    import numpy as np
    from oracle import oracle as O

    g, s = w.grid, w.survey
    R = O.RefSession(g, s)
    kwn, _, _, _ = R.rwaven()
    R.precal()
    r = O.ref()
    r.ref_forward_sample.restype = ctypes.c_double
    r.ref_sens_sample.restype = ctypes.c_double
    K, E = kwn.size, g.eanz
    nk = min(K, threads)
    sig = np.ascontiguousarray(w.sigma, np.complex128)
    t_f = r.ref_forward_sample(O.ptr(sig), threads, nk, fwd_electrodes)
    # the reference spreads the forward over wavenumbers only (src/fem.f90:257-277)
    waves = -(-K // nk)
    fwd_full = waves * E * t_f / fwd_electrodes
    t_s = r.ref_sens_sample(threads, sens_meas)
    sens_full = t_s * s.nanz / sens_meas
    return dict(threads=threads, forward_s=fwd_full, sens_s=sens_full, sample_forward_s=t_f, sample_sens_s=t_s,
                nk=nk, fwd_electrodes=fwd_electrodes, sens_meas=sens_meas)


def cpu_baseline(w, threads: int, sens_meas: int, fwd_electrodes: int):
    """Reference OpenMP Fortran build (oracle/_ref, amdflang -O2 -fopenmp) on a bounded
    sample of the same workload, on every core this process may use and on one
    thread; the C restatement (one thread, forward only) when the reference build is
    absent."""
    from oracle import oracle as O

    g, s = w.grid, w.survey
    if O.ref_available():
        a = _ref_sample(w, threads, sens_meas, fwd_electrodes)
        one = _ref_sample(w, 1, max(4, sens_meas // 12), fwd_electrodes)
        sample = (f"reference OpenMP Fortran (amdflang -O2 -fopenmp) on {threads} threads: forward of {a['nk']} "
                  f"wavenumbers x {fwd_electrodes} electrodes ({a['sample_forward_s']:.2f} s), bsensi of "
                  f"{sens_meas} of {s.nanz} measurements ({a['sample_sens_s']:.2f} s), scaled to the full step; "
                  f"one thread: {one['nk']} wavenumber x {fwd_electrodes} electrodes "
                  f"({one['sample_forward_s']:.2f} s) and {one['sens_meas']} measurements "
                  f"({one['sample_sens_s']:.2f} s)")
        return dict(value=1.0 / (a["forward_s"] + a["sens_s"]), unit="solves/s", cores=threads, kind="reference",
                    sample=sample, forward_s=a["forward_s"], sens_s=a["sens_s"], cpu_model=cpu_model(),
                    one_thread=dict(value=1.0 / (one["forward_s"] + one["sens_s"]), forward_s=one["forward_s"],
                                    sens_s=one["sens_s"]))
    kwn, kwi, _, _ = O.rwaven(g, s)
    elbg, relbg, kg, lb = O.precal(g, kwn)
    K, E = kwn.size, g.eanz
    systems = [(l, 1) for l in range(1, fwd_electrodes + 1)]
    t0 = time.perf_counter()
    O.forward(g, w.sigma, elbg, relbg, kg, lb, systems=systems)
    t_f = time.perf_counter() - t0
    fwd_full = t_f * K * E / fwd_electrodes
    sample = f"C restatement, 1 thread: forward of {fwd_electrodes} systems scaled to {K * E}"
    return dict(value=1.0 / fwd_full, unit="solves/s", cores=1, kind="port", sample=sample,
                forward_s=fwd_full, sens_s=None, cpu_model=cpu_model())


def other_rooflines(g, s, K, E, nsys, phase, bits, workload):
    """Algorithmic HBM bytes over the measured phase time (HIP events on the library's
    stream) for the HBM-bound kernels of one step: substitution, assembly (diagonal
    and value passes) and the Jacobian.  Bytes per unit as in the design notes:
      substitution: the factor's band read once, 16 (mb+1) sanz per right-hand side (8
        with FP32 panels), and the kpot write 16 sanz;
      assembly: elbg slice 8 elanz npk, sigma 16 elanz, the CSR values 16 nnz (nnz: the
        lower-triangle entries incl. the diagonal), per system;
      Jacobian: J write 16 nanz elanz, kpot 16 sanz E K, elbg 8 elanz npk K."""
    nrhs = E * K
    pb = 8 if bits == 32 else 16
    solve_b = nrhs * (pb * (g.mb + 1) * g.sanz + 16 * g.sanz)
    import numpy as np
    el = np.asarray(g.nrel[:g.elanz], np.int64)
    pairs = [np.maximum(el[:, a], el[:, c]) * (g.sanz + 1) + np.minimum(el[:, a], el[:, c])
             for a in range(el.shape[1]) for c in range(a + 1)]
    pk = np.concatenate(pairs)
    nnz = np.unique(pk[(pk % (g.sanz + 1)) > 0]).size
    asm_b = nsys * (8 * g.elanz * g.npk + 16 * g.elanz + 16 * nnz)
    sens_b = 16.0 * s.nanz * g.elanz + 16.0 * g.sanz * E * K + 8.0 * g.elanz * g.npk * K
    out = {}
    nt = (E + 15) // 16  # the Jacobian kernel's electrode tiles
    for name, key, b, kern in (("substitution (crt_solve16_kernel, first solve)", "solve", solve_b, None),
                               ("assembly (crt_diag_kernel + crt_assemble_kernel)", "diag", asm_b, None),
                               ("Jacobian (crt_sens_gp_kernel)", "sens", sens_b, f"crt_sens_gp_kernel<{nt}>")):
        ms = phase.get(key, 0.0)
        if ms > 0:
            gbs = b / (ms * 1e-3) / 1e9
            out[name] = dict(bound="hbm", achieved=round(gbs, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                             frac=round(gbs / HBM_PEAK_GBS, 4), bytes=b, ms=round(ms, 4))
            if kern:
                out[name].update(kernel=kern, traffic=measured_traffic(kern, workload, "condensed"))
    return out


def gn_step(ctx, w, ltri: int):
    """One complex Gauss-Newton step of the inversion loop (src/inv.f90: forward on
    the current model, bvolti, bsensi, then update -> cjg) on the device, with the
    regularisation (bsmatm) and the starting lambda (blam0) also built there.
    Synthetic data: voltages of the workload's model with 3 % + 1e-4 V Gaussian
    noise; start model homogeneous 100 ohm m, -5 mrad; CG to eps = 1e-12 on the
    squared residual (relative residual 1e-6), ncgmax = manz / 10
    (src/rall.f90:735-738).  Returns phase times in ms and the CG figures."""
    import numpy as np
    from crtomo_stable_amd import lib
    g = w.grid
    ctx.set_sigma(w.sigma)
    ctx.forward()
    v_true, _ = ctx.voltages(1)
    rng = np.random.default_rng(11)
    err = 0.03 * np.abs(v_true) + 1e-4
    v_obs = v_true + err * rng.standard_normal(v_true.size) * v_true / np.abs(v_true)
    dat = -np.log(v_obs)
    wmatd = 1.0 / (err / np.abs(v_true)) ** 2
    wdfak = np.ones(v_true.size, np.int32)
    sig0 = np.full(g.elanz, (1.0 / 100.0) * np.exp(-1j * 5e-3))
    par = np.log(sig0)
    manz = g.elanz
    ncgmax = manz // 10
    device_sync()
    t = {}
    t0 = time.perf_counter()
    ctx.set_sigma(sig0)
    ctx.forward()
    v64, _ = ctx.voltages(1)
    device_sync()
    t["forward"] = time.perf_counter() - t0
    tmf = ctx.times()
    t1 = time.perf_counter()
    ctx.sens(1)
    device_sync()
    t["sens"] = time.perf_counter() - t1
    t1 = time.perf_counter()
    ctx.bsmatm(ltri, 1.0, 1.0, nx=g.nx, nz=getattr(g, "nz", 0))
    lam = ctx.blam0(wmatd, wdfak, 1.0, 1.0)
    device_sync()
    t["bsmatm_blam0"] = time.perf_counter() - t1
    t1 = time.perf_counter()
    r = ctx.update(lam, 1e-12, ncgmax, wmatd, wdfak, dat, par)
    device_sync()
    t["update"] = time.perf_counter() - t1
    t["step"] = time.perf_counter() - t0
    tm = ctx.times()
    s64 = ctx.get_sens()  # host copy for the mixed-precision comparison, outside the step
    # the same update in the other CG forms: complex FP32 copy of J in the iterations,
    # the FSAI preconditioner (pattern 2, and the stencil pattern 1) on top of the Jacobi
    # scaling, and the learned preconditioner's analytic prior alone (kind 4)
    d64 = r["dpar"]
    forms = {}
    for name, jbits, prec in (("fp32_jacobian", 32, 0), ("fsai", 64, 2), ("fp32_jacobian_fsai", 32, 2),
                              ("fp32_jacobian_fsai_stencil", 32, 1), ("fp32_jacobian_learned_prior", 32, 4)):
        ctx.set_cg_precision(jbits)
        ctx.set_cg_preconditioner(prec)
        t1 = time.perf_counter()
        rv = ctx.update(lam, 1e-12, ncgmax, wmatd, wdfak, dat, par)
        device_sync()
        tv = time.perf_counter() - t1
        tmv = ctx.times()
        forms[name] = {"update_ms": round(1e3 * tv, 3), "ncg": rv["ncg"],
                       "cg_ms_per_iteration": round(tmv["cg_matvec"], 4),
                       "dpar_rel_diff_vs_fp64_jacobi": float(np.max(np.abs(rv["dpar"] - d64)) / np.max(np.abs(d64))),
                       "step_ms": round(1e3 * (t["step"] - t["update"] + tv), 3)}
    # learned preconditioner: trained on the system of a neighbouring lambda trial
    # (2 lam0), then used at lam0; the training time is reported apart from the step
    ctx.set_cg_precision(64)
    ctx.set_cg_preconditioner(0)
    ctx.update(2.0 * lam, 1e-12, 1, wmatd, wdfak, dat, par)
    t1 = time.perf_counter()
    l0, l1 = ctx.pc_train(300, 3e-3)
    device_sync()
    t_train = time.perf_counter() - t1
    for name, jbits in (("learned", 64), ("fp32_jacobian_learned", 32)):
        ctx.set_cg_precision(jbits)
        ctx.set_cg_preconditioner(3)
        t1 = time.perf_counter()
        rv = ctx.update(lam, 1e-12, ncgmax, wmatd, wdfak, dat, par)
        device_sync()
        tv = time.perf_counter() - t1
        tmv = ctx.times()
        forms[name] = {"update_ms": round(1e3 * tv, 3), "ncg": rv["ncg"],
                       "cg_ms_per_iteration": round(tmv["cg_matvec"], 4),
                       "dpar_rel_diff_vs_fp64_jacobi": float(np.max(np.abs(rv["dpar"] - d64)) / np.max(np.abs(d64))),
                       "step_ms": round(1e3 * (t["step"] - t["update"] + tv), 3),
                       "pc_train_ms": round(1e3 * t_train, 1), "pc_loss": [l0, l1]}
    ctx.set_cg_precision(64)
    ctx.set_cg_preconditioner(0)
    # the same step with the mixed-precision forward (FP32 factorisation + FP64
    # refinement), FP64 Jacobi CG; errors against the FP64 step per voltage,
    # sensitivity row and model update
    mixed = None
    try:
        ctx.set_forward_precision(32)
    except lib.CRTError as e:
        mixed = {"unsupported": str(e)}
    if mixed is None:
        try:
            t1 = time.perf_counter()
            ctx.set_sigma(sig0)
            ctx.forward()
            v32, _ = ctx.voltages(1)
            device_sync()
            tf = time.perf_counter() - t1
            sweeps, ratio, fallbacks = ctx.refine_info()
            tm32 = ctx.times()
            t1 = time.perf_counter()
            ctx.sens(1)
            device_sync()
            tsn = time.perf_counter() - t1
            s32 = ctx.get_sens()
        finally:
            ctx.set_forward_precision(64)
        t1 = time.perf_counter()
        r32 = ctx.update(lam, 1e-12, ncgmax, wmatd, wdfak, dat, par)
        device_sync()
        tu = time.perf_counter() - t1
        srow = np.abs(s32 - s64).max(axis=1) / np.abs(s64).max(axis=1)
        mixed = {"forward_ms": round(1e3 * tf, 3), "factor_ms": round(tm32["factor"], 3),
                 "first_solve_ms": round(tm32["solve"], 3), "refine_ms": round(tm32["refine"], 3),
                 "sweeps": sweeps, "final_correction_ratio": ratio, "fp64_fallback_batches": fallbacks,
                 "sens_ms": round(1e3 * tsn, 3), "update_ms": round(1e3 * tu, 3), "ncg": r32["ncg"],
                 "volt_rel_err_vs_fp64": float(np.max(np.abs(v32 - v64) / np.abs(v64))),
                 "sens_row_rel_err_vs_fp64": float(srow.max()),
                 "dpar_rel_diff_vs_fp64": float(np.max(np.abs(r32["dpar"] - d64)) / np.max(np.abs(d64))),
                 "step_ms": round(1e3 * (tf + tsn + t["bsmatm_blam0"] + tu), 3)}
    # the DC inversion's step (ldc: kompadc..vredc forward on DBLE(sigma), bvolti's DC
    # branch, bsendc, update + cjggdc on the real Jacobian) and the final phase
    # improvement's update (lfpi: cjggdc on DBLE(sens) from DIMAG(bvec)), same data
    dc = {}
    try:
        ctx.set_inversion_kind(1)
        td = {}
        device_sync()
        t0 = time.perf_counter()
        ctx.set_sigma(sig0)
        ctx.forward()
        ctx.voltages(1)
        device_sync()
        td["forward"] = time.perf_counter() - t0
        t1 = time.perf_counter()
        ctx.sens(1)
        device_sync()
        td["sens"] = time.perf_counter() - t1
        t1 = time.perf_counter()
        ctx.bsmatm(ltri, 1.0, 1.0, nx=g.nx, nz=getattr(g, "nz", 0))
        lam_dc = ctx.blam0(wmatd, wdfak, 1.0, 1.0)
        device_sync()
        td["bsmatm_blam0"] = time.perf_counter() - t1
        t1 = time.perf_counter()
        rd = ctx.update(lam_dc, 1e-12, ncgmax, wmatd, wdfak, dat, par)
        device_sync()
        td["update"] = time.perf_counter() - t1
        td["step"] = time.perf_counter() - t0
        itd = ctx.times()["cg_matvec"]
        dc["dc_step"] = {"phase_ms": {k: round(1e3 * v, 3) for k, v in td.items()}, "lambda0": lam_dc,
                         "ncg": rd["ncg"], "cg_ms_per_iteration": round(itd, 4),
                         "cg_matvec_GBps": round(2.0 * 8.0 * v_true.size * manz / (itd * 1e-3) / 1e9, 1)
                         if itd > 0 else None}
        # the same DC update with the FP32 copy of the real J and FSAI on the wider pattern
        ctx.set_cg_precision(32)
        ctx.set_cg_preconditioner(2)
        try:
            t1 = time.perf_counter()
            rdf = ctx.update(lam_dc, 1e-12, ncgmax, wmatd, wdfak, dat, par)
            device_sync()
            tuf = time.perf_counter() - t1
            itf = ctx.times()["cg_matvec"]
        finally:
            ctx.set_cg_preconditioner(0)
            ctx.set_cg_precision(64)
        dc["dc_step_fp32j_fsai"] = {
            "update_ms": round(1e3 * tuf, 3), "ncg": rdf["ncg"], "cg_ms_per_iteration": round(itf, 4),
            "step_ms": round(1e3 * (td["forward"] + td["sens"] + td["bsmatm_blam0"] + tuf), 3),
            "dpar_rel_diff_vs_jacobi": float(np.max(np.abs(rdf["dpar"] - rd["dpar"])) / np.max(np.abs(rd["dpar"])))}
        ctx.set_inversion_kind(2)
        ctx.set_sigma(sig0)
        ctx.forward()
        ctx.voltages(1)
        ctx.sens(1)
        device_sync()
        t1 = time.perf_counter()
        rf = ctx.update(lam, 1e-12, ncgmax, wmatd, wdfak, dat, par)
        device_sync()
        itf = ctx.times()["cg_matvec"]
        dc["fpi_update"] = {"update_ms": round(1e3 * (time.perf_counter() - t1), 3), "ncg": rf["ncg"],
                            "cg_ms_per_iteration": round(itf, 4)}
    finally:
        ctx.set_inversion_kind(0)
    nanz = v_true.size
    j_bytes = 16.0 * nanz * manz
    it_ms = tm["cg_matvec"]
    return {"phase_ms": {k: round(1e3 * v, 3) for k, v in t.items()}, "lambda0": lam, "ncg": r["ncg"],
            "ncgmax": ncgmax, "cg_rel_residual": float(np.sqrt(r["cgres"][r["ncg"]])) if r["ncg"] > 0 else None,
            "cg_ms_per_iteration": round(it_ms, 4),
            "cg_matvec_GBps": round(2.0 * j_bytes / (it_ms * 1e-3) / 1e9, 1) if it_ms > 0 else None,
            "J_bytes": j_bytes, "cg_forms": forms, "forward_device_ms": {k: round(tmf[k], 3) for k in
                                                                       ("diag", "factor", "solve", "volt")},
            "mixed_precision_forward": mixed, "dc_and_fpi": dc}


def gn_inversion(nit: int, train_models: int, pc_steps: int, eps: float = 1e-4):
    """The metric's second half on the 400 x 100 quad workload (96 electrodes, 4500
    measurements, 3 % + 1e-4 V noise, start model homogeneous 100 ohm-m / -5 mrad),
    with each form of the Gauss-Newton step:
      fp64_jacobi    the plain FP64 GPU path, the reference's algorithms on the device:
                     band factorisation of every (electrode, wavenumber) system, FP64 J,
                     Jacobi-scaled CG;
      fp64_jacobi_condensed  the same with the condensed forward;
      fp32j_fsai     condensed forward, complex FP32 copy of J in the iterations, FSAI
                     preconditioner;
      fp32j_learned  condensed forward, FP32 J copy, learned preconditioner trained
                     beforehand on the CG systems of `train_models` inversions of other
                     models (seeded Gaussian random fields, 10..1000 ohm-m, correlation
                     2..20 m) on the same grid and survey, none of them the scored model.
    gn_step (equal update): the first Gauss-Newton step from the start model, each form
    from the same model and lambda (blam0), the CG to eps 1e-12 on the squared residual
    (relative residual 1e-6, where every form converges): Jacobian + update + the
    forward of the new model with its misfit (src/inv.f90:415-486, src/update.f90,
    src/dmisft.f90).  The updates agree within the stop's accuracy and so do the
    misfits after the step; the step times are compared.
    inversion (default eps): `nit` steps of the loop of crtomo_stable_amd.inversion at
    the reference's default eps (1e-4, src/rall.f90:210; lam0 halved per step, a step
    that does not lower the misfit retried at half length) with each form; at that
    stop the forms' CG updates differ, so their trajectories and final misfits do too
    (reported as they come)."""
    import numpy as np
    from crtomo_stable_amd import configs, inversion as INV, lib

    w = configs.c3()
    g = w.grid
    manz = g.elanz
    ctx = lib.Context(0)
    ctx.set_grid(g)
    ctx.set_survey(w.survey)
    ctx.rwaven()
    ctx.precal()
    dat, wmatd = INV.synthetic_data(ctx, w.sigma)
    wdfak = np.ones(dat.size, np.int32)
    par0 = np.full(manz, np.log((1.0 / 100.0) * np.exp(-1j * 5e-3)))
    out = {"workload": "c3: 400x100 quads (40000 cells), 96 electrodes, 4500 dipole-dipole + Wenner, 3 % + 1e-4 V "
                       "noise; complex Gauss-Newton from 100 ohm-m / -5 mrad, lambda lam0 (blam0)",
           "gn_step": {"cg_eps": 1e-12, "forms": {}}, "inversion_default_eps": {"cg_eps": eps, "steps": nit,
                                                                                 "forms": {}}}
    forms = (("fp64_jacobi", 64, 0, 0), ("fp64_jacobi_condensed", 64, 0, 1), ("fp32j_fsai", 32, 2, 1),
             ("fp32j_learned", 32, 3, 1))

    def train():
        # training inversions of other models (their CG in the FP32-J FSAI form), each
        # system added to the training set and trained on right after its update
        losses = []
        ctx.set_forward_method(1)
        ctx.set_cg_precision(32)
        ctx.set_cg_preconditioner(2)
        t0 = time.perf_counter()

        def hook(c, it):
            losses.append(c.pc_train(pc_steps, 3e-3))
        for m in range(train_models):
            rng = np.random.default_rng(100 + m)
            sig_m = configs.grf_sigma(g, float(rng.uniform(2.0, 20.0)), seed=1000 + m)
            dat_m, w_m = INV.synthetic_data(ctx, sig_m, seed=200 + m)
            INV.gauss_newton(ctx, g, dat_m, w_m, par0, nit, ltri=0, eps=eps, lam_factor=0.5, on_update=hook)
        device_sync()
        return {"models": train_models, "adam_steps_per_system": pc_steps,
                "seconds": round(time.perf_counter() - t0, 1), "systems": len(losses),
                "loss_first_last_of_each_call": losses[:1] + losses[-1:]}

    ctx.bsmatm(0, 1.0, 1.0, nx=g.nx, nz=g.nz)
    lam0, d_ref, training = None, None, None
    for name, jbits, prec, method in forms:
        if prec == 3 and training is None:
            training = train()
            ctx.bsmatm(0, 1.0, 1.0, nx=g.nx, nz=g.nz)
        ctx.set_forward_method(method)
        ctx.set_cg_precision(jbits)
        ctx.set_cg_preconditioner(prec)
        ctx.set_sigma(np.exp(par0))
        ctx.forward()
        _, sa0 = ctx.voltages(1)
        device_sync()
        t0 = time.perf_counter()
        ctx.sens(1)
        device_sync()
        t_sens = time.perf_counter() - t0
        if lam0 is None:
            lam0 = ctx.blam0(wmatd, wdfak, 1.0, 1.0)
        t0 = time.perf_counter()
        u = ctx.update(lam0, 1e-12, manz // 10, wmatd, wdfak, dat, par0)
        device_sync()
        t_upd = time.perf_counter() - t0
        t0 = time.perf_counter()
        ctx.set_sigma(np.exp(u["par"]))
        ctx.forward()
        _, sa1 = ctx.voltages(1)
        device_sync()
        t_fwd = time.perf_counter() - t0
        tm = ctx.times()
        if d_ref is None:
            d_ref = u["dpar"]
        f = {"forward_method": "band per system" if method == 0 else "condensed",
             "sens_ms": round(1e3 * t_sens, 2), "update_ms": round(1e3 * t_upd, 2),
             "forward_ms": round(1e3 * t_fwd, 2), "step_ms": round(1e3 * (t_sens + t_upd + t_fwd), 2),
             "ncg": u["ncg"], "cg_ms_per_iteration": round(tm["cg_matvec"], 4),
             "rms_before": round(INV.nrmsd(dat, sa0, wmatd), 4), "rms_after": INV.nrmsd(dat, sa1, wmatd),
             "dpar_rel_diff_vs_fp64_jacobi": float(np.max(np.abs(u["dpar"] - d_ref)) / np.max(np.abs(d_ref)))}
        if prec == 3:
            f["training"] = training
        out["gn_step"]["forms"][name] = f
    base = out["gn_step"]["forms"]["fp64_jacobi"]
    for name, f in out["gn_step"]["forms"].items():
        f["speedup_vs_fp64_jacobi"] = round(base["step_ms"] / f["step_ms"], 3)
        f["rms_after_rel_diff_vs_fp64_jacobi"] = abs(f["rms_after"] - base["rms_after"]) / base["rms_after"]
    out["gn_step"]["lambda0"] = lam0
    # the inversion loop at the reference's default eps
    for name, jbits, prec, method in forms:
        ctx.set_forward_method(method)
        ctx.set_cg_precision(jbits)
        ctx.set_cg_preconditioner(prec)
        r = INV.gauss_newton(ctx, g, dat, wmatd, par0, nit, lamfix=lam0, ltri=0, eps=eps, lam_factor=0.5)
        st = r.step_ms()
        its = [i for i in r.iterations if i.ncg]
        out["inversion_default_eps"]["forms"][name] = {
            "rms": [round(i.rms, 4) for i in r.iterations], "final_rms": r.final_rms, "steps": len(st),
            "ncg": [i.ncg for i in its], "step_ms": [round(x, 2) for x in st], "total_ms": round(float(np.sum(st)), 2),
            "phase_ms_total": {k: round(float(np.sum([i.ms.get(k, 0.0) for i in its])), 2)
                               for k in ("sens", "update", "forward", "forward_half_step")}}
    b2 = out["inversion_default_eps"]["forms"]["fp64_jacobi"]["total_ms"]
    for f in out["inversion_default_eps"]["forms"].values():
        f["time_ratio_vs_fp64_jacobi"] = round(b2 / f["total_ms"], 3)
    ctx.set_cg_precision(64)
    ctx.set_cg_preconditioner(0)
    ctx.set_forward_method(-1)
    ctx.close()
    return out


def buncert_step(ctx, w, ltri: int):
    """Model uncertainty (src/bmcm_mod.f90:49-175) on the step's Jacobian with unit
    data weights and the device regularisation: Re(J^H W J) (dsyrk), the reference's
    regularisation fill, the inverse by Cholesky, res = cov ata and cov2 = res cov;
    only the five diagonals come back to the host."""
    import numpy as np
    g = w.grid
    nanz, manz = w.survey.nanz, g.elanz
    ctx.bsmatm(ltri, 1.0, 1.0, nx=getattr(g, "nx", 0), nz=getattr(g, "nz", 0))
    wmatd = np.ones(nanz); wdfak = np.ones(nanz, np.int32)
    lam = 1.0
    device_sync()
    t0 = time.perf_counter()
    err, d = ctx.buncert(lam, wmatd, wdfak, gauss=False, full=False)
    device_sync()
    t = time.perf_counter() - t0
    flop = 2.0 * nanz * manz * manz + manz ** 3 / 3.0 + 2.0 * manz ** 3 / 3.0 + 2 * 2.0 * manz ** 3
    return {"ms": round(1e3 * t, 1), "errnr": err, "manz": manz, "nanz": nanz, "ltri": ltri,
            "flop": flop, "TFLOPs": round(flop / t / 1e12, 2),
            "note": "dsyrk 2 nanz manz^2 (as a full product count) + potrf manz^3/3 + potri 2 manz^3/3 + "
                    "two dgemm 2 manz^3, FP64; timed with host transfers of the diagonals only"}


def bench_c2(args):
    """Batched forward (the C2 workload): ``args.models`` seeded models of 5k..40k
    cells with 48..96 electrodes each, every (electrode, wavenumber) system of all
    models assembled, factorised and solved by one graph launch per step."""
    import numpy as np
    from crtomo_stable_amd import configs, lib

    t0 = time.perf_counter()
    ctxs, works, flop, nsys = [], [], 0.0, 0
    for i in range(args.models):
        w = configs.c2_model(i)
        c = lib.Context(0)
        c.set_grid(w.grid)
        c.set_survey(w.survey)
        kwn, _, _, _ = c.rwaven()
        lbeta = c.precal()
        c.set_sigma(w.sigma)
        n = kwn.size * (w.grid.eanz if lbeta else 1)
        flop += n * 8.0 * w.grid.sanz * w.grid.mb * (w.grid.mb + 1) / 2.0
        nsys += n
        ctxs.append(c)
        works.append(w)
        if (i + 1) % 32 == 0:
            print(f"[c2] {i + 1} models set up ({time.perf_counter() - t0:.0f} s)", file=sys.stderr, flush=True)
    b = lib.Batch(ctxs, autotune=True)
    for _ in range(args.warmup):
        b.forward()
    device_sync()
    dev_ms = 0.0
    t1 = time.perf_counter()
    for _ in range(args.steps):
        b.forward()
        dev_ms += b.info()[0]
    device_sync()
    elapsed = time.perf_counter() - t1
    ms_step = 1e3 * elapsed / args.steps
    dev_ms /= args.steps
    _, n_launch, variant, tune = b.info()
    assert n_launch == nsys
    mbs = np.array([w.grid.mb for w in works])
    cells = np.array([w.grid.elanz for w in works])
    achieved = flop / (dev_ms * 1e-3) / 1e12
    out = {
        "metric": "forward solves/s",
        "value": round(nsys * 1e3 / ms_step, 1),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128",
        "data": "synthetic",
        "config": {
            "workload": f"c2: {args.models} seeded quad models, {cells.min()}..{cells.max()} cells, mb "
                        f"{mbs.min()}..{mbs.max()}, every (electrode, wavenumber) system of all models in one "
                        f"graph launch (FP64 band Cholesky)",
            "models": args.models, "systems_per_step": nsys, "parallelism": "1 GPU",
        },
        "graph_ms": round(dev_ms, 3),
        "roofline": dict(bound="mfma", achieved=round(achieved, 3), peak=FP64_PEAK_TFLOPS, unit="TFLOP/s",
                         frac=round(achieved / FP64_PEAK_TFLOPS, 4), traffic=None,
                         kernel="whole graph (band-Cholesky FLOP over the graph's device time, so a lower "
                                "bound of the factor kernels' rate)",
                         flop_per_step=flop),
        "factor_variant_per_D": {int(D): int(variant[D]) for D in range(17) if variant[D]},
        "tune_ms": {int(D): [round(float(x), 3) for x in tune[D]] for D in range(17) if tune[D].any()},
        "setup_s": round(t1 - t0, 1),
    }
    # the mixed-precision variant (FP32 factorisation on the matrix cores, FP64 refinement
    # in the graph) on the first 32 models, against the FP64 batch of the same models
    b.close()
    nm = min(32, len(ctxs))
    sub = ctxs[:nm]
    nsub = sum(c.kwnanz * w.grid.eanz for c, w in zip(sub, works[:nm]))
    try:
        res = {}
        for label, mixed in (("fp64", False), ("mixed", True)):
            bx = lib.Batch(sub, mixed=mixed)
            for it in range(1 + args.steps):
                print(f"[c2] {label} batch of {nm} models: launch {it}", file=sys.stderr, flush=True)
                t2 = time.perf_counter()
                bx.forward()
                device_sync()
                ms = 1e3 * (time.perf_counter() - t2)
            res[label] = (ms, bx.info()[0], bx.refine_info() if mixed else None,
                          [c.voltages(1)[0] for c in sub])
            bx.close()
        (ms64, dev64, _, v64), (msm, devm, (sweeps, ratio, fallbacks), vm) = res["fp64"], res["mixed"]
        err = max(float(np.abs(a / r - 1).max()) for a, r in zip(vm, v64))
        out["mixed"] = dict(models=nm, systems=nsub, ms_per_launch=round(msm, 3), graph_ms=round(devm, 3),
                            fp64_ms_per_launch=round(ms64, 3), speedup_vs_fp64=round(ms64 / msm, 3),
                            value=round(nsub * 1e3 / msm, 1), unit="solves/s", max_sweeps=sweeps,
                            max_accepted_ratio=ratio, models_redone_fp64=fallbacks,
                            max_rel_voltage_err_vs_fp64=err)
    except Exception as e:
        out["mixed"] = {"error": str(e)}
    if not args.no_cpu_baseline:
        try:
            w0 = works[0]
            cb = cpu_baseline(w0, args.cpu_threads, 1, args.cpu_fwd_electrodes)
            # per-system time of model 0, scaled to every system by sanz mb^2
            n0 = ctxs[0].kwnanz * w0.grid.eanz
            t_sys = cb["forward_s"] / n0
            work0 = w0.grid.sanz * w0.grid.mb ** 2
            t_all = sum(t_sys * (c.kwnanz * w.grid.eanz) * (w.grid.sanz * w.grid.mb ** 2) / work0
                        for c, w in zip(ctxs, works))
            out["cpu_baseline"] = dict(value=round(nsys / t_all, 4), unit="solves/s", cores=cb["cores"],
                                       kind=cb["kind"],
                                       sample=cb["sample"] + f"; model 0 only (sanz {w0.grid.sanz}, mb "
                                              f"{w0.grid.mb}), its forward time per system scaled to all "
                                              f"{nsys} systems by sanz*mb^2")
        except Exception as e:
            out["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(out), flush=True)
    for c in ctxs:
        c.close()


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c1", choices=["c1", "c0", "c2", "c3"])
    ap.add_argument("--models", type=int, default=256, help="models of the c2 batch")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-gn-inversion", action="store_true",
                    help="skip the Gauss-Newton comparison of CG forms on the 400x100 workload")
    ap.add_argument("--gn-iterations", type=int, default=4)
    ap.add_argument("--gn-train-models", type=int, default=2)
    ap.add_argument("--gn", action="store_true",
                    help="also time one Gauss-Newton step (forward, bsensi, bsmatm, blam0, update/cjg)")
    ap.add_argument("--forward-method", type=int, default=-1, choices=[-1, 0, 1],
                    help="-1: the library default (condensed where the grid allows), 0: band factorisation of "
                         "every system, 1: condensed")
    ap.add_argument("--forward-precision", type=int, default=64, choices=[64, 32],
                    help="32: FP32 band factorisation with FP64 iterative refinement in the timed step")
    ap.add_argument("--buncert", action="store_true",
                    help="also time the model uncertainty (crt_buncert: Re(J^H W J), +lam R, inverse, "
                         "resolution, cov2) on the step's Jacobian")
    ap.add_argument("--cpu-threads", type=int, default=0, help="OpenMP threads of the CPU baseline (0: every "
                    "core this process may use)")
    ap.add_argument("--cpu-sens-meas", type=int, default=0, help="0: max(96, 3 x threads)")
    ap.add_argument("--cpu-fwd-electrodes", type=int, default=4)
    args = ap.parse_args()
    if args.cpu_threads <= 0:
        args.cpu_threads = host_cores()
    if args.cpu_sens_meas <= 0:
        args.cpu_sens_meas = max(96, 3 * args.cpu_threads)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only; the multi-GPU layout is described in "
              "DESIGN.md (shard the (electrode, wavenumber) systems, all-reduce the CG vectors)",
              file=sys.stderr)
        sys.exit(2)

    from crtomo_stable_amd import configs, lib
    if not os.path.exists(lib.LIB_PATH):
        from crtomo_stable_amd import build
        build.build()

    if args.workload == "c2":
        return bench_c2(args)
    w = {"c0": configs.c0, "c1": configs.c1, "c3": configs.c3}[args.workload]()
    g, s = w.grid, w.survey
    ctx = lib.Context(0)
    ctx.set_grid(g)
    ctx.set_survey(s)
    kwn, kwi, _, _ = ctx.rwaven()
    lbeta = ctx.precal()
    ctx.set_sigma(w.sigma)
    K, E = kwn.size, g.eanz
    nsys = K * E if lbeta else K
    if args.forward_precision == 32:
        ctx.set_forward_precision(32)

    def step():
        ctx.forward()
        ctx.voltages(1)
        ctx.sens(1)

    keys = ("diag", "factor", "solve", "prefix", "schur", "product", "refine", "volt", "sens")

    def timed(nsteps, nwarm):
        for _ in range(nwarm):
            step()
        device_sync()
        acc = {k: 0.0 for k in keys}
        t0 = time.perf_counter()
        for _ in range(nsteps):
            step()
            t = ctx.times()
            for key in acc:
                acc[key] += t[key]
        device_sync()
        return 1e3 * (time.perf_counter() - t0) / nsteps, {k: v / nsteps for k, v in acc.items()}

    def band_roofline(phase):
        # band Cholesky: 8 sanz mb(mb+1)/2 real FLOP per (electrode, wavenumber) system
        flop_sys = 8.0 * g.sanz * g.mb * (g.mb + 1) / 2.0
        fac_ms = phase["factor"]
        achieved = flop_sys * nsys / (fac_ms * 1e-3) / 1e12
        d16 = -(-g.mb // 16)
        peak = FP64_PEAK_TFLOPS
        if args.forward_precision == 32:
            kernel = f"crt_factor32st_kernel<{d16}>"
            peak = FP32_PEAK_TFLOPS
        elif os.environ.get("CRT_FACTOR8") or d16 > 16:
            kernel = "crt_factor_kernel"
        elif d16 == 1 or os.environ.get("CRT_FACTOR_V2") and d16 <= 10:
            kernel = f"crt_factor16_kernel<{d16}>"
        else:
            kernel = f"crt_factor16{'la' if d16 <= 10 else 'st'}_kernel<{d16}>"
        return dict(bound="mfma", achieved=round(achieved, 3), peak=peak, unit="TFLOP/s",
                    frac=round(achieved / peak, 4), traffic=measured_traffic(kernel, args.workload, "band"),
                    kernel=kernel, flop_per_system=flop_sys, systems_per_launch=nsys, launch_ms=round(fac_ms, 4))

    def condensed_rooflines(phase):
        """Algorithmic FLOP of each kernel of the condensed forward over its measured
        phase time (HIP events on the library's stream):
          interior band factorisation: 8 nI mbI(mbI+1)/2 per wavenumber;
          substitution (forward and back, both passes): 2 x 8 nI mbI per right-hand side,
            nG + E right-hand sides per wavenumber;
          separator elimination: (4/3) (nB^3 - nG^3) per wavenumber (the first nS columns of
            the complex factorisation over B = [S, G], nB = nS + nG, with the update of G);
          dense boundary factorisation: (4/3) nG^3 + 16 nG^2 per (electrode, wavenumber);
          interior potentials: 8 nI nG per (electrode, wavenumber)."""
        info = ctx.condense_info()
        nI, nG, mbI, DI = info["interior"], info["boundary"], info["band"], info["tiles"]
        nS = info["separator"]
        JS, nt = (nS + 15) // 16, (nG + 15) // 16
        NW = 4 if nt <= 20 else 8  # waves per source system (the library's choice)
        NI, NIp = (nt + NW - 1) // NW, (JS + 7) // 8  # the prefix kernel holds the separator block alone
        re = "false"  # the complex kind (the DC kind runs the real instantiations)
        ent = {
            "factor": (f"crt_factor16{'la' if DI <= 10 else 'st'}_kernel<{DI}>",
                       K * 8.0 * nI * mbI * (mbI + 1) / 2.0, K),
            "solve": (f"crt_cond_trsm_kernel<{DI}, false, {re}> + crt_cond_trsm_kernel<{DI}, true, {re}>",
                      K * (nG + E) * 2.0 * 8.0 * nI * mbI, K * (nG + E)),
            "prefix": (f"crt_cond_dense_kernel<{NIp}, true, 8, {re}>", K * 4.0 / 3.0 * ((nS + nG) ** 3 - nG ** 3),
                       K),
            "schur": (f"crt_cond_dense_kernel<{NI}, false, {NW}, {re}>",
                      K * E * (4.0 / 3.0 * nG ** 3 + 16.0 * nG ** 2), K * E),
            "product": (f"crt_cond_gemm_kernel<{re}>", K * E * 8.0 * nI * nG, K * E),
        }
        out = {}
        for key, (kernel, flop, units) in ent.items():
            ms = phase[key]
            if ms <= 0:
                continue
            a = flop / (ms * 1e-3) / 1e12
            out[key] = dict(bound="mfma", achieved=round(a, 3), peak=FP64_PEAK_TFLOPS, unit="TFLOP/s",
                            frac=round(a / FP64_PEAK_TFLOPS, 4), traffic=measured_traffic(kernel.split(" + ")[0],
                                                                                          args.workload, "condensed"),
                            kernel=kernel, flop_per_launch=flop, units_per_launch=units, launch_ms=round(ms, 4))
        return info, out

    ctx.set_forward_method(args.forward_method)
    ms_step, phase = timed(args.steps, args.warmup)
    method = ctx.forward_method()
    if method == 1:
        info, rl = condensed_rooflines(phase)
        top = max(rl, key=lambda k: rl[k]["launch_ms"])
        roofline = rl[top]
        others = {k: v for k, v in rl.items() if k != top}
        others.update({k: v for k, v in other_rooflines(g, s, K, E, nsys, phase, args.forward_precision, args.workload).items()
                       if k.startswith("Jacobian")})
    else:
        info = None
        roofline = band_roofline(phase)
        others = other_rooflines(g, s, K, E, nsys, phase, args.forward_precision, args.workload)

    out = {
        "metric": "forward+Jacobian solves/s",
        "value": round(1e3 / ms_step, 4),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128",
        "data": "synthetic",
        "config": {
            "workload": {"c1": "c1: seeded Delaunay mesh 100x30 m, 2 m sinusoidal topography, complex GRF model, "
                               "complex forward + bsensi Jacobian",
                         "c0": "c0: 60x20 quads, 30 electrodes, 188 dipole-dipole, forward + bsensi",
                         "c3": "c3: 400x100 quads (40501 nodes, mb 201), 96 electrodes, 4500 dipole-dipole + "
                               "Wenner, layered model with three blocks, forward + bsensi"}[args.workload],
            "elements": g.elanz, "nodes": g.sanz, "mb": g.mb, "electrodes": E, "measurements": s.nanz,
            "wavenumbers": K, "systems_per_step": nsys, "parallelism": "1 GPU",
            "forward_method": "condensed" if method == 1 else "band per system",
        },
        "forward_systems_per_s": round(nsys * 1e3 / ms_step, 1),
        "phase_ms": {k: round(v, 4) for k, v in phase.items()},
        "roofline": roofline,
        "roofline_other_kernels": others,
    }
    if info is not None:
        out["condensation"] = info
    if method == 1 and args.forward_precision == 64:
        # the reference's algorithm on the same workload: every system factorised
        ctx.set_forward_method(0)
        try:
            ms_b, phase_b = timed(max(2, args.steps // 2), 1)
            out["band_per_system"] = {"ms_per_step": round(ms_b, 3), "value": round(1e3 / ms_b, 4),
                                      "phase_ms": {k: round(v, 4) for k, v in phase_b.items()},
                                      "roofline": band_roofline(phase_b),
                                      "speedup_of_condensed": round(ms_b / ms_step, 3)}
        finally:
            ctx.set_forward_method(args.forward_method)
    if args.forward_precision == 32:
        sweeps, ratio, fallbacks = ctx.refine_info()
        out["dtype"] = "c64 factorisation, c128 refinement"
        out["refinement"] = {"sweeps": sweeps, "final_correction_ratio": ratio, "fp64_fallback_batches": fallbacks}
        ctx.set_forward_precision(64)
    if args.gn:
        out["gn_step"] = gn_step(ctx, w, 0 if getattr(g, "nx", 0) else 1)
    if args.buncert:
        out["buncert"] = buncert_step(ctx, w, 0 if getattr(g, "nx", 0) else 1)
    if not args.no_gn_inversion and args.workload == "c1":
        ctx.close()
        ctx = None
        out["gn_inversion"] = gn_inversion(args.gn_iterations, args.gn_train_models, 100)
    if not args.no_cpu_baseline:
        try:
            cb = cpu_baseline(w, args.cpu_threads, args.cpu_sens_meas, args.cpu_fwd_electrodes)
            out["cpu_baseline"] = {k: (round(v, 6) if isinstance(v, float) else v) for k, v in cb.items()}
        except Exception as e:  # the baseline is reported, never the target
            out["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(out), flush=True)
    if ctx is not None:
        ctx.close()


if __name__ == "__main__":
    main()
