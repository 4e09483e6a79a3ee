"""Writes the crtomo.cfg / CRTomo-run fixtures from the reference's own committed tomodirs
(data the reference holds: every exe/crtomo.cfg with the inv.ctr its run wrote).

crtomo_cfg_echo.json: per tomodir the text of crtomo.cfg, which of its side files
existed at run time (the prior model, crt.noisemod) and the configuration the reference
echoed into inv.ctr (src/kont1.f90) after reading the file (src/rall.f90:141-593).

crtomo_runs.json: the committed runs of the reference's tomodirs with noise-free data
(tests/minimal, tests/2.5D_mixed_surf, tests/DC_Problem, tests/non-ordered_boundary_conditions,
frontcover/inversion, examples/uncert/test_dev: crtomo.cfg, the inv.ctr iteration lines of
src/kont2.f90, the REAL*8 misfit of rho00.mag, the forward voltages of volt00.dat, error.dat
where the run stopped, and for the uncertainty run its coverage and .diag files) for the
inversion tests; their grid, electrodes and measurements are copied to
tests/golden/reference_fixtures/<inputs>.

    python tests/golden/make_crtomo_golden.py
"""
import glob
import json
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"

PARAMS = [  # the kont1 lines with a "! comment", in order
    ("nx", "nx-switch"), ("nz", "nz-switch"), ("alfx", "smoothing parameter in x"), ("alfz", "smoothing parameter in z"),
    ("itmax", "max. # inversion iterations"), ("ldc", "DC inversion"), ("lrobust", "robust inversion"),
    ("lfphai", "final phase improvement"), ("stabw0", "rel. resistance error"), ("stabm0", "min. abs. resistance"),
    ("stabpA1", "parameter A1 (mrad"), ("stabpB", "parameter B  (-)"), ("stabpA2", "parameter A2 (%)"),
    ("stabp0", "parameter A3 (mrad)"), ("lrho0", "homogeneous background"), ("bet0", "background magnitude"),
    ("pha0", "background phase"), ("lagain", "Another dataset"), ("swrtr", "2D (=0) or 2.5D"),
    ("lsink", "fictitious sink ?"), ("nsink", "fictitious sink node"), ("lrandb2", "boundary values ?"),
]
FLAGS = [
    ("lstart", "Read start model?"), ("lsens", "Write coverage?"), ("lcov1", "Write MCM 1?"),
    ("lres", "Write resolution?"), ("lcov2", "Write MCM 2?"), ("lgauss", "Using Gauss ols?"),
    ("lphi0", "Forcing negative phase?"), ("lsytop", "Calculate sytop?"), ("lverb", "Verbose?"),
    ("lelerr", "Error Ellipses?"), ("lffhom", "Restart FPI with homogenous phase?"),
    ("llamf_fixed", "Fixed lambda?"), ("lnse", "Add data noise ?"), ("couple", "Couple to Err. Modl?"),
]


def val(tok):
    if tok in ("T", "F"):
        return tok == "T"
    try:
        return int(tok)
    except ValueError:
        return float(tok)


def echo(inv_ctr):
    lines = open(inv_ctr, errors="replace").read().splitlines()
    i = next(k for k, l in enumerate(lines) if l.strip() == "##" and k > 2 and lines[k - 1].startswith("## Created"))
    body = lines[i + 2:]
    e = {}
    first = body[0].split()
    e["mswitch"] = 0 if body[0].strip() == "***FILES***" else int(first[0])
    for k, name in enumerate(("elem", "elec", "volt", "ramd")):
        e[name] = body[1 + k].split(" # ")[0].strip()
    e["ldiff_or_prior"] = body[5].split()[0] == "T"
    e["dd0"], e["dm0"], e["dfm0"] = (body[6 + k].strip() for k in range(3))
    rest = body[9:]
    for name, key in PARAMS:
        hit = [l for l in rest if "!" in l and key in l.split("!", 1)[1]]
        if hit:
            e[name] = val(hit[0].split()[0])
    # after the boundary-values switch: drandb, then ltri (with the fixed-lambda bits 32
    # and 64 in the versions that print them)
    k = next(k for k, l in enumerate(rest) if "boundary values ?" in l)
    e["drandb"] = rest[k + 1].strip()
    e["ltri"] = int(rest[k + 2].split()[0]) % 32
    for name, label in FLAGS:
        hit = [l for l in lines if l.startswith(label)]
        if hit:
            e[name] = hit[0][len(label):].strip().split()[0] == "T"
    hit = [l for l in lines if l.startswith("Lambda=")]
    if hit:
        e["lamfix"] = float(hit[0].split()[1])
    return e


def cfg_cases():
    out = []
    for cfg in sorted(glob.glob(f"{REF}/**/exe/crtomo.cfg", recursive=True)):
        exe = os.path.dirname(cfg)
        body = [l for l in open(cfg, errors="replace").read().splitlines() if not l.lstrip().startswith("#")]
        ramd = body[4].split()[0] if len(body) > 4 and body[4].split() else ""
        inv = os.path.normpath(os.path.join(exe, ramd, "inv.ctr"))
        if not os.path.exists(inv):
            continue
        e = echo(inv)
        side = []
        if e.get("lstart") and e.get("dm0"):
            side.append(e["dm0"])
        if e.get("lnse") and e.get("couple") is False:
            side.append("crt.noisemod")
        out.append(dict(name=os.path.relpath(os.path.dirname(exe), REF), cfg=open(cfg, errors="replace").read(),
                        side_files=side, echo=e))
    return out


# the reference's own test tomodirs with a committed run, and where their grid,
# electrodes and measurements lie among the fixtures
RUNS = {
    "minimal": ("tests/minimal/Reference", "minimal"),
    "mixed_surf": ("tests/2.5D_mixed_surf/Reference", "mixed_surf"),
    "dc_problem_complex": ("tests/DC_Problem/Complex", "dc_problem"),
    "dc_problem_dc": ("tests/DC_Problem/DC", "dc_problem"),
    "frontcover": ("frontcover/inversion", "frontcover"),
    "non_ordered_bc": ("tests/non-ordered_boundary_conditions/Reference", "non_ordered_bc"),
    "uncert_dc": ("examples/uncert/test_dev", "uncert_dc"),
}


def numeric_rows(path):
    """The numeric rows of a list-directed output file (header and Max/Min lines dropped)."""
    out = []
    for l in open(path).read().splitlines()[1:]:
        t = l.split()
        try:
            out.append([float(x) for x in t])
        except ValueError:
            continue
    return out


def committed_run(tomodir, inputs):
    d = f"{REF}/{tomodir}"
    rows = []
    for l in open(f"{d}/inv/inv.ctr").read().splitlines():
        if re.match(r"^(PIT|PUP| IT| UP)\s", l):
            rows.append(dict(id=l[0:3].strip(), it=int(l[4:7]), text=l))
    out = dict(tomodir=tomodir, inputs=inputs, cfg=open(f"{d}/exe/crtomo.cfg").read(), rows=rows)
    mag = f"{d}/inv/rho00.mag"
    if os.path.exists(mag):
        out["betrms0"] = float(open(mag).readline().split()[-1])
    volt = f"{d}/inv/volt00.dat"
    if os.path.exists(volt):
        out["volt00"] = [list(map(float, l.split()[2:4])) for l in open(volt).read().splitlines()[1:] if l.split()]
    err = f"{d}/exe/error.dat"
    if os.path.exists(err) and not out["rows"][-1:] == []:
        out["error_dat"] = open(err).read()
    # the run ended normally when inv.ctr ends with ***finished***
    out["finished"] = open(f"{d}/inv/inv.ctr").read().rstrip().endswith("***finished***")
    # the uncertainty run's final outputs: coverage and the five main diagonals
    if tomodir.endswith("uncert/test_dev"):
        out["final"] = {f: numeric_rows(f"{d}/inv/{f}") for f in
                        ("coverage.mag", "ata.diag", "ata_reg.diag", "cov1_m.diag", "res_m.diag", "cov2_m.diag")}
    return out


def copy_inputs(tomodir, inputs):
    dst = os.path.join(HERE, "reference_fixtures", inputs)
    os.makedirs(dst, exist_ok=True)
    for src in ("grid/elem.dat", "grid/elec.dat", "mod/volt.dat"):
        target = os.path.join(dst, os.path.basename(src))
        data = open(f"{REF}/{tomodir}/{src}", "rb").read()
        if os.path.exists(target):
            assert open(target, "rb").read() == data, target
        else:
            open(target, "wb").write(data)


if __name__ == "__main__":
    cases = cfg_cases()
    with open(os.path.join(HERE, "crtomo_cfg_echo.json"), "w") as f:
        json.dump(cases, f, indent=1)
    runs = {}
    for name, (tomodir, inputs) in RUNS.items():
        copy_inputs(tomodir, inputs)
        runs[name] = committed_run(tomodir, inputs)
    with open(os.path.join(HERE, "crtomo_runs.json"), "w") as f:
        json.dump(runs, f, indent=1)
    print(len(cases), "crtomo.cfg cases")
