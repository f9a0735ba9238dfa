"""printout's files, the restart file and the full-state checkpoint on the
GPU, against the flang-built reference (oracle/_ref/libref_66x34x34.so, run
on the CPU in a child process) and against themselves.

- Same state, same bytes: the reference runs 3 steps and calls printout at
  t = 100 (every output fires: tprint 5, ntecplot 50, the pressure plane,
  tsave 100); its final fields and scalars are copied into the GPU, whose
  dg_printout(100) must write mainNNNN.plt, pressureNNNN.plt, velocity.plt,
  spurious.plt, normv.plt and hysteresis.plt byte for byte as the
  reference, printout's lines of outputdata.out and of the screen byte for
  byte, and save.dat byte for byte as the records the reference's savedata
  writes (to unit 2, fort.2).
- A whole run at the output cadence (dg_run with the files open) against the
  reference's loop with printout: the same files and records, numbers equal
  to 1e-6 of their column's scale (the fields themselves differ from the
  reference's by the whole-step parity bar, 1e-10).
- The reference's restart: save.dat from the GPU's savedata, read by the
  reference's iniget and by dg_init with restart = 1; the fields after
  initialize agree bit for bit and 2 steps later to the whole-step bar.
- The full state: a run resumed with dg_load_state from dg_save_state after
  step 2 is bit-identical at step 4 to the run that went on, for CSF, SGIP
  with CalcH3d heights and the tolerance-mode solver.
- The reference's own program (main.f without the time-step units, linked
  against the library, oracle/link_dropin.py) run on the GPU: its own
  printout and savedata write the same files as the CPU reference's run.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import refrun
from oracle_ffi import clipped_cc, inputdata_text, shipped_text

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DROPIN = os.path.join(REPO, "oracle", "_ref", "three_dim_dropin")
FILES = ("velocity.plt", "spurious.plt", "normv.plt", "hysteresis.plt")
PRINTOUT_KEYS = ("*****", "timestep=", "time=", "tau=", "CFL =", "max |v|=", "total x momentum=", "max mut/mu1=")

# 66x34x34 (the reference's compiled extents), 8 mm long, a 0.8 mm droplet
# falling at 1 m/s: the reference stays finite over these steps
BASE = dict(nx=66, ny=34, nz=34, length=0.008, radius=0.0008, y0=0.002)


def need_ref():
    if not refrun.have_ref(66, 34, 34):
        pytest.skip("oracle/_ref/libref_66x34x34.so not built (oracle/build_ref.sh 6,5,5)")


def read(path):
    with open(path, "rb") as f:
        return f.read()


STARS = "*" * 50 + " " + "*" * 25


def printout_lines(text):
    """printout's lines of outputdata.out or of the screen (main.f:5035-5067),
    from its first star line on (initialize's echo of the input comes
    before)."""
    lines = text.splitlines()
    first = next((q for q, ln in enumerate(lines) if STARS in ln), len(lines))
    out = []
    for ln in lines[first:]:
        s = ln.strip()
        if any(s.startswith(k) for k in PRINTOUT_KEYS) or s.startswith("ntecplot print"):
            out.append(ln)
    return out


def numbers(path):
    """The records of a .plt file (header lines skipped) as an array."""
    rows = []
    with open(path) as f:
        for ln in f:
            s = ln.split()
            if not s or s[0].startswith(("VARIABLES", "zone")):
                continue
            rows.append([float(x) for x in s])
    return np.array(rows)


def close_arrays(a, b, rel=1e-6):
    assert a.shape == b.shape, (a.shape, b.shape)
    scale = np.maximum(np.abs(b).max(axis=0), 1e-300)
    err = (np.abs(a - b) / scale).max()
    assert err <= rel, err


def restart_records(path):
    """savedata's records (main.f:5418-5465): the header values and the
    (n, 4) cell records."""
    raw = read(path)
    pos, vals = 0, []
    while pos < len(raw):
        n = int(np.frombuffer(raw[pos:pos + 4], dtype="<i4")[0])
        vals.append(raw[pos + 4:pos + 4 + n])
        assert raw[pos + 4 + n:pos + 8 + n] == raw[pos:pos + 4]
        pos += 8 + n
    head, cells = vals[:36], vals[36:]
    return head, np.frombuffer(b"".join(cells), dtype="<f8").reshape(-1, 4)


def test_printout_same_state_same_bytes(dg, tmp_path, capfd):
    need_ref()
    txt = inputdata_text(tmax=6, tprint=5, ntecplot=50, tsave=100, **BASE)
    print("case printout at t=100 of the reference's state after 3 steps, 66x34x34", flush=True)
    ra, gb = tmp_path / "ref", tmp_path / "gpu"
    ra.mkdir()
    gb.mkdir()
    ref = refrun.run_reference(txt, 3, workdir=str(ra), printout_ts=[100],
                               want_fields=("cc", "u", "v", "w", "p", "mut"))
    sc = ref["scalars"]
    g = dg.DropletGPU(inputdata_text=txt)
    try:
        for nm, a in ref["fields"].items():
            g.set(nm, a)
        g.set_state(tau=sc["tau"], time=sc["time"], tauh=sc["tauh"], tauh2=sc["tauh2"], tswap=sc["tswap"],
                    npre=sc["npre"], ntop=sc["ntop"], npost=sc["npost"])
        capfd.readouterr()
        g.output_open(str(gb), echo_stdout=True)
        g.printout(100)
        g.output_flush()
        screen = capfd.readouterr().out
    finally:
        g.output_close()
        g.close()
    for name in ("main0002.plt", "pressure0001.plt") + FILES:
        assert read(gb / name) == read(ra / name), name
    assert read(gb / "save.dat") == read(ra / "fort.2")
    ours = (gb / "outputdata.out").read_text().splitlines()
    theirs = printout_lines((ra / "outputdata.out").read_text())
    assert len(ours) == 8 and ours == theirs[-8:], (ours, theirs[-8:])
    ref_screen = printout_lines((ra / "stdout.txt").read_text())
    assert printout_lines(screen) == ref_screen[-9:], (screen, ref_screen[-9:])


def test_run_with_output_matches_reference_loop(dg, tmp_path):
    need_ref()
    nsteps = 7
    txt = inputdata_text(tmax=nsteps - 1, tprint=2, ntecplot=3, tsave=4, **BASE)
    print(f"case dg_run with printout, {nsteps} steps, tprint 2, ntecplot 3, tsave 4", flush=True)
    ra, gb = tmp_path / "ref", tmp_path / "gpu"
    ra.mkdir()
    gb.mkdir()
    refrun.run_reference(txt, nsteps, workdir=str(ra), loop_printout=True, want_fields=("cc",))
    g = dg.DropletGPU(inputdata_text=txt)
    try:
        g.output_open(str(gb))
        g.run(nsteps, t0=0)
        g.output_flush()
        stats = g.output_stats()
    finally:
        g.output_close()
        g.close()
    assert stats["jobs"] >= 5
    for name in ("main0000.plt", "main0001.plt", "main0002.plt", "pressure0000.plt") + FILES:
        a, b = gb / name, ra / name
        assert a.exists(), name
        ta, tb = a.read_text().splitlines(), b.read_text().splitlines()
        assert len(ta) == len(tb), name
        assert [ln for ln in ta if "VARIABLES" in ln or "zone" in ln] == \
               [ln for ln in tb if "VARIABLES" in ln or "zone" in ln], name
        na, nb = numbers(a), numbers(b)
        if nb.size:
            close_arrays(na, nb)
    ours = (gb / "outputdata.out").read_text().splitlines()
    theirs = printout_lines((ra / "outputdata.out").read_text())
    assert len(ours) == len(theirs) == 4 * 8
    for x, y in zip(ours, theirs):
        if "*" in x or "timestep" in x:
            assert x == y
            continue
        fx = [float(w) for w in x.replace("=", " ").split() if w[0].isdigit() or w[0] in "-."]
        fy = [float(w) for w in y.replace("=", " ").split() if w[0].isdigit() or w[0] in "-."]
        assert np.allclose(fx, fy, rtol=1e-8, atol=1e-10), (x, y)
    ha, ca = restart_records(gb / "save.dat")
    hb, cb = restart_records(ra / "fort.2")
    assert ha == hb
    for q in range(4):
        den = max(np.linalg.norm(cb[:, q]), 1e-300)
        assert np.linalg.norm(ca[:, q] - cb[:, q]) / den <= 1e-10, q


def rel_l2(a, b):
    a, b = a[1:-1, 1:-1, 1:-1], b[1:-1, 1:-1, 1:-1]
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_restart_file_matches_reference_iniget(dg, tmp_path):
    need_ref()
    print("case savedata after 3 steps, restart = 1 on the GPU and in the reference", flush=True)
    txt = inputdata_text(tmax=6, **BASE)
    g = dg.DropletGPU(inputdata_text=txt)
    try:
        g.run(3, t0=0)
        g.savedata(str(tmp_path / "save.dat"))
        saved = {nm: g.get(nm) for nm in ("u", "v", "w", "cc")}
    finally:
        g.close()
    rtxt = inputdata_text(tmax=6, restart=1, **BASE)
    (tmp_path / "inputdata.in").write_text(rtxt)
    ref0 = refrun.run_reference(rtxt, 0, workdir=str(tmp_path), want_fields=("cc", "u", "v", "w", "rhon", "mun"))
    ref2 = refrun.run_reference(rtxt, 2, workdir=str(tmp_path), want_fields=("cc", "u", "v", "w", "p"))
    (tmp_path / "inputdata.in").write_text(rtxt)
    g = dg.DropletGPU(inputdata_path=str(tmp_path / "inputdata.in"))
    try:
        t0, tmax = g.run_limits()
        assert (t0, tmax) == (0, 6)
        for nm in ("u", "v", "w", "cc"):
            assert np.array_equal(g.get(nm)[1:-1, 1:-1, 1:-1], saved[nm][1:-1, 1:-1, 1:-1]), nm
        for nm, a in ref0["fields"].items():
            assert np.array_equal(g.get(nm), a), nm
        g.run(2, t0=t0)
        errs = {nm: rel_l2(g.get(nm), a) for nm, a in ref2["fields"].items()}
        print("after 2 steps from the restart", {k: f"{v:.1e}" for k, v in errs.items()}, flush=True)
        for nm, e in errs.items():
            assert e <= 1e-10, (nm, e)
    finally:
        g.close()


STATE_FIELDS = ("cc", "u", "v", "w", "p", "s11", "s22", "s33", "s12", "s13", "s23", "rhon", "mun", "a", "b", "c",
                "ei", "res", "rhs", "pprev")


@pytest.mark.parametrize("case", ["csf", "sgip", "csf-tolerance"])
def test_full_state_resume_is_bit_identical(dg, tmp_path, case):
    print(f"case full-state resume after step 2, {case}", flush=True)
    if case == "sgip":
        txt = shipped_text(surfacet=5)
        cc0 = clipped_cc(txt)
        fields = STATE_FIELDS + ("hxmin", "hxmax", "hymin", "hymax", "hzmin", "hzmax")
    else:
        txt = inputdata_text(tmax=6, **BASE)
        cc0 = None
        fields = STATE_FIELDS

    def start():
        g = dg.DropletGPU(inputdata_text=txt)
        if cc0 is not None:
            g.set("cc", cc0)
        if case == "csf-tolerance":
            g.set_solver(mode=2, tol=1e-8)
        return g

    g = start()
    try:
        for t in range(3):
            g.step(t)
        g.save_state(str(tmp_path / "state.bin"))
        for t in (3, 4):
            g.step(t)
        a = {nm: g.get(nm) for nm in fields}
        sa = g.state()
    finally:
        g.close()
    g = start()
    try:
        g.load_state(str(tmp_path / "state.bin"))
        for t in (3, 4):
            g.step(t)
        b = {nm: g.get(nm) for nm in fields}
        sb = g.state()
    finally:
        g.close()
    for nm in fields:
        assert np.array_equal(a[nm], b[nm], equal_nan=True), nm
    for k in ("tau", "time", "tauh", "tauh2", "tswap", "npre", "ntop", "npost"):
        assert sa[k] == sb[k], k


def test_state_file_rejects_other_grid(dg, tmp_path):
    print("case state file of another grid", flush=True)
    g = dg.DropletGPU(inputdata_text=inputdata_text(nx=18, ny=18, nz=18))
    try:
        g.save_state(str(tmp_path / "s.bin"))
    finally:
        g.close()
    g = dg.DropletGPU(inputdata_text=inputdata_text(nx=34, ny=18, nz=18))
    try:
        with pytest.raises(dg.DropletError, match="is for 18 x 18 x 18"):
            g.load_state(str(tmp_path / "s.bin"))
        (tmp_path / "bad.bin").write_bytes(b"not a state")
        with pytest.raises(dg.DropletError, match="not a state file"):
            g.load_state(str(tmp_path / "bad.bin"))
    finally:
        g.close()


def test_reference_program_linked_to_library(dg, tmp_path):
    """main.f minus the time-step units, linked against libdroplet_gpu.so
    (built by oracle/link_dropin.py), runs its unchanged loop on the GPU;
    its own printout and savedata see the fields the library downloads at
    the end of each solvelpres_ and write the reference's files."""
    need_ref()
    if not os.path.exists(DROPIN):
        pytest.skip("oracle/_ref/three_dim_dropin not built (python oracle/link_dropin.py)")
    nsteps = 7
    txt = inputdata_text(tmax=nsteps - 1, tprint=2, ntecplot=3, tsave=4, **BASE)
    print(f"case the reference's program on the library, {nsteps} steps", flush=True)
    ra, gb = tmp_path / "ref", tmp_path / "gpu"
    ra.mkdir()
    gb.mkdir()
    refrun.run_reference(txt, nsteps, workdir=str(ra), loop_printout=True, want_fields=("cc",))
    (gb / "inputdata.in").write_text(txt)
    r = subprocess.run([DROPIN], cwd=str(gb), capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, (r.returncode, r.stdout[-3000:], r.stderr[-3000:])
    for name in ("main0000.plt", "main0001.plt", "main0002.plt", "pressure0000.plt") + FILES:
        na, nb = numbers(gb / name), numbers(ra / name)
        if nb.size:
            close_arrays(na, nb)
        else:
            assert na.size == 0
    ha, ca = restart_records(gb / "fort.2")
    hb, cb = restart_records(ra / "fort.2")
    assert ha == hb
    for q in range(4):
        assert np.linalg.norm(ca[:, q] - cb[:, q]) / max(np.linalg.norm(cb[:, q]), 1e-300) <= 1e-10, q
    ours = printout_lines((gb / "outputdata.out").read_text())
    theirs = printout_lines((ra / "outputdata.out").read_text())
    assert len(ours) == len(theirs)
    shutil.rmtree(str(gb), ignore_errors=True)
