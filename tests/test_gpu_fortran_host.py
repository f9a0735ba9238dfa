"""The Fortran host (ISO_C_BINDING over the C-ABI) runs the GPU time step
from an unchanged inputdata.in; its fields match the CPU restatement to the
parity bar (1e-10 relative L2)."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from oracle_ffi import Oracle, inputdata_text

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(REPO, "3d_droplet_simulation_2007_amd", "droplet_host")


def test_fortran_host_matches_restatement(dg):
    if not os.path.exists(HOST):
        pytest.fail("droplet_host was not built (flang missing at build time?)")
    print("case Fortran host 18^3, 6 steps", flush=True)
    txt = inputdata_text(nx=18, ny=18, nz=18, tmax=5)
    with tempfile.TemporaryDirectory() as d:
        inp = os.path.join(d, "inputdata.in")
        out = os.path.join(d, "fields.bin")
        with open(inp, "w") as f:
            f.write(txt)
        r = subprocess.run([HOST, inp, "6", "0", out], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        assert "final t=" in r.stdout
        data = np.fromfile(out, dtype=np.float64).reshape(5, 18, 18, 18)
    o = Oracle(txt)
    o.run(6)
    for q, name in enumerate(("cc", "u", "v", "w", "p")):
        a, b = data[q][1:-1, 1:-1, 1:-1], o.field(name)[1:-1, 1:-1, 1:-1]
        err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)
        assert err <= 1e-10, (name, err)
