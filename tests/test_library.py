"""The C-ABI library builds, loads without a GPU and exports every entry
point include/droplet_gpu.h declares; without a GPU it fails loudly (there
is no CPU fallback)."""
import ctypes
import os
import re

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
HEADER = os.path.join(REPO, "include", "droplet_gpu.h")


def declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|double|const char \*)\s*(\w+)\s*\(", txt, flags=re.M)))


def test_header_declares_entry_points():
    names = declared()
    for must in ("dg_init", "dg_step", "dg_solvelpres", "dg_vofmethod", "dg_csf", "dg_checktime", "dg_relax",
                 "dg_poisson_solve", "dg_get_field", "dg_last_error", "vofmethod_", "csf_", "solvelpres_",
                 "checktime_", "kepsmodel_", "cfl_", "dg_sync_to_host_"):
        assert must in names


def test_library_exports_every_declared_symbol(dg):
    lib = ctypes.CDLL(os.path.join(REPO, "3d_droplet_simulation_2007_amd", "libdroplet_gpu.so"))
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_no_cpu_fallback_without_gpu(dg):
    """On a machine without a GPU, initialisation reports a HIP error instead
    of computing anything on the CPU."""
    if os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK):
        pytest.skip("a GPU is present")
    txt = dg.make_inputdata(nx=18, ny=18, nz=18)
    with pytest.raises(dg.DropletError):
        dg.DropletGPU(inputdata_text=txt)


def test_make_inputdata_matches_reference_order(dg):
    """make_inputdata writes the 38 values in readinput's order (main.f:188-234)."""
    txt = dg.make_inputdata(nx=34, ny=18, nz=18, tmax=7, surfacet=2, radius=0.001)
    vals = [ln.split()[0] for ln in txt.splitlines()]
    assert len(vals) == 38
    assert vals[:4] == ["34", "18", "18", "7"]
    assert float(vals[31]) == 0.001 and vals[37] == "2"
    assert np.isclose(float(vals[22]), 1000.0) and np.isclose(float(vals[21]), 1.2)
