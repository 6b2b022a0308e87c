"""CPU tests of the drop-in boundary: the libraries load and export every
symbol their headers declare; the ctypes tables match the headers.  No
compute calls (there is no GPU here)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))


def header_functions(name):
    text = open(os.path.join(ROOT, "include", name)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:opcuda|ophip|opm)_\w+)\s*\(", text)))


def exported(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True,
                         check=True).stdout
    return {l.split()[-1] for l in out.splitlines() if " T " in l}


def test_ophip_exports_every_declared_symbol():
    names = header_functions("ophip.h")
    assert len(names) >= 55
    ex = exported(os.path.join(ROOT, "oplib_amd", "libophip.so"))
    missing = [n for n in names if n not in ex]
    assert not missing, missing


def test_opmodel_exports_every_declared_symbol():
    names = header_functions("opmodel.h")
    ex = exported(os.path.join(ROOT, "oplib_amd", "libopmodel.so"))
    missing = [n for n in names if n not in ex]
    assert not missing, missing


def test_ctypes_table_matches_header():
    from oplib_amd import ophip
    names = set(header_functions("ophip.h"))
    assert names == set(ophip.SIGNATURES), names ^ set(ophip.SIGNATURES)


def test_reference_entry_points_all_present():
    # every opcuda_* the reference's C# layer imports for this path
    # (CModel Device.cs:52-105, CLVModel.cs:141-143, CArray.cs:186-211,
    # CStopWatch.cs:44-54, gsgemm4/gsgemv4/gmt Program.cs)
    wanted = """opcuda_cublas_init opcuda_shutdown opcuda_thread_synchronize opcuda_set_device
    opcuda_device_get_count opcuda_get_status opcuda_mem_alloc opcuda_mem_free_device
    opcuda_mem_alloc_host opcuda_mem_free_host opcuda_memcpy_h2d opcuda_memcpy_d2h
    opcuda_memcpy_d2d opcuda_create_timer opcuda_reset_timer opcuda_stop_timer
    opcuda_get_timer_value opcuda_scopy opcuda_scopy1 opcuda_ssetall opcuda_ssetone
    opcuda_sgemm opcuda_sgemv opcuda_ssqmm opcuda_sgemm4 opcuda_sgemv3 opcuda_sgemv4
    opcuda_sek opcuda_sck opcuda_sglvg opcuda_mc_load_mt_gpu opcuda_mc_nrng
    opcuda_mc_status_sz opcuda_mc_setseed opcuda_mc1f opcuda_mt_benchmark""".split()
    ex = exported(os.path.join(ROOT, "oplib_amd", "libophip.so"))
    assert not [w for w in wanted if w not in ex]


def test_status_word_without_device():
    from oplib_amd import ophip
    # reading the status never touches the device
    assert isinstance(ophip.lib.opcuda_get_status(), int)


def test_no_oracle_in_product():
    # the product never imports, links or loads the oracle
    pkg = os.path.join(ROOT, "oplib_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".cpp", ".h")) or f == "Makefile":
                text = open(os.path.join(dirpath, f), errors="ignore").read()
                assert "liboracle" not in text and "from oracle" not in text, f
    out = subprocess.run(["ldd", os.path.join(pkg, "libopmodel.so")], capture_output=True,
                         text=True).stdout
    assert "oracle" not in out


def test_opmodel_grid_matches_oracle():
    from oplib_amd import opmodel as M
    from oracle import oracle as O
    import numpy as np
    for cfg in (M.config_lv1d(), M.config_rs2048(), M.config_small()):
        x, x0 = M.grid(cfg)
        xo, x0o = O.xgrid(cfg.nx, cfg.xlo, cfg.xhi, cfg.spot)
        assert x0 == x0o and np.array_equal(x, xo)
