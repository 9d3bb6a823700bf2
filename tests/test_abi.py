"""CPU checks of the drop-in boundary: libdfg.so loads and exports exactly the
entry points include/dfg.h declares (no compute calls without a GPU), and the
host-side configuration rules agree with the oracle's independent layout."""
import ctypes
import os
import re
import subprocess

import pytest

import dfguide as G
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dfg.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dfg_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_python_mirror():
    assert _declared() == sorted(G.EXPORTS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(G.LIB_PATH):
        pytest.skip("libdfg.so not built")
    lib = ctypes.CDLL(G.LIB_PATH)
    for name in _declared():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", G.LIB_PATH], capture_output=True, text=True).stdout
    exported = sorted(set(re.findall(r" T (dfg_[a-z_0-9]+)$", out, flags=re.M)))
    assert exported == _declared()
    G.load_library()  # argtypes for every symbol resolve


def test_header_compiles_as_c():
    src = '#include "dfg.h"\nint main(void){ dfg_info i; (void)i; return DFG_OK; }\n'
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), "-x", "c", "-",
                        "-fsyntax-only"], input=src, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.parametrize("cfg", [G.reference_config(), G.hash_config(), G.plain_head_config(),
                                 G.hash_config(n_bins_theta=64, n_bins_phi=64)])
def test_configs_layouts_are_consistent(cfg):
    th, ph, nrc = O.config_layouts(cfg)
    assert th["enc_width"] % 16 == 0 and ph["enc_width"] % 16 == 0
    assert th["enc_width"] <= 128 and ph["enc_width"] <= 128
    if nrc:
        assert nrc["enc_width"] == 112 and nrc["n_out_pad"] == 16
