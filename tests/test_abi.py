"""The C-ABI library loads and exports every entry point of include/cdet_hip.h
(no compute call: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

import cdet_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "cdet_hip.h")).read()
    text += "\n".join(l for l in open(os.path.join(ROOT, "include", "cdet_mcmcmc.h")).read().splitlines()
                      if "cdet_eval_batch_host" in l or "cdet_ctx_set_batch_mode" in l)
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cdet_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    if not os.path.exists(cdet_amd.LIB_PATH):
        pytest.fail("libcdet_hip.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(cdet_amd.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the host mirror binds exactly the declared set
    assert sorted(cdet_amd._signatures()) == names


def test_params_struct_layout_matches_header():
    assert ctypes.sizeof(cdet_amd.CdetParams) == 4 * 4 + 13 * 8
    assert cdet_amd.CdetParams.beta.offset == 16


def test_params_derive_is_host_only():
    p = cdet_amd.make_params(4, 4, 96, 6.0, 2.0, 1.0, 1.0, 1.0)
    assert p.scale == 192.0 ** 0.5
    assert p.diag_up_shift == 0.5 and p.diag_down_shift == 0.5


def test_driver_library_exports():
    import mcmcmc
    lib = ctypes.CDLL(mcmcmc.LIB_PATH)
    for name in ("cdet_mcmcmc_run", "cdet_mcmcmc_last_error"):
        assert hasattr(lib, name)
