"""The C-ABI library loads and exports every entry point of include/cdet_hip.h
(no compute call: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

import cdet_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


HIP_FROM_DRIVER_HEADER = ("cdet_eval_batch_host", "cdet_ctx_set_batch_mode", "cdet_seed_cond_batch_host",
                          "cdet_eval_batch_submit", "cdet_eval_batch_wait", "cdet_ctx_set_pair_reuse",
                          "cdet_host_alloc", "cdet_host_free")


def functions_in(header):
    text = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", header)).read(), flags=re.S)
    return set(re.findall(r"\b(cdet_[a-z0-9_]+)\s*\(", text))


def declared_functions():
    names = functions_in("cdet_hip.h") | functions_in("cdet_chunks.h")
    names |= set(HIP_FROM_DRIVER_HEADER)
    return sorted(names)


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
    names = functions_in("cdet_mcmcmc.h") - set(HIP_FROM_DRIVER_HEADER) - {"cdet_batch_eval_fn", "cdet_batch_cond_fn"}
    assert len(names) >= 9
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layouts_match_headers(tmp_path):
    """sizeof/offsetof of every ctypes mirror against the C compiler's."""
    import subprocess
    import mcmcmc
    mirrors = {"cdet_params": cdet_amd.CdetParams, "cdet_lattice": cdet_amd.CdetLattice,
               "cdet_chunk_spec": cdet_amd.ChunkSpec, "cdet_mcmcmc_config": mcmcmc.MCMCMCConfig,
               "cdet_mcmcmc_chain_result": mcmcmc.ChainResult, "cdet_histogram_seed": mcmcmc.HistogramSeed,
               "cdet_mcmcmc_backend": mcmcmc.Backend}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "cdet_chunks.h"', '#include "cdet_mcmcmc.h"',
             "int main(void) {"]
    for cname, cls in mirrors.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in cls._fields_:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                       check=True).stdout.splitlines())
    for cname, cls in mirrors.items():
        assert int(got[cname]) == ctypes.sizeof(cls), cname
        for f, _ in cls._fields_:
            assert int(got[f"{cname}.{f}"]) == getattr(cls, f).offset, (cname, f)
