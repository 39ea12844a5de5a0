"""CPU-side checks of the C ABI library: it loads and exports every symbol
include/ipuma_hip.h declares.  No device calls."""
import ctypes
import subprocess

import ipuma_amd as I


def test_header_declares_entry_points():
    syms = I.header_symbols()
    for name in ("ipuma_engine_create", "ipuma_engine_destroy", "ipuma_compare_local",
                 "ipuma_compare_mn_local", "ipuma_compare_device", "ipuma_last_error"):
        assert name in syms


def test_library_exports_every_header_symbol():
    lib = I.lib()
    missing = [s for s in I.header_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", I.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert set(I.header_symbols()) <= exported


def test_default_configs_match_reference():
    lib = I.lib()
    sw = I._SWConfig()
    al = I._AlgoConfig()
    lib.ipuma_default_sw_config(ctypes.byref(sw))
    lib.ipuma_default_algo_config(ctypes.byref(al))
    # src/ipu_config.h:11-31
    assert (sw.gap_init, sw.gap_extend, sw.match_value, sw.mismatch_value,
            sw.ambiguity_value) == (0, -1, 1, -1, -1)
    assert (al.tiles_used, al.max_ab, al.max_batches, al.bufsize) == (1, 300, 20, 3000)


def test_null_engine_is_an_error_not_a_crash():
    lib = I.lib()
    assert lib.ipuma_set_batching(None, 0) == -1
    assert lib.ipuma_synchronize(None) == -1
    assert b"NULL" in lib.ipuma_last_error()
