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


def _unbound_engine():
    # validation runs before the C call; a NULL handle makes any call that
    # slipped through return EINVAL instead of touching a device
    eng = I.SWAlgorithm.__new__(I.SWAlgorithm)
    eng._h = None
    return eng


def test_device_entries_refuse_host_memory():
    import numpy as np
    import pytest
    import torch
    eng = _unbound_engine()
    n = 4
    cpu = dict(codes=torch.zeros(64, dtype=torch.uint8), a_off=torch.zeros(n, dtype=torch.int64),
               a_len=torch.ones(n, dtype=torch.int32), b_off=torch.zeros(n, dtype=torch.int64),
               b_len=torch.ones(n, dtype=torch.int32))
    outs = [torch.zeros(n, dtype=torch.int32) for _ in range(3)]
    with pytest.raises(TypeError, match="not on the GPU"):
        eng.compare_device(cpu["codes"], 64, cpu["a_off"], cpu["a_len"], cpu["b_off"],
                           cpu["b_len"], *outs)
    with pytest.raises(TypeError, match="numpy"):
        eng.compare_device(np.zeros(64, np.uint8), 64, cpu["a_off"], cpu["a_len"], cpu["b_off"],
                           cpu["b_len"], *outs)
    with pytest.raises(TypeError, match="numpy"):
        eng.search_device(1, 64, 2, 3, 4, 5, 2, None, np.zeros(8, np.int32),
                          np.zeros(8, np.int32), nq=4, ndb=4)
    with pytest.raises(TypeError, match="raw pointer carries no length"):
        eng.compare_device(1, 64, 2, 3, 4, 5, 6, 7, 8)
    with pytest.raises(TypeError, match="not on the GPU"):
        eng.encode_device(torch.zeros(8, dtype=torch.uint8), 8, 1)


def test_device_argument_checks():
    import pytest
    import torch

    class FakeCuda:
        """Stands in for a CUDA tensor in the checks (never reaches the C call)."""
        is_cuda = True
        device = "cuda:0"

        def __init__(self, dtype, numel, contiguous=True):
            self.dtype, self._n, self._c = dtype, numel, contiguous

        def numel(self):
            return self._n

        def is_contiguous(self):
            return self._c

        def data_ptr(self):
            return 0x1000

    assert I._dev(FakeCuda(torch.int32, 5), "x", "int32", numel=5) == 0x1000
    assert I._dev(1234, "x", "int32") == 1234
    assert I._dev(None, "x", "int32", optional=True) is None
    with pytest.raises(TypeError, match="dtype int64"):
        I._dev(FakeCuda(torch.int64, 5), "a_len", "int32")  # an int64 length array
    with pytest.raises(ValueError, match="contiguous"):
        I._dev(FakeCuda(torch.int32, 5, contiguous=False), "x", "int32")
    with pytest.raises(ValueError, match="expected 4"):
        I._dev(FakeCuda(torch.int64, 5), "a_off", "int64", numel=4)
    with pytest.raises(ValueError, match="at least 6"):
        I._dev(FakeCuda(torch.int32, 5), "scores", "int32", min_numel=6)
    with pytest.raises(ValueError, match="required"):
        I._dev(None, "top_idx", "int32")
    eng = _unbound_engine()
    n = 3
    good = dict(codes=FakeCuda(torch.uint8, 64), a_off=FakeCuda(torch.int64, n),
                a_len=FakeCuda(torch.int32, n), b_off=FakeCuda(torch.int64, n),
                b_len=FakeCuda(torch.int32, n))
    with pytest.raises(ValueError, match="top_idx"):  # nq * k = 2 * 5 outputs needed
        eng.search_device(good["codes"], 64, FakeCuda(torch.int64, 2), FakeCuda(torch.int32, 2),
                          good["b_off"], good["b_len"], 5, None, FakeCuda(torch.int32, 9),
                          FakeCuda(torch.int32, 10))
    with pytest.raises(ValueError, match="codes"):  # codes shorter than codes_len
        eng.compare_device(good["codes"], 65, good["a_off"], good["a_len"], good["b_off"],
                           good["b_len"], *[FakeCuda(torch.int32, n) for _ in range(3)])
    # everything valid: the NULL handle is refused by the C ABI itself
    with pytest.raises(I.IpumaError):
        eng.compare_device(good["codes"], 64, good["a_off"], good["a_len"], good["b_off"],
                           good["b_len"], *[FakeCuda(torch.int32, n) for _ in range(3)])
