"""The serving load generator's request pools (host only): every window hands
the server one byte per residue, the characters of the generator batch's own
sequences, with n + 1 offsets."""
import ctypes

import numpy as np

from ipuma_amd import loadgen as LG
from ipuma_amd import workloads as W


def test_pool_windows_carry_the_batch_sequences():
    pools = LG.make_pools(pairs=200, seed=3)
    for kind, (pool, _) in pools.items():
        protein = kind == "protein"
        b = pool.batch
        assert pool.a_chars.dtype == np.uint8 and pool.b_chars.dtype == np.uint8
        start, n = 17, 40
        k, ap, ao, bp, bo, cells = pool.window(start, n)
        assert k == n
        a_offs = np.ctypeslib.as_array(ctypes.cast(ao, ctypes.POINTER(ctypes.c_int64)), (n + 1,))
        b_offs = np.ctypeslib.as_array(ctypes.cast(bo, ctypes.POINTER(ctypes.c_int64)), (n + 1,))
        for j in range(n):
            p = start + j
            for chars_ptr, offs, side in ((ap, a_offs, "a"), (bp, b_offs, "b")):
                ln = int(offs[j + 1] - offs[j])
                got = ctypes.string_at(chars_ptr + int(offs[j]), ln).decode()
                o, l = int(b[f"{side}_off"][p]), int(b[f"{side}_len"][p])
                assert got == W.decode(b["codes"][o:o + l], protein), (kind, p, side)
        assert cells == int((b["a_len"][start:start + n].astype(np.int64) *
                             b["b_len"][start:start + n]).sum())
