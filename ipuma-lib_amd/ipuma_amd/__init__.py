"""Python binding of the MI355X Smith-Waterman engine (ctypes over the C ABI).

The C ABI (include/ipuma_hip.h) is the product boundary; this module only
marshals arguments.  It mirrors the reference's ``ipu::batchaffine::SWAlgorithm``
surface (src/ipu_batch_affine.h:52-118): ``SWAlgorithm(sw_config, algo_config)``,
``compare_local(A, B)``, ``compare_mn_local(seqs, cmps)``, ``get_result()``.
Loading fails loudly when the HIP library has not been built; there is no CPU
fallback.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass, field
from functools import lru_cache

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libipuma_amd.so")
HEADER = os.path.join(os.path.dirname(PKG_ROOT), "include", "ipuma_hip.h")

IPUMA_OK = 0
IPUMA_EINVAL, IPUMA_ENOMEM, IPUMA_EDEVICE, IPUMA_ECONFIG = -1, -2, -3, -4
IPUMA_ETOOLONG, IPUMA_EBUSY, IPUMA_ECLOSED = -5, -6, -7
ERRORS = {-1: "EINVAL", -2: "ENOMEM", -3: "EDEVICE", -4: "ECONFIG", -5: "ETOOLONG",
          -6: "EBUSY", -7: "ECLOSED"}

SIMILARITY = {"na": 0, "blosum50": 1, "blosum62": 2, "simple": 3}
DATATYPE = {"na": 0, "aa": 1, "str": 2}
VERTEX_TYPE = {"cpp": 0, "assembly": 1, "multi": 2, "multiasm": 3}
FILL_ALGO = {"fillfirst": 0, "roundrobin": 1, "greedy": 2}
BATCH_BUCKETED, BATCH_PLAIN, BATCH_PLAIN_TILE = 0, 1, 2
PREC_AUTO, PREC_INT32 = 0, 1


class IpumaError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class _SWConfig(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in (
        "gap_init", "gap_extend", "match_value", "mismatch_value", "ambiguity_value",
        "similarity", "datatype")]


class _AlgoConfig(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in (
        "tiles_used", "max_ab", "max_batches", "bufsize", "vtype", "fill_algo", "forward_only",
        "use_remote_buffer", "transmission_programs", "io_tiles")]


class BatchStats(ctypes.Structure):
    _fields_ = [("pairs", ctypes.c_int64), ("cells", ctypes.c_int64),
                ("padded_cells", ctypes.c_int64), ("groups", ctypes.c_int64),
                ("rerun_pairs", ctypes.c_int64), ("lanes_per_pair_max", ctypes.c_int32),
                ("rows_per_lane", ctypes.c_int32), ("band_groups", ctypes.c_int32),
                ("head_groups", ctypes.c_int32), ("fast_kernel_ms", ctypes.c_double),
                ("batch_ms", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class _ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_int64), ("max_batch_pairs", ctypes.c_int32),
                ("max_wait_us", ctypes.c_int32), ("max_inflight", ctypes.c_int32),
                ("block_on_full", ctypes.c_int32), ("merge_requests", ctypes.c_int32)]


class ServerStats(ctypes.Structure):
    _fields_ = [("requests", ctypes.c_int64), ("batches", ctypes.c_int64),
                ("pairs", ctypes.c_int64), ("cells", ctypes.c_int64),
                ("max_batch_pairs_seen", ctypes.c_int64), ("busy_ms", ctypes.c_double),
                ("arena_used_bytes", ctypes.c_int64), ("pending_pairs", ctypes.c_int64),
                ("span_ms", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


@dataclass
class SWConfig:
    """ipu::SWConfig (src/ipu_config.h:11-19) with the reference defaults."""
    gapInit: int = 0
    gapExtend: int = -1
    matchValue: int = 1
    mismatchValue: int = -1
    ambiguityValue: int = -1
    similarity: str = "na"
    datatype: str = "na"

    def _c(self):
        return _SWConfig(self.gapInit, self.gapExtend, self.matchValue, self.mismatchValue,
                         self.ambiguityValue, SIMILARITY[self.similarity],
                         DATATYPE[self.datatype])


@dataclass
class IPUAlgoConfig:
    """ipu::IPUAlgoConfig (src/ipu_config.h:21-31) with the reference defaults."""
    tilesUsed: int = 1
    maxAB: int = 300
    maxBatches: int = 20
    bufsize: int = 3000
    vtype: str = "cpp"
    fillAlgo: str = "fillfirst"
    forwardOnly: bool = False
    useRemoteBuffer: bool = False
    transmissionPrograms: int = 1
    ioTiles: int = 0

    def _c(self):
        return _AlgoConfig(self.tilesUsed, self.maxAB, self.maxBatches, self.bufsize,
                           VERTEX_TYPE[self.vtype], FILL_ALGO[self.fillAlgo],
                           int(self.forwardOnly), int(self.useRemoteBuffer),
                           self.transmissionPrograms, self.ioTiles)


def build() -> None:
    """Compile the HIP library for gfx950 (in-tree)."""
    subprocess.run(["make", "-s", "-j8", "-C", PKG_ROOT], check=True)


_P = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64


@lru_cache(maxsize=1)
def lib():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `make -C {PKG_ROOT}` "
                          "(there is no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    L.ipuma_last_error.restype = ctypes.c_char_p
    L.ipuma_engine_create.argtypes = [_P, _P, ctypes.POINTER(_P)]
    L.ipuma_engine_destroy.argtypes = [_P]
    L.ipuma_engine_destroy.restype = None
    L.ipuma_compare_local.argtypes = [_P, _i32, _P, _P, _P, _P, _P, _P, _P]
    L.ipuma_compare_mn_local.argtypes = [_P, _i32, _P, _P, _i32, _P, _P, _P, _P]
    L.ipuma_compare_device.argtypes = [_P, _i32, _P, _i64, _P, _P, _P, _P, _P, _P, _P, _P]
    L.ipuma_compare_device_noalloc.argtypes = [_P, _i32, _P, _i64, _P, _P, _P, _P, _P, _P, _P,
                                               _P]
    L.ipuma_set_batching.argtypes = [_P, _i32]
    L.ipuma_set_precision.argtypes = [_P, _i32]
    L.ipuma_set_forward_only.argtypes = [_P, _i32]
    L.ipuma_set_lanes_per_pair.argtypes = [_P, _i32]
    L.ipuma_set_work_rounds.argtypes = [_P, _i32]
    L.ipuma_set_step_cap_scale.argtypes = [_P, ctypes.c_float]
    L.ipuma_set_timing.argtypes = [_P, _i32]
    L.ipuma_encode_device.argtypes = [_P, _P, _i64, _P, _P]
    L.ipuma_last_batch_stats.argtypes = [_P, ctypes.POINTER(BatchStats)]
    L.ipuma_synchronize.argtypes = [_P]
    L.ipuma_search_device.argtypes = [_P, _P, _i64, _i32, _P, _P, _i32, _P, _P, _i32, _P, _P, _P,
                                      _P]
    L.ipuma_default_server_config.argtypes = [_P]
    L.ipuma_default_server_config.restype = None
    L.ipuma_server_create.argtypes = [_P, _P, _P, ctypes.POINTER(_P)]
    L.ipuma_server_destroy.argtypes = [_P]
    L.ipuma_server_destroy.restype = None
    L.ipuma_server_submit.argtypes = [_P, _i32, _P, _P, _P, _P, ctypes.POINTER(_i64)]
    L.ipuma_server_wait.argtypes = [_P, _i64, _P, _P, _P, ctypes.POINTER(ctypes.c_double)]
    L.ipuma_server_get_stats.argtypes = [_P, ctypes.POINTER(ServerStats)]
    L.ipuma_server_engine.argtypes = [_P, ctypes.POINTER(_P)]
    L.ipuma_input_buffer_words.argtypes = [_P]
    L.ipuma_input_buffer_words.restype = _i64
    L.ipuma_prepare_remote.argtypes = [_P, _P, _i32, _P, _P, _P, _P, _P, _i64, _P, _P]
    L.ipuma_prepare_remote_mn.argtypes = [_P, _P, _i32, _P, _P, _i32, _P, _P, _i64, _P, _P]
    L.ipuma_transfer_results.argtypes = [_P, _i64, _P, _i32, _P, _P, _P]
    L.ipuma_prepared_compare.argtypes = [_P, _P, _i64, _P, _i64, _i32]
    L.ipuma_prepared_submit.argtypes = [_P, _P, _i64, _P, _i64, ctypes.POINTER(_i64)]
    L.ipuma_prepared_join.argtypes = [_P, _i64]
    L.ipuma_traceback_local.argtypes = [_P, _i32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]
    L.ipuma_last_traceback_stats.argtypes = [_P, _P, _P]
    return L


def _check(rc: int):
    if rc != IPUMA_OK:
        raise IpumaError(rc, lib().ipuma_last_error().decode(errors="replace"))


def header_symbols(path: str = HEADER):
    """Function names the C ABI header declares."""
    import re
    text = open(path).read()
    return sorted(set(re.findall(
        r"^\s*(?:int64_t|int|void|const char \*)\s*\*?\s*(ipuma_\w+)\s*\(", text, re.M)))


def _pack(seqs):
    enc = [s.encode("latin-1") if isinstance(s, str) else bytes(s) for s in seqs]
    offs = np.zeros(len(enc) + 1, np.int64)
    if enc:
        offs[1:] = np.cumsum([len(x) for x in enc])
    buf = b"".join(enc) + b"\0"
    return buf, offs


# --- prepared input buffers (the reference's wire format) --------------------
def input_buffer_words(algo: IPUAlgoConfig) -> int:
    """IPUAlgoConfig::getInputBufferSize32b (src/ipu_config.cpp:14)."""
    a = algo._c()
    return int(lib().ipuma_input_buffer_words(ctypes.byref(a)))


def _i32buf(x, n=None):
    if x is None:
        return np.zeros(n, np.int32)
    if not (isinstance(x, np.ndarray) and x.dtype == np.int32 and x.flags.c_contiguous):
        raise TypeError("expected a contiguous numpy int32 array")
    return x


def prepare_remote(sw: SWConfig, algo: IPUAlgoConfig, A, B, inputs=None):
    """SWAlgorithm::prepare_remote (src/ipu_batch_affine.cpp:891-950): the input
    buffer (numpy int32, input_buffer_words long unless given), the slot of
    every comparison and the largest tile fill."""
    if len(A) != len(B):
        raise ValueError("A and B must have the same number of sequences")
    inputs = _i32buf(inputs, input_buffer_words(algo))
    mapping = np.zeros(len(A), np.int32)
    abuf, aoff = _pack(A)
    bbuf, boff = _pack(B)
    mb = ctypes.c_int32(0)
    s, a = sw._c(), algo._c()
    _check(lib().ipuma_prepare_remote(ctypes.byref(s), ctypes.byref(a), len(A), abuf,
                                      aoff.ctypes.data, bbuf, boff.ctypes.data,
                                      inputs.ctypes.data, len(inputs), mapping.ctypes.data,
                                      ctypes.byref(mb)))
    return inputs, mapping, mb.value


def prepare_remote_mn(sw: SWConfig, algo: IPUAlgoConfig, seqs, cmps, inputs=None):
    """compare_mn_local's preparation (fillMNBuckets + fill_input_buffer)."""
    inputs = _i32buf(inputs, input_buffer_words(algo))
    c = np.ascontiguousarray(np.asarray(cmps, np.int32).reshape(-1, 2))
    mapping = np.zeros(len(c), np.int32)
    buf, offs = _pack(seqs)
    mb = ctypes.c_int32(0)
    s, a = sw._c(), algo._c()
    _check(lib().ipuma_prepare_remote_mn(ctypes.byref(s), ctypes.byref(a), len(seqs), buf,
                                         offs.ctypes.data, len(c), c.ctypes.data,
                                         inputs.ctypes.data, len(inputs), mapping.ctypes.data,
                                         ctypes.byref(mb)))
    return inputs, mapping, mb.value


def transfer_results(results, mapping):
    """SWAlgorithm::transferResults (src/ipu_batch_affine.cpp:674-695)."""
    results = _i32buf(results)
    mapping = np.ascontiguousarray(mapping, np.int32)
    n = len(mapping)
    out = [np.zeros(n, np.int32) for _ in range(3)]
    _check(lib().ipuma_transfer_results(results.ctypes.data, len(results), mapping.ctypes.data,
                                        n, *[o.ctypes.data for o in out]))
    return tuple(out)


class PreparedJob:
    """A submitted prepared buffer (src/ipu_batch_affine.h:41-48); holds its
    buffers until joined."""

    def __init__(self, engine, job_id, inputs, results):
        self._engine, self.id, self.inputs, self.results = engine, job_id, inputs, results
        self.joined = False

    def join(self):
        if not self.joined:
            self.joined = True
            _check(lib().ipuma_prepared_join(self._engine._h, self.id))
        return self.results


def _count(x, name, given=None) -> int:
    """Number of elements of a length array given to a device entry: the
    tensor's, or `given` where the array is a raw pointer."""
    if given is not None:
        if hasattr(x, "numel") and int(x.numel()) != int(given):
            raise ValueError(f"{name}: {int(x.numel())} elements, expected {int(given)}")
        return int(given)
    if hasattr(x, "numel"):
        return int(x.numel())
    raise TypeError(f"{name}: a raw pointer carries no length; pass a tensor or the count")


def _dev(x, name, dtype, numel=None, min_numel=None, optional=False):
    """Device address of an argument of a device entry point.  A CUDA tensor
    is checked for dtype, contiguity and size; an int passes as a raw device
    pointer.  Host arrays (numpy, CPU tensors) are refused: the kernels would
    dereference a host address."""
    if x is None:
        if optional:
            return None
        raise ValueError(f"{name}: required")
    if isinstance(x, bool):
        raise TypeError(f"{name}: expected a CUDA tensor or a device pointer")
    if isinstance(x, int):
        return x
    if isinstance(x, np.ndarray):
        raise TypeError(f"{name}: a numpy array is host memory; pass a CUDA tensor")
    if not hasattr(x, "data_ptr") or not hasattr(x, "is_cuda"):
        raise TypeError(f"{name}: expected a CUDA tensor or a device pointer, got {type(x)}")
    if not x.is_cuda:
        raise TypeError(f"{name}: tensor is on {x.device}, not on the GPU")
    want = (dtype,) if isinstance(dtype, str) else tuple(dtype)
    got = str(x.dtype).replace("torch.", "")
    if got not in want:
        raise TypeError(f"{name}: dtype {got}, expected {' or '.join(want)}")
    if not x.is_contiguous():
        raise ValueError(f"{name}: tensor must be contiguous")
    if numel is not None and int(x.numel()) != int(numel):
        raise ValueError(f"{name}: {int(x.numel())} elements, expected {int(numel)}")
    if min_numel is not None and int(x.numel()) < int(min_numel):
        raise ValueError(f"{name}: {int(x.numel())} elements, at least {int(min_numel)} needed")
    return x.data_ptr()


@dataclass
class BlockAlignmentResults:
    """ipu::batchaffine::BlockAlignmentResults (src/ipu_batch_affine.h:22-26)."""
    scores: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    a_range_result: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    b_range_result: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))


class SWAlgorithm:
    """Engine handle with the reference's SWAlgorithm calling convention."""

    def __init__(self, config: SWConfig | None = None, algoconfig: IPUAlgoConfig | None = None):
        self.config = config or SWConfig()
        self.algoconfig = algoconfig or IPUAlgoConfig()
        h = _P()
        sw = self.config._c()
        al = self.algoconfig._c()
        _check(lib().ipuma_engine_create(ctypes.byref(sw), ctypes.byref(al), ctypes.byref(h)))
        self._h = h
        self.slot_cap = 1  # remote-buffer slots per layout lot (slotToken = lot * slot_cap + slot)
        n = self.algoconfig.tilesUsed * self.algoconfig.maxBatches
        self._result = BlockAlignmentResults(np.zeros(n, np.int32), np.zeros(n, np.int32),
                                             np.zeros(n, np.int32))

    def close(self):
        if getattr(self, "_h", None):
            lib().ipuma_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # --- policies -----------------------------------------------------------
    def set_batching(self, mode: int):
        _check(lib().ipuma_set_batching(self._h, mode))

    def set_precision(self, mode: int):
        _check(lib().ipuma_set_precision(self._h, mode))

    def set_forward_only(self, flag: bool):
        _check(lib().ipuma_set_forward_only(self._h, int(bool(flag))))

    def set_timing(self, on: bool):
        _check(lib().ipuma_set_timing(self._h, int(bool(on))))

    def set_lanes_per_pair(self, lanes: int):
        _check(lib().ipuma_set_lanes_per_pair(self._h, int(lanes)))

    def set_work_rounds(self, rounds: int):
        """Groups per resident wave the length bucketing aims for (1..64)."""
        _check(lib().ipuma_set_work_rounds(self._h, int(rounds)))

    def set_step_cap_scale(self, scale: float):
        """Scale of the bucketing's step cap per group (0.25..4, default 1)."""
        _check(lib().ipuma_set_step_cap_scale(self._h, float(scale)))

    # --- reference surface --------------------------------------------------
    def _store(self, s, a, b):
        n = len(s)
        cap = max(len(self._result.scores), n)
        for name, v in (("scores", s), ("a_range_result", a), ("b_range_result", b)):
            arr = getattr(self._result, name)
            if len(arr) < cap:
                arr = np.concatenate([arr, np.zeros(cap - len(arr), np.int32)])
            arr[:n] = v
            setattr(self._result, name, arr)

    def compare_local(self, A, B, errcheck: bool = False):
        """compare_local (src/ipu_batch_affine.cpp:564-663)."""
        if len(A) != len(B):
            raise ValueError("A and B must have the same number of sequences")
        s, a, b = self.compare(A, B)
        self._store(s, a, b)

    def compare_mn_local(self, seqs, cmps, errcheck: bool = False):
        """compare_mn_local (src/ipu_batch_affine.cpp:697-750)."""
        s, a, b = self.compare_mn(seqs, cmps)
        self._store(s, a, b)

    def prepared_remote_compare(self, inputs, results=None, slot_token: int = 0):
        """prepared_remote_compare (src/ipu_batch_affine.cpp:557-562): runs a
        prepared buffer; returns the results buffer (3 x tilesUsed*maxBatches
        words unless given: [scores | a ranges | b ranges])."""
        inputs = _i32buf(inputs)
        n = self.algoconfig.tilesUsed * self.algoconfig.maxBatches
        results = _i32buf(results, 3 * n)
        lot = int(slot_token) // self.slot_cap
        _check(lib().ipuma_prepared_compare(self._h, inputs.ctypes.data, len(inputs),
                                            results.ctypes.data, len(results), lot))
        return results

    def async_submit_prepared_remote_compare(self, inputs, results=None) -> PreparedJob:
        """async_submit_prepared_remote_compare (src/ipu_batch_affine.cpp:496-505)."""
        inputs = _i32buf(inputs)
        n = self.algoconfig.tilesUsed * self.algoconfig.maxBatches
        results = _i32buf(results, 3 * n)
        job = ctypes.c_int64(0)
        _check(lib().ipuma_prepared_submit(self._h, inputs.ctypes.data, len(inputs),
                                           results.ctypes.data, len(results),
                                           ctypes.byref(job)))
        return PreparedJob(self, job.value, inputs, results)

    def blocking_join_prepared_remote_compare(self, job: PreparedJob):
        """blocking_join_prepared_remote_compare (src/ipu_batch_affine.cpp:507-555)."""
        return job.join()

    def traceback(self, A, B, scores, a_range, b_range):
        """CIGAR runs (length << 4 | op; 0 M, 1 I = A only, 2 D = B only) of each
        comparison's alignment, from the ranges of a compare with starts."""
        n = len(A)
        abuf, aoff = _pack(A)
        bbuf, boff = _pack(B)
        sc = np.ascontiguousarray(scores, np.int32)
        ar = np.ascontiguousarray(a_range, np.int32)
        br = np.ascontiguousarray(b_range, np.int32)
        u = ar.astype(np.int64) & 0xFFFFFFFF
        v = br.astype(np.int64) & 0xFFFFFFFF
        cap = np.where(sc > 0, ((u >> 16) - (u & 0xFFFF) + 1) + ((v >> 16) - (v & 0xFFFF) + 1), 0)
        offs = np.zeros(n + 1, np.int64)
        offs[1:] = np.cumsum(cap)
        cig = np.zeros(max(int(offs[-1]), 1), np.uint32)
        lens = np.zeros(n, np.int32)
        _check(lib().ipuma_traceback_local(self._h, n, abuf, aoff.ctypes.data, bbuf,
                                           boff.ctypes.data, sc.ctypes.data, ar.ctypes.data,
                                           br.ctypes.data, cig.ctypes.data, offs.ctypes.data,
                                           lens.ctypes.data))
        return [cig[offs[k]:offs[k] + lens[k]].copy() for k in range(n)]

    def traceback_stats(self) -> dict:
        """Region cells of the last traceback call and, with timing on, the time
        of its traceback kernels in ms (HIP events on the engine's stream)."""
        cells = ctypes.c_int64(0)
        ms = ctypes.c_double(0.0)
        _check(lib().ipuma_last_traceback_stats(self._h, ctypes.byref(cells), ctypes.byref(ms)))
        return {"region_cells": int(cells.value), "kernel_ms": float(ms.value)}

    def get_result(self) -> BlockAlignmentResults:
        """get_result (src/ipu_batch_affine.cpp:408): copies sized tilesUsed*maxBatches."""
        r = self._result
        return BlockAlignmentResults(r.scores.copy(), r.a_range_result.copy(),
                                     r.b_range_result.copy())

    # --- array API ----------------------------------------------------------
    def compare(self, A, B):
        n = len(A)
        abuf, aoff = _pack(A)
        bbuf, boff = _pack(B)
        s = np.zeros(n, np.int32)
        ar = np.zeros(n, np.int32)
        br = np.zeros(n, np.int32)
        _check(lib().ipuma_compare_local(
            self._h, n, abuf, aoff.ctypes.data, bbuf, boff.ctypes.data, s.ctypes.data,
            ar.ctypes.data, br.ctypes.data))
        return s, ar, br

    def compare_mn(self, seqs, cmps):
        buf, offs = _pack(seqs)
        c = np.ascontiguousarray(np.asarray(cmps, dtype=np.int32).reshape(-1, 2))
        n = c.shape[0]
        s = np.zeros(n, np.int32)
        ar = np.zeros(n, np.int32)
        br = np.zeros(n, np.int32)
        _check(lib().ipuma_compare_mn_local(
            self._h, len(seqs), buf, offs.ctypes.data, n, c.ctypes.data, s.ctypes.data,
            ar.ctypes.data, br.ctypes.data))
        return s, ar, br

    def compare_device(self, codes, codes_len, a_off, a_len, b_off, b_len, scores, a_range,
                       b_range, stream=None, n=None):
        """Device batch: every argument is a CUDA tensor (uint8 codes, int64
        offsets, int32 lengths and outputs, contiguous) or a raw device pointer
        (int), which the caller vouches for; n is needed when a_len is one."""
        n = _count(a_len, "a_len", n)
        _check(lib().ipuma_compare_device(
            self._h, n, _dev(codes, "codes", "uint8", min_numel=codes_len), int(codes_len),
            _dev(a_off, "a_off", "int64", numel=n), _dev(a_len, "a_len", "int32", numel=n),
            _dev(b_off, "b_off", "int64", numel=n), _dev(b_len, "b_len", "int32", numel=n),
            _dev(scores, "scores", "int32", min_numel=n),
            _dev(a_range, "a_range", "int32", min_numel=n),
            _dev(b_range, "b_range", "int32", min_numel=n), stream))

    def compare_device_noalloc(self, codes, codes_len, a_off, a_len, b_off, b_len, scores,
                               a_range, b_range, stream=None, n=None):
        """compare_device without device allocation (ipuma_compare_device_noalloc):
        forward-only engines, n <= 65536, |A| <= 4096, work space in a static
        array per GPU shared by every call on it."""
        n = _count(a_len, "a_len", n)
        _check(lib().ipuma_compare_device_noalloc(
            self._h, n, _dev(codes, "codes", "uint8", min_numel=codes_len), int(codes_len),
            _dev(a_off, "a_off", "int64", numel=n), _dev(a_len, "a_len", "int32", numel=n),
            _dev(b_off, "b_off", "int64", numel=n), _dev(b_len, "b_len", "int32", numel=n),
            _dev(scores, "scores", "int32", min_numel=n),
            _dev(a_range, "a_range", "int32", min_numel=n),
            _dev(b_range, "b_range", "int32", min_numel=n), stream))

    def search_device(self, codes, codes_len, q_off, q_len, db_off, db_len, k, scores,
                      top_idx, top_score, stream=None, nq=None, ndb=None):
        """Every query against every database sequence (CUDA tensors or raw
        device pointers, as for compare_device; nq / ndb are needed when the
        length arrays are raw pointers); scores may be None (only the top k per
        query is kept)."""
        nq = _count(q_len, "q_len", nq)
        ndb = _count(db_len, "db_len", ndb)
        _check(lib().ipuma_search_device(
            self._h, _dev(codes, "codes", "uint8", min_numel=codes_len), int(codes_len), nq,
            _dev(q_off, "q_off", "int64", numel=nq), _dev(q_len, "q_len", "int32", numel=nq), ndb,
            _dev(db_off, "db_off", "int64", numel=ndb),
            _dev(db_len, "db_len", "int32", numel=ndb), int(k),
            _dev(scores, "scores", "int32", min_numel=nq * ndb, optional=True),
            _dev(top_idx, "top_idx", "int32", min_numel=nq * int(k)),
            _dev(top_score, "top_score", "int32", min_numel=nq * int(k)), stream))

    def encode_device(self, chars, length, codes, stream=None):
        """Characters to residue codes on the device (CUDA uint8/int8 tensors or
        raw device pointers)."""
        _check(lib().ipuma_encode_device(
            self._h, _dev(chars, "chars", ("uint8", "int8"), min_numel=length), int(length),
            _dev(codes, "codes", "uint8", min_numel=length), stream))

    def stats(self) -> dict:
        st = BatchStats()
        _check(lib().ipuma_last_batch_stats(self._h, ctypes.byref(st)))
        return st.as_dict()

    def synchronize(self):
        _check(lib().ipuma_synchronize(self._h))


def cigar_string(runs) -> str:
    """'12M1I3M' from CIGAR runs."""
    return "".join(f"{int(r) >> 4}{'MID'[int(r) & 15]}" for r in runs)


def unpack_range(x):
    """(begin, end) the way the reference tests unpack (test/simple-test.cpp:64-70)."""
    x = np.asarray(x).astype(np.int64) & 0xFFFFFFFF
    begin = (x & 0xFFFF).astype(np.uint16).astype(np.int16)
    end = (x >> 16).astype(np.uint16).astype(np.int16)
    return begin, end


class Server:
    """Request server (replaces the reference's IPUMultiDriver batch loop,
    ipusw/ipumultidriver.hpp:10-145, and its async submit / join,
    src/ipu_batch_affine.h:105-108): submit() from any thread returns a
    ticket, wait(ticket) the results and the latency in ms."""

    def __init__(self, config: SWConfig = None, algoconfig: IPUAlgoConfig = None,
                 hbm_budget_bytes: int = 1 << 30, max_batch_pairs: int = 1 << 16,
                 max_wait_us: int = 2000, max_inflight: int = 2, block_on_full: bool = True,
                 merge_requests: bool = True):
        config = config or SWConfig()
        algoconfig = algoconfig or IPUAlgoConfig()
        sc = _ServerConfig(int(hbm_budget_bytes), int(max_batch_pairs), int(max_wait_us),
                           int(max_inflight), int(block_on_full), int(merge_requests))
        h = _P()
        sw, al = config._c(), algoconfig._c()
        _check(lib().ipuma_server_create(ctypes.byref(sw), ctypes.byref(al), ctypes.byref(sc),
                                         ctypes.byref(h)))
        self._h = h
        self.max_batch_pairs = int(max_batch_pairs)

    def submit(self, A, B) -> int:
        if len(A) != len(B):
            raise ValueError("A and B must have the same number of sequences")
        abuf, aoff = _pack(A)
        bbuf, boff = _pack(B)
        t = _i64()
        _check(lib().ipuma_server_submit(self._h, len(A), abuf, aoff.ctypes.data, bbuf,
                                         boff.ctypes.data, ctypes.byref(t)))
        return int(t.value)

    def submit_raw(self, n: int, a_chars: int, a_offsets: int, b_chars: int,
                   b_offsets: int) -> int:
        """Zero-copy submit: host pointers to characters and to n + 1 int64
        offsets per side (the caller keeps them alive until this returns)."""
        t = _i64()
        _check(lib().ipuma_server_submit(self._h, int(n), a_chars, a_offsets, b_chars,
                                         b_offsets, ctypes.byref(t)))
        return int(t.value)

    def wait(self, ticket: int, n: int):
        """(scores, a_range, b_range, latency_ms) of a request of n comparisons."""
        s = np.zeros(n, np.int32)
        ar = np.zeros(n, np.int32)
        br = np.zeros(n, np.int32)
        lat = ctypes.c_double()
        _check(lib().ipuma_server_wait(self._h, int(ticket), s.ctypes.data, ar.ctypes.data,
                                       br.ctypes.data, ctypes.byref(lat)))
        return s, ar, br, lat.value

    def stats(self) -> dict:
        st = ServerStats()
        _check(lib().ipuma_server_get_stats(self._h, ctypes.byref(st)))
        return st.as_dict()

    def close(self):
        if getattr(self, "_h", None):
            lib().ipuma_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
