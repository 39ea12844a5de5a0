"""Parity oracle for the batched Smith-Waterman affine-gap path.

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of bench.py may import this module, and only as the
checker or as the timed CPU baseline.  The product (ipuma-lib_amd) never
imports it.

Three independent checkers live here:

* ``liboracle.so`` (oracle/sw_oracle.c) -- a C restatement of the reference's
  SWAffine vertex (src/codelets/codelet_batch_affine.cpp:37-124) with its
  tables (src/swatlib/similarity.h:17-94) and encodings
  (src/swatlib/encoding.h:27-73); OpenMP over pairs.
* ``bruteforce_pair`` -- an exhaustive enumerator of alignments for tiny
  inputs.  It shares no code or recurrence with the DP: it walks every
  alignment path between every pair of substrings and applies the gap model
  of src/driver.hpp:47-49 (a gap of length k costs open + (k-1)*extend).
* ``libssw_ref.so`` (oracle/_ref, built from the reference's vendored SSW by
  oracle/Makefile) -- the reference's own CPU comparator
  (test/parity-test.cpp:70-87).  Present only where the reference tree is
  (built there, the library travels with the tree); bench.py's CPU baseline
  times it over a batch (``ssw_batch_codes``).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from functools import lru_cache

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libssw_ref.so")

SIM_NA, SIM_BLOSUM50, SIM_BLOSUM62, SIM_SIMPLE = 0, 1, 2, 3
DT_NA, DT_AA, DT_STRING = 0, 1, 2

_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_i8p = ctypes.POINTER(ctypes.c_int8)


def build(ref_root: str | None = None) -> None:
    """Compile liboracle.so (and oracle/_ref when the reference tree exists)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    root = ref_root or "/root/reference"
    if os.path.isdir(os.path.join(root, "extern", "include", "ssw")):
        subprocess.run(["make", "-s", "-C", HERE, "ref", f"REF_ROOT={root}"], check=True)


@lru_cache(maxsize=1)
def _lib():
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    lib.oracle_select_matrix.argtypes = [ctypes.c_int] * 4 + [_i8p]
    lib.oracle_select_matrix.restype = ctypes.c_int
    lib.oracle_encode.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_int64, _u8p]
    lib.oracle_sw_batch.argtypes = [
        ctypes.c_int64, _u8p, _i64p, _i32p, _i64p, _i32p, _i8p,
        ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
        _i32p, _i32p, _i32p]
    lib.oracle_sw_batch.restype = ctypes.c_int
    lib.oracle_max_threads.restype = ctypes.c_int
    lib.oracle_cigar.argtypes = [_u8p, _u8p] + [ctypes.c_int] * 4 + [
        ctypes.POINTER(ctypes.c_int8), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
        ctypes.POINTER(ctypes.c_uint32), ctypes.c_int]
    lib.oracle_cigar.restype = ctypes.c_int
    return lib


def have_ref() -> bool:
    return os.path.exists(REF_PATH)


@lru_cache(maxsize=1)
def _ref():
    lib = ctypes.CDLL(REF_PATH)
    lib.ssw_ref_align_dna.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p,
                                      ctypes.c_int] + [ctypes.c_int] * 5 + [_i32p]
    lib.ssw_ref_align_matrix.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p,
                                         ctypes.c_int, _i8p, ctypes.c_int, _i8p,
                                         ctypes.c_int, ctypes.c_int, _i32p]
    if hasattr(lib, "ssw_ref_align_matrix_cigar"):
        lib.ssw_ref_align_matrix_cigar.argtypes = [
            ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int, _i8p, ctypes.c_int,
            _i8p, ctypes.c_int, ctypes.c_int, _i32p, ctypes.c_char_p, ctypes.c_int]
    return lib


def _ptr(a: np.ndarray, t):
    return a.ctypes.data_as(t)


def ssw_batch_codes(codes, a_off, a_len, b_off, b_len, mat: np.ndarray, gap_open: int,
                    gap_extend: int, threads: int) -> np.ndarray:
    """The reference's CPU comparator (SSW through cpusw/aligners.hpp:128-152: a query
    profile of A, then ssw_align against B, end positions only) over a coded batch,
    OpenMP over pairs.  gap_open / gap_extend are SSW's positive penalties (the
    reference's gapInit = -2, gapExtend = -1 is SSW's 3 / 1).  Scores only."""
    lib = _ref()
    lib.ssw_ref_batch_codes.argtypes = [ctypes.c_int32, _u8p, _i64p, _i32p, _i64p, _i32p, _i8p,
                                        ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                        ctypes.c_int32, _i32p]
    lib.ssw_ref_batch_codes.restype = None
    n = len(a_len)
    codes = np.ascontiguousarray(codes, np.uint8)
    a_off = np.ascontiguousarray(a_off, np.int64)
    b_off = np.ascontiguousarray(b_off, np.int64)
    a_len = np.ascontiguousarray(a_len, np.int32)
    b_len = np.ascontiguousarray(b_len, np.int32)
    mat = np.ascontiguousarray(mat, np.int8)
    out = np.zeros(n, np.int32)
    lib.ssw_ref_batch_codes(n, _ptr(codes, _u8p), _ptr(a_off, _i64p), _ptr(a_len, _i32p),
                            _ptr(b_off, _i64p), _ptr(b_len, _i32p), _ptr(mat, _i8p),
                            mat.shape[0], int(gap_open), int(gap_extend), int(threads),
                            _ptr(out, _i32p))
    return out


def select_matrix(similarity: int, match: int = 1, mismatch: int = -1,
                  ambiguity: int = -1) -> np.ndarray:
    """selectMatrix (src/swatlib/similarity.cpp:89-104), as a dim x dim int8."""
    out = np.zeros(25 * 25, dtype=np.int8)
    dim = _lib().oracle_select_matrix(similarity, match, mismatch, ambiguity, _ptr(out, _i8p))
    if dim < 0:
        raise RuntimeError("No matrix for simple penalty available.")
    return out[: dim * dim].reshape(dim, dim).copy()


def encode(datatype: int, s: str | bytes) -> np.ndarray:
    """Encoding::encode (src/swatlib/encoding.cpp:53-62)."""
    b = s.encode("latin-1") if isinstance(s, str) else bytes(s)
    out = np.zeros(len(b), dtype=np.uint8)
    _lib().oracle_encode(datatype, b, len(b), _ptr(out, _u8p))
    return out


def arena(seqs_a, seqs_b, datatype: int):
    """Encode pairs into one arena: returns codes, a_off, a_len, b_off, b_len."""
    parts, a_off, a_len, b_off, b_len = [], [], [], [], []
    pos = 0
    for a, b in zip(seqs_a, seqs_b):
        ea, eb = encode(datatype, a), encode(datatype, b)
        a_off.append(pos); a_len.append(len(ea)); parts.append(ea); pos += len(ea)
        b_off.append(pos); b_len.append(len(eb)); parts.append(eb); pos += len(eb)
    codes = np.concatenate(parts) if parts else np.zeros(0, np.uint8)
    return (codes, np.array(a_off, np.int64), np.array(a_len, np.int32),
            np.array(b_off, np.int64), np.array(b_len, np.int32))


def sw_batch_codes(codes, a_off, a_len, b_off, b_len, matrix: np.ndarray, gap_init: int,
                   gap_extend: int, forward_only: bool = False, nthreads: int = 0):
    """Run the C oracle on an encoded arena; returns (scores, aranges, branges)."""
    n = len(a_len)
    codes = np.ascontiguousarray(codes, dtype=np.uint8)
    if codes.size == 0:
        codes = np.zeros(1, np.uint8)
    a_off = np.ascontiguousarray(a_off, dtype=np.int64)
    b_off = np.ascontiguousarray(b_off, dtype=np.int64)
    a_len = np.ascontiguousarray(a_len, dtype=np.int32)
    b_len = np.ascontiguousarray(b_len, dtype=np.int32)
    mat = np.ascontiguousarray(matrix, dtype=np.int8)
    scores = np.zeros(n, np.int32)
    ar = np.zeros(n, np.int32)
    br = np.zeros(n, np.int32)
    rc = _lib().oracle_sw_batch(n, _ptr(codes, _u8p), _ptr(a_off, _i64p), _ptr(a_len, _i32p),
                                _ptr(b_off, _i64p), _ptr(b_len, _i32p), _ptr(mat, _i8p),
                                mat.shape[0], gap_init, gap_extend, int(bool(forward_only)),
                                nthreads, _ptr(scores, _i32p), _ptr(ar, _i32p), _ptr(br, _i32p))
    if rc != 0:
        raise RuntimeError("oracle_sw_batch failed")
    return scores, ar, br


def sw_batch(seqs_a, seqs_b, similarity=SIM_NA, datatype=DT_NA, match=1, mismatch=-1,
             ambiguity=-1, gap_init=0, gap_extend=-1, forward_only=False, nthreads=0):
    """compare_local semantics on strings (src/ipu_batch_affine.cpp:564-663)."""
    mat = select_matrix(similarity, match, mismatch, ambiguity)
    codes, ao, al, bo, bl = arena(seqs_a, seqs_b, datatype)
    return sw_batch_codes(codes, ao, al, bo, bl, mat, gap_init, gap_extend, forward_only, nthreads)


def cigar(a_codes, b_codes, a_range: int, b_range: int, score: int, matrix: np.ndarray,
          gap_init: int, gap_extend: int):
    """oracle_cigar: runs (length << 4 | op), op 0 = M, 1 = I (A only), 2 = D (B only),
    of the alignment between the reported start and end positions."""
    a = np.ascontiguousarray(a_codes, np.uint8)
    b = np.ascontiguousarray(b_codes, np.uint8)
    ae, as_ = (a_range >> 16) & 0xFFFF, a_range & 0xFFFF
    be, bs = (b_range >> 16) & 0xFFFF, b_range & 0xFFFF
    if score <= 0:
        return np.zeros(0, np.uint32)
    cap = (ae - as_ + 1) + (be - bs + 1) + 1
    out = np.zeros(cap, np.uint32)
    m = np.ascontiguousarray(matrix, np.int8)
    n = _lib().oracle_cigar(_ptr(a, _u8p), _ptr(b, _u8p), as_, ae, bs, be,
                            m.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)), m.shape[0],
                            gap_init, gap_extend, score,
                            out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), cap)
    if n < 0:
        raise ValueError(f"oracle_cigar failed ({n})")
    return out[:n].copy()


def cigar_string(runs) -> str:
    return "".join(f"{int(r) >> 4}{'MID'[int(r) & 15]}" for r in runs)


def cigar_score(a_codes, b_codes, a_start: int, b_start: int, runs, matrix, gap_init, gap_extend):
    """Score of the alignment the runs describe (affine gaps)."""
    go, ge = -(gap_init + gap_extend), -gap_extend
    i, j, s = a_start, b_start, 0
    for r in runs:
        ln, op = int(r) >> 4, int(r) & 15
        if op == 0:
            for _ in range(ln):
                s += int(matrix[a_codes[i], b_codes[j]])
                i += 1
                j += 1
        else:
            s -= go + (ln - 1) * ge
            if op == 1:
                i += ln
            else:
                j += ln
    return s, i - 1, j - 1


def max_threads() -> int:
    return _lib().oracle_max_threads()


def unpack_range(x: int):
    """(begin, end) as the reference's tests unpack them (test/simple-test.cpp:64-70)."""
    x = int(x) & 0xFFFFFFFF
    begin = np.int16(np.uint16(x & 0xFFFF))
    end = np.int16(np.uint16(x >> 16))
    return int(begin), int(end)


# ----------------------------------------------------------------------------
# Exhaustive enumerator (tiny inputs only): independent of the DP.
# ----------------------------------------------------------------------------

def _best_global(a, b, mat, gopen, gext):
    """Max over every alignment of a vs b (both fully consumed) of its score.

    Enumerates alignment columns recursively; a run of k gap columns of one
    kind costs gopen + (k-1)*gext.  Returns -inf-ish for impossible cases.
    """
    best = -10 ** 9
    m, n = len(a), len(b)

    def rec(i, j, last, acc):
        nonlocal best
        if i == m and j == n:
            best = max(best, acc)
            return
        if i < m and j < n:
            rec(i + 1, j + 1, "M", acc + int(mat[a[i], b[j]]))
        if i < m:  # a[i] against a gap
            rec(i + 1, j, "X", acc - (gext if last == "X" else gopen))
        if j < n:  # b[j] against a gap
            rec(i, j + 1, "Y", acc - (gext if last == "Y" else gopen))

    rec(0, 0, "", 0)
    return best


def bruteforce_pair(a_codes, b_codes, mat, gap_init, gap_extend, forward_only=False):
    """Score, end and start positions by enumeration.

    H[i][j] (i over B, j over A) is the best score of any alignment of some
    A[a0..j] with some B[b0..i] (or 0).  The end is the first (i, j) in
    B-major order reaching the maximum; the start is the last (i, j) in
    B-major order inside [0..Bend]x[0..Aend] whose best alignment of
    A[j..a1] with B[i..b1] (a1 <= Aend, b1 <= Bend) reaches the maximum --
    the order in which the reference's reverse pass visits cells
    (codelet_batch_affine.cpp:96-114).
    """
    a = [int(x) for x in a_codes]
    b = [int(x) for x in b_codes]
    m, n = len(a), len(b)
    gopen, gext = -(gap_init + gap_extend), -gap_extend
    if m == 0 or n == 0:
        return 0, 0, 0
    H = [[0] * m for _ in range(n)]
    for i in range(n):
        for j in range(m):
            v = 0
            for b0 in range(i + 1):
                for a0 in range(j + 1):
                    v = max(v, _best_global(a[a0:j + 1], b[b0:i + 1], mat, gopen, gext))
            H[i][j] = v
    score, bend, aend = 0, 0, 0
    for i in range(n):
        for j in range(m):
            if H[i][j] > score:
                score, bend, aend = H[i][j], i, j
    bstart = astart = 0
    if not forward_only:
        s2 = 0
        for i in range(bend, -1, -1):
            for j in range(aend, -1, -1):
                v = 0
                for b1 in range(i, bend + 1):
                    for a1 in range(j, aend + 1):
                        v = max(v, _best_global(a[j:a1 + 1], b[i:b1 + 1], mat, gopen, gext))
                if v > s2:
                    s2, bstart, astart = v, i, j
    arange = ((aend & 0xFFFF) << 16) | (astart & 0xFFFF)
    brange = ((bend & 0xFFFF) << 16) | (bstart & 0xFFFF)
    return score, np.int32(np.uint32(arange)), np.int32(np.uint32(brange))


# ----------------------------------------------------------------------------
# SSW (reference comparator) access.
# ----------------------------------------------------------------------------

def ssw_dna(query: str, ref: str, match=1, mismatch=1, gap_open=1, gap_extend=1, ambiguity=1):
    """SSW DNA aligner, penalties as positive magnitudes like the reference tests
    (test/parity-test.cpp:72).  Returns (score, qbegin, qend, rbegin, rend)."""
    out = np.zeros(5, np.int32)
    rc = _ref().ssw_ref_align_dna(query.encode(), len(query), ref.encode(), len(ref), match,
                                  mismatch, gap_open, gap_extend, ambiguity, _ptr(out, _i32p))
    if rc != 0:
        raise RuntimeError("ssw align failed")
    return tuple(int(x) for x in out)


def ssw_matrix(query: str, ref: str, matrix: np.ndarray, translate: np.ndarray, gap_open: int,
               gap_extend: int):
    out = np.zeros(5, np.int32)
    mat = np.ascontiguousarray(matrix, dtype=np.int8)
    tr = np.ascontiguousarray(translate, dtype=np.int8)
    assert tr.size == 128
    rc = _ref().ssw_ref_align_matrix(query.encode(), len(query), ref.encode(), len(ref),
                                     _ptr(mat, _i8p), mat.shape[0], _ptr(tr, _i8p), gap_open,
                                     gap_extend, _ptr(out, _i32p))
    if rc != 0:
        raise RuntimeError("ssw align failed")
    return tuple(int(x) for x in out)


def ssw_matrix_cigar(query: str, ref: str, matrix: np.ndarray, translate: np.ndarray,
                     gap_open: int, gap_extend: int):
    """ssw_matrix plus SSW's CIGAR string (query = A, ref = B)."""
    out = np.zeros(5, np.int32)
    mat = np.ascontiguousarray(matrix, dtype=np.int8)
    tr = np.ascontiguousarray(translate, dtype=np.int8)
    buf = ctypes.create_string_buffer(4 * (len(query) + len(ref)) + 64)
    rc = _ref().ssw_ref_align_matrix_cigar(query.encode(), len(query), ref.encode(), len(ref),
                                           _ptr(mat, _i8p), mat.shape[0], _ptr(tr, _i8p),
                                           gap_open, gap_extend, _ptr(out, _i32p), buf,
                                           len(buf))
    if rc != 0:
        raise RuntimeError("ssw align failed")
    return tuple(int(x) for x in out), buf.value.decode()


def translate_table(datatype: int) -> np.ndarray:
    """128-entry char->code table of the reference's encoding, for SSW."""
    t = np.zeros(128, np.int8)
    for c in range(128):
        t[c] = encode(datatype, bytes([c]))[0]
    return t
