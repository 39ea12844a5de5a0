"""Scoring tables against the reference's own (src/swatlib/similarity.h:17-93,
extracted as numbers into tests/golden/similarity_tables.json by
tests/golden/make_tables.py): the oracle's and the engine's (csrc/sw_tables.cpp)
transcriptions entry by entry, through selectMatrix's scaling rule
(src/swatlib/similarity.cpp:19-33, 89-104).  Host only."""
import ctypes
import json
import os

import numpy as np
import pytest

import oracle as O
import ipuma_amd as I
from ipuma_amd import workloads as W

GOLD = os.path.join(os.path.dirname(__file__), "golden")
TABLES = json.load(open(os.path.join(GOLD, "similarity_tables.json")))
SIMS = {"NA_SUBSTITUTION": O.SIM_NA, "BLOSUM50": O.SIM_BLOSUM50, "BLOSUM62": O.SIM_BLOSUM62}
SCALINGS = [(1, -1, -1), (2, -3, -7), (5, -4, 0)]


def reference_select(name, match, mismatch, ambiguity):
    """selectMatrix on the extracted table: the NA compatibility table scaled
    (1 -> match, 0 -> mismatch, row and column 0 -> ambiguity), BLOSUM as is."""
    t = TABLES[name]
    m = np.array(t["values"], np.int64).reshape(t["side"], t["side"])
    if name != "NA_SUBSTITUTION":
        return m
    out = np.where(m != 0, match, mismatch)
    out[0, :] = ambiguity
    out[:, 0] = ambiguity
    return out


def engine_select(sim, match, mismatch, ambiguity):
    fn = getattr(I.lib(), "_ZN5ipuma13select_matrixEiiiiPa")
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_int] * 4 + [ctypes.c_void_p]
    out = np.zeros(25 * 25, np.int8)
    dim = fn(sim, match, mismatch, ambiguity, out.ctypes.data)
    assert dim > 0
    return out[: dim * dim].reshape(dim, dim).astype(np.int64)


def test_fixture_shape():
    assert TABLES["BLOSUM62"]["labels"] == list("ARNDCQEGHILKMFPSTWYVBJZX*")
    assert TABLES["NA_SUBSTITUTION"]["labels"] == list("NATCGURYKMSWBDHV")
    for name in ("BLOSUM50", "BLOSUM62"):
        m = reference_select(name, 1, -1, -1)
        assert (m == m.T).all(), name


def check_against_reference(got, name, scaling, who):
    exp = reference_select(name, *scaling)
    side = exp.shape[0]
    if name == "NA_SUBSTITUTION":
        # a 17th code (the encoder's unknown, 16) lies outside the reference's
        # 16 x 16 table; both copies score it as the ambiguity value
        assert got.shape == (17, 17), who
        assert (got[16, :] == scaling[2]).all() and (got[:, 16] == scaling[2]).all(), who
    else:
        assert got.shape == (side, side), who
    bad = np.argwhere(got[:side, :side] != exp)
    assert len(bad) == 0, f"{name} {scaling}: {who} differs at {bad[:5].tolist()}"


@pytest.mark.parametrize("name", sorted(SIMS))
@pytest.mark.parametrize("scaling", SCALINGS)
def test_oracle_tables_equal_reference(name, scaling):
    check_against_reference(O.select_matrix(SIMS[name], *scaling).astype(np.int64), name,
                            scaling, "the oracle")


@pytest.mark.parametrize("name", sorted(SIMS))
@pytest.mark.parametrize("scaling", SCALINGS)
def test_engine_tables_equal_reference(name, scaling):
    check_against_reference(engine_select(SIMS[name], *scaling), name, scaling,
                            "sw_tables.cpp")


def test_planted_slip_is_caught():
    # a symmetric pair changed in both copies (the slip symmetry checks miss)
    exp = reference_select("BLOSUM62", 1, -1, -1)
    slipped = exp.copy()
    slipped[3, 7] = slipped[7, 3] = slipped[3, 7] + 1
    assert (slipped == slipped.T).all()
    with pytest.raises(AssertionError):
        check_against_reference(slipped, "BLOSUM62", (1, -1, -1), "a slipped copy")


def test_generator_blosum62_restriction():
    # the protein generator's substitution weights use BLOSUM62 on the 20
    # standard residues
    exp = reference_select("BLOSUM62", 1, -1, -1)[:20, :20]
    assert np.array_equal(np.asarray(W._B62_20, np.int64), exp)
