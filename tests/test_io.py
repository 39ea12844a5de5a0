"""Dataset files (ipusw/run_comparison.hpp:22-61, src/swatlib/fasta.cpp:15-51):
the Python loader and the C++ facade's loader read the same sequences, and
both match the reference's reading rules written out by hand below.  Host only."""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "ipuma-lib_amd"))
from ipuma_amd import io as IO  # noqa: E402

IO_BIN = os.path.join(ROOT, "ipuma-lib_amd", "lib", "io_test")

CASES = {
    # name: (bytes, count, expected sequences)
    "records.fa": (b"junk before the first header\n>r1 first\nACGT\nAC GT\n\n>r2\r\nTT\r\nG\r\n"
                   b">r3 last\nNNN", 0, ["ACGTACGT", "TTG", "NNN"]),
    "counted.fa": (b">a\nAA\n>b\nCC\n>c\nGG\n", 2, ["AA", "CC"]),
    "header_only.fa": (b">only\n", 0, [""]),
    "empty.fa": (b"", 0, [""]),
    "lines.txt": (b"ACGT\n\nTT\r\nG", 0, ["ACGT", "", "TT\r", "G"]),
}


def _cpp(path, count):
    out = subprocess.run([IO_BIN, path, str(count)], capture_output=True, check=True).stdout
    lines = out.decode("latin-1").split("\n")
    n = int(lines[0])
    return [ln[1:-1] for ln in lines[1:1 + n]]


@pytest.mark.parametrize("name", sorted(CASES))
def test_loaders(tmp_path, name):
    data, count, expected = CASES[name]
    p = tmp_path / name
    p.write_bytes(data)
    assert IO.load_data(str(p), count) == expected
    if os.path.exists(IO_BIN):
        assert _cpp(str(p), count) == expected


def test_read_fasta_descriptions(tmp_path):
    p = tmp_path / "d.fa"
    p.write_bytes(b">first record  \t\nAC\n>second\nGT\n")
    assert IO.read_fasta(str(p)) == [(">first record", "AC"), (">second", "GT")]
    if os.path.exists(IO_BIN):
        out = subprocess.run([IO_BIN, str(p), "0", "first"], capture_output=True,
                             check=True).stdout.decode()
        assert out.strip().split("\n")[-1] == "{>first record}{AC}"
