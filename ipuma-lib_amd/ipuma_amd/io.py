"""Dataset files as the reference's batch driver reads them.

``load_data`` follows ipusw/run_comparison.hpp:22-61: a ``.txt`` file holds
one sequence per line; any other file is FASTA (src/swatlib/fasta.cpp:15-51:
a record is a '>' header line, trailing space removed, and the following lines
joined with all whitespace removed; lines before the first header are
skipped).  Like the reference, the record being read when the file ends is
appended, so a file without any record yields one empty sequence.
"""
from __future__ import annotations


_SPACE = " \t\n\v\f\r"  # isspace() in the C locale
_DROP = {ord(c): None for c in _SPACE}


def _lines(path: str):
    """The file's lines as std::getline splits them: at '\n' only."""
    with open(path, "rb") as fh:
        data = fh.read().decode("latin-1")
    parts = data.split("\n")
    if parts and parts[-1] == "":
        parts.pop()
    return parts


def read_fasta(path: str, count: int = 0):
    """(description, sequence) of the file's records, the first ``count`` if > 0."""
    out = []
    desc, seq, have = "", [], False
    for line in _lines(path):
        if line.startswith(">"):
            if have:
                out.append((desc, "".join(seq)))
                if count and len(out) >= count:
                    return out
            desc, seq, have = line.rstrip(_SPACE), [], True
        elif have:
            seq.append(line.translate(_DROP))
    if have or not out:
        out.append((desc, "".join(seq)))
    return out[:count] if count else out


def load_sequences(path: str):
    """One sequence per line (ipusw/run_comparison.hpp:22-32)."""
    return _lines(path)


def load_data(path: str, count: int = 0):
    """ipusw/run_comparison.hpp:34-61."""
    if path.endswith(".txt"):
        return load_sequences(path)
    return [s for _, s in read_fasta(path, count)]
