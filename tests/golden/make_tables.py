"""Extract the reference's scoring tables into tests/golden/similarity_tables.json.

Reads src/swatlib/similarity.h of the reference tree (the NA_SUBSTITUTION
16 x 16 compatibility table and the BLOSUM50 / BLOSUM62 25 x 25 tables,
similarity.h:17-93) as text and writes the numbers, with the header's column
labels, as a fixture.  Runs in the build container only; the fixture is data.
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(sys.argv[1] if len(sys.argv) > 1 else "/root/reference",
                   "src", "swatlib", "similarity.h")


def table(text, name, side):
    m = re.search(r"const std::vector<int8_t> " + name + r" = \{(.*?)\};", text, re.S)
    if not m:
        raise SystemExit(f"{name} not found in {SRC}")
    body = m.group(1)
    labels = None
    values = []
    for line in body.splitlines():
        code, _, comment = line.partition("//")
        if not code.strip() and comment.strip() and labels is None:
            labels = comment.split()  # the column-label line
            continue
        values += [int(v) for v in re.findall(r"-?\d+", code)]
    if len(values) != side * side:
        raise SystemExit(f"{name}: {len(values)} values, expected {side * side}")
    return {"side": side, "labels": labels, "values": values}


def main():
    text = open(SRC).read()
    out = {"source": "src/swatlib/similarity.h:17-93",
           "NA_SUBSTITUTION": table(text, "NA_SUBSTITUTION", 16),
           "BLOSUM50": table(text, "BLOSUM50", 25),
           "BLOSUM62": table(text, "BLOSUM62", 25)}
    with open(os.path.join(HERE, "similarity_tables.json"), "w") as f:
        json.dump(out, f, indent=None, separators=(",", ":"))
        f.write("\n")


if __name__ == "__main__":
    main()
