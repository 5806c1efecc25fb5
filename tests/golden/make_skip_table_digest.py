"""Writes tests/golden/mrg_skip_table.json: the SHA-256 digest of the
reference's MRG skip table (src/mrg_transitions.cu:14, 24 x 256 transitions of
9 u32 words, hashed as little-endian u32 in table order) and its shape.

Only the digest is kept, not the table: tests regenerate the table from the
MRG recurrence and compare digests. Run from the repository root while the
reference checkout is present:  python tests/golden/make_skip_table_digest.py
"""
import hashlib
import json
import os
import re
import struct
import sys

SRC = "/root/reference/src/mrg_transitions.cu"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mrg_skip_table.json")


def parse(path):
    text = open(path).read()
    body = text[text.index("cuda_mrg_skip_matrices"):]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    rows = re.findall(r"\{([0-9,\s]+)\}", body)
    vals = [[int(x) for x in r.split(",")] for r in rows]
    assert all(len(v) == 9 for v in vals), "unexpected row width"
    return vals


def main():
    vals = parse(SRC)
    assert len(vals) == 24 * 256, len(vals)
    blob = b"".join(struct.pack("<9I", *v) for v in vals)
    doc = {
        "source": "src/mrg_transitions.cu:14 (cuda_mrg_skip_matrices)",
        "shape": [24, 256, 9],
        "encoding": "little-endian uint32, table order",
        "sha256": hashlib.sha256(blob).hexdigest(),
    }
    json.dump(doc, open(OUT, "w"), indent=1)
    print(doc)


if __name__ == "__main__":
    sys.exit(main())
