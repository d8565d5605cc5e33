"""Writes tests/golden/poseidon_fused.json from the reference's Poseidon
constant tables.

The reference's kernels run Poseidon in a fused form whose tables are
written out in native/poseidon_constants_template.cuh (lines 43-124: the
fused constant layer, the fused dense matrix, the fused S-box constants, and
the VS / W_HATS vectors of the 22 sparse partial rounds).  Three of them (the
dense matrix, VS and W_HATS) depend only on the MDS matrix, so a correct
derivation must reproduce them exactly; the other two also depend on boojum's
round constants, which the reference does not hold.

The fixture keeps, for each table, its size, a SHA-256 of its values (u64,
little-endian, in the template's order) and its first and last values, so a
test can compare a derivation against the reference without the tables
themselves.  This script reads the template as text (no code runs), and only
where /root/reference exists.  Run:  python tests/golden/gen_poseidon_golden.py
"""
import hashlib
import json
import os
import re
import struct

HERE = os.path.dirname(os.path.abspath(__file__))
TEMPLATE = "/root/reference/native/poseidon_constants_template.cuh"
TABLES = [
    "ROUND_CONSTANTS_FUSED_LAST_FULL_AND_FIRST_PARTIAL",
    "FUSED_DENSE_MATRIX_LAST_FULL_AND_FIRST_PARTIAL",
    "ROUND_CONSTANTS_FOR_FUSED_S_BOXES",
    "VS_FOR_PARTIAL_ROUNDS",
    "W_HATS_FOR_PARTIAL_ROUNDS",
]


def table(text, name):
    # the live definition (commented-out template placeholders precede it)
    bodies = re.findall(r"\n__constant__ constexpr base_field " + name + r"\[[^=]*=\s*\{(.*?)\};", text, re.S)
    pairs = re.findall(r"\{(0x[0-9a-f]+),(0x[0-9a-f]+)\}", bodies[-1])
    return [int(lo, 16) | (int(hi, 16) << 32) for lo, hi in pairs]


def digest(values):
    return hashlib.sha256(b"".join(struct.pack("<Q", v) for v in values)).hexdigest()


def main():
    text = open(TEMPLATE).read()
    out = {"source": "native/poseidon_constants_template.cuh:43-124", "tables": {}}
    for name in TABLES:
        v = table(text, name)
        out["tables"][name] = {"count": len(v), "sha256": digest(v), "first": hex(v[0]), "last": hex(v[-1])}
    mds = re.search(r"MDS_MATRIX_EXPS\[STATE_WIDTH\] = \{([0-9,]+)\};", text).group(1)
    out["mds_matrix_exps"] = [int(x) for x in mds.strip(",").split(",")]
    mu = re.search(r"LOG_MU_MINUS_ONE\[STATE_WIDTH\] = \{([0-9, ]+)\};", text).group(1)
    out["poseidon2_log_mu_minus_one"] = [int(x) for x in mu.replace(" ", "").strip(",").split(",")]
    with open(os.path.join(HERE, "poseidon_fused.json"), "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
