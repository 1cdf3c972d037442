"""Writes tests/golden/transfer_tables.npz: the numeric values of the
reference's constant transfer tables (the 'inNODE' trilinear tables of
src/Solving_Operator4MultiGridRestrictionAndInterpolation.m:4-31 (span 2) and
:118-243 (span 4), and src/Solving_SubEleNodMat.m:4-76), read as data from the
reference checkout.  Run here (where /root/reference exists); the fixture is
committed, tests never read the reference.
"""
import os
import re

import numpy as np

REF = "/root/reference/src"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "transfer_tables.npz")


def numbers_block(text, start_pat):
    i = text.index(start_pat)
    j = text.index("[", i)
    k = text.index("]", j)
    rows = []
    for line in text[j + 1:k].splitlines():
        vals = [float(v) for v in re.findall(r"-?\d+(?:\.\d+)?(?:e-?\d+)?", line)]
        if vals:
            rows.append(vals)
    return np.array(rows)


def main():
    op = open(os.path.join(REF, "Solving_Operator4MultiGridRestrictionAndInterpolation.m")).read()
    case2 = op.index("case 2")
    case4 = op.index("case 4")
    ri2 = numbers_block(op[case2:case4], "strcmp(opt, 'inNODE')")
    ri4 = numbers_block(op[case4:], "strcmp(opt, 'inNODE')")
    sub = open(os.path.join(REF, "Solving_SubEleNodMat.m")).read()
    s2 = numbers_block(sub[sub.index("case 2"):sub.index("case 4")], "val =").astype(np.int32)
    s4 = numbers_block(sub[sub.index("case 4"):], "val =").astype(np.int32)
    assert ri2.shape == (27, 8) and ri4.shape == (125, 8), (ri2.shape, ri4.shape)
    assert s2.shape == (8, 8) and s4.shape == (64, 8), (s2.shape, s4.shape)
    np.savez(OUT, RI2=ri2, RI4=ri4, sub2=s2, sub4=s4)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
