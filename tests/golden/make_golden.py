"""Writes tests/golden/oracle_golden.json: small graphs whose expected results were derived by hand
(the derivation is written next to each case), used to pin the CPU oracle.

The reference ships no golden vectors (its tests only print, and read datasets that are not in its
tree), so these cases are this repository's own known answers.
"""
import json
import os

INF = 0xFFFFFFFF

# Graph G5 (directed): 0->1 w4, 0->2 w1, 1->3 w1, 2->1 w2, 2->3 w5, 3->4 w3.
G5 = {"row_offsets": [0, 2, 3, 5, 6, 6], "col_indices": [1, 2, 3, 1, 3, 4], "weights": [4, 1, 1, 2, 5, 3]}

CASES = [
    # d(0)=0; d(2)=1 (0->2); d(1)=min(4, 1+2)=3; d(3)=min(3+1, 1+5)=4; d(4)=4+3=7
    dict(kind="sssp", src=0, expect=[0, 3, 1, 4, 7], **G5),
    # from 2: d(1)=2, d(3)=min(2+1, 5)=3, d(4)=6; 0 unreachable
    dict(kind="sssp", src=2, expect=[INF, 2, 0, 3, 6], **G5),
    # hops from 0: {1,2} at 1, 3 at 2, 4 at 3
    dict(kind="bfs", src=0, expect=[0, 1, 1, 2, 3], **G5),
    # from 4: nothing reachable
    dict(kind="bfs", src=4, expect=[INF, INF, INF, INF, 0], **G5),
    # 3x3: row0 = {0:w1, 2:w5}, row1 empty, row2 = {1:w2}; x = [10, INF, 3]
    # y0 = min(1+10, 5+3) = 8; y1 = INF (empty row); y2 = INF (x[1] is INF, saturating combine)
    dict(kind="spmv_minplus", row_offsets=[0, 2, 2, 3], col_indices=[0, 2, 1], weights=[1, 5, 2],
         x=[10, INF, 3], expect=[8, INF, INF]),
    # unit weights on the same matrix: y0 = min(11, 4) = 4
    dict(kind="spmv_minplus", row_offsets=[0, 2, 2, 3], col_indices=[0, 2, 1], weights=None,
         x=[10, INF, 3], expect=[4, INF, INF]),
]

if __name__ == "__main__":
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "oracle_golden.json")
    with open(out, "w") as f:
        json.dump(CASES, f, indent=1)
    print("wrote", out)
