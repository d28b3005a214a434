"""Writes tests/golden/kat.json: known answers derived by hand from the mathematics, not
from any implementation.

unit_tet: v0=(1,0,0), v1=(0,1,0), v2=(0,0,1), v3=(0,0,0). The P1 basis is
phi0=x, phi1=y, phi2=z, phi3=1-x-y-z, so the gradients are e_x, e_y, e_z, -(1,1,1) and
the reference's det = a.(b x c) = -1. lap K_ij = g_i.g_j. ela blocks with mu=1,
lambda=0.25: (g_i.g_j) I + 1.25 g_i g_j^T. The lap prec is 1/diag(K); the ela block
(3,3) is 3I + 1.25*11^T whose inverse is (1/3)(I - (5/27) 11^T) (checked in the test).

two_tet: e0=(1,2,3,4), e1=(2,3,4,5). CSR rows in first-seen order over each node's
elements (ascending) and their vertices: node 1 sees e0 only, nodes 2-4 see e0 then e1,
node 5 sees e1 only. The two tets share a face, so they get colors 0 and 1.

mask: E_ESSBCM codes 52 -> x, 53 -> y, 54 -> z, other nonzero -> all, 0 -> none.
"""
import json
import os

unit = {
    "gradients": [[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0], [-1.0, -1.0, -1.0]],
    "lap_K": [[1, 0, 0, -1], [0, 1, 0, -1], [0, 0, 1, -1], [-1, -1, -1, 3]],
    "lap_prec": [1.0, 1.0, 1.0, 1.0 / 3.0],
    "ela_block_00": [2.25, 0, 0, 0, 1, 0, 0, 0, 1],
    "ela_block_33": [4.25, 1.25, 1.25, 1.25, 4.25, 1.25, 1.25, 1.25, 4.25],
    "ela_block_03": [-2.25, -1.25, -1.25, 0, -1, 0, 0, 0, -1],
}
two = {
    "elemToNode": [1, 2, 3, 4, 2, 3, 4, 5],
    "row_ptr": [0, 4, 9, 14, 19, 23],
    "col_idx": [1, 2, 3, 4, 1, 2, 3, 4, 5, 1, 2, 3, 4, 5, 1, 2, 3, 4, 5, 2, 3, 4, 5],
    "colors": [0, 1],
    "colorToElem": [0, 1, 2],
}
mask = {
    "codes": [52, 53, 54, 10, 0, 200, 50],
    "mask": [[1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1], [0, 0, 0], [1, 1, 1], [1, 1, 1]],
}

if __name__ == "__main__":
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kat.json")
    with open(out, "w") as f:
        json.dump({"unit_tet": unit, "two_tet": two, "mask": mask}, f, indent=1)
    print("wrote", out)
