"""Writes the golden fixture of the elastic path's staggered moduli from the
data file the reference ships (run in a container that has the reference tree
mounted).

Source file (raw little-endian float32, written by substagger,
forward/yang_version_2016_2_2/Iso_multishot_fd_subroutine_fsbc_zy_m2_pub.f90:539-542,
for the program's own configuration iso_2d_modelling_sin_test.par: padded
shot grid nxea = 903, nzea = 603, written column by column, z fastest):
  forward/yang_version_2016_2_2/musge_test   musge [903][603]

The model is laterally invariant (a four-layer model), so every column but
the last is the same: the fixture keeps column 0, the last column, and the
sha256 / L2 of the whole file, after checking that columns 1..901 equal
column 0.

Fixture written here (data only):
  yang_musge_columns.npz   col0 [603], col_last [603] float32, meta (sha256, shape, l2)
"""
import hashlib
import os
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
SRC = os.path.join(REF, "forward", "yang_version_2016_2_2", "musge_test")
OUT = os.path.dirname(os.path.abspath(__file__))


def main():
    raw = open(SRC, "rb").read()
    a = np.frombuffer(raw, dtype="<f4").reshape(903, 603)
    if not (a[1:902] == a[0][None, :]).all():
        raise SystemExit("musge_test is not laterally invariant; the two-column fixture would not be complete")
    meta = np.array([hashlib.sha256(raw).hexdigest(), "903x603", repr(float(np.linalg.norm(a.astype(np.float64))))])
    np.savez(os.path.join(OUT, "yang_musge_columns.npz"), col0=a[0].copy(), col_last=a[902].copy(), meta=meta)


if __name__ == "__main__":
    main()
