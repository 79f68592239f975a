"""Writes the golden fixtures of the acoustic path from the data files the
reference ships (run in a container that has the reference tree mounted).

Source files (raw little-endian float32, written by the reference program
forward/Acoustic_FD_by_ss/src/Acoustic_1order_FD.cpp at its own hard-coded
configuration):
  forward/Acoustic_FD_by_ss/file/record.dat  [100 traces][2000 samples] (:385-393)
  forward/Acoustic_FD_by_ss/file/Snap.dat    Vx interior [100][100] at k=500 (:343-355)
  forward/Acoustic_FD_by_ss/file/Model.dat   Vp interior [100][100] (:145-158)

Fixtures written here (data only):
  acoustic_reference_program.json  sha256, L2 norm, max|.| of each full file
  acoustic_record_traces.npy       traces 0, 25, 49, 50, 51, 75, 99 (float32)
  acoustic_snap_vx500.npy          the full Snap.dat (float32 [100][100])
"""
import hashlib
import json
import os
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
SRC = os.path.join(REF, "forward", "Acoustic_FD_by_ss", "file")
OUT = os.path.dirname(os.path.abspath(__file__))
TRACES = [0, 25, 49, 50, 51, 75, 99]


def main():
    meta = {}
    for name, shape in [("record.dat", (100, 2000)), ("Snap.dat", (100, 100)), ("Model.dat", (100, 100))]:
        raw = open(os.path.join(SRC, name), "rb").read()
        a = np.frombuffer(raw, dtype="<f4").reshape(shape)
        meta[name] = {"sha256": hashlib.sha256(raw).hexdigest(), "shape": list(shape),
                      "l2": float(np.linalg.norm(a.astype(np.float64))), "maxabs": float(np.abs(a).max())}
    rec = np.fromfile(os.path.join(SRC, "record.dat"), dtype="<f4").reshape(100, 2000)
    snap = np.fromfile(os.path.join(SRC, "Snap.dat"), dtype="<f4").reshape(100, 100)
    meta["record_traces"] = TRACES
    np.save(os.path.join(OUT, "acoustic_record_traces.npy"), rec[TRACES].astype(np.float32))
    np.save(os.path.join(OUT, "acoustic_snap_vx500.npy"), snap.astype(np.float32))
    with open(os.path.join(OUT, "acoustic_reference_program.json"), "w") as f:
        json.dump(meta, f, indent=1)


if __name__ == "__main__":
    main()
