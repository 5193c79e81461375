"""Reader for the golden vectors in tests/golden/ (written by make_goldens.mjs).

Each case is <name>.json (metadata and a table of contents) plus
<name>.bin.gz (the arrays, concatenated little-endian)."""
import gzip
import json
import os

import numpy as np

GOLDEN_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

_DT = {"f4": np.float32, "f8": np.float64, "u1": np.uint8, "i4": np.int32}


def load(name):
    with open(os.path.join(GOLDEN_DIR, name + ".json")) as fh:
        meta = json.load(fh)
    with gzip.open(os.path.join(GOLDEN_DIR, name + ".bin.gz"), "rb") as fh:
        blob = fh.read()
    arrays = {}
    for e in meta["toc"]:
        dt = np.dtype(_DT[e["dtype"]])
        arrays[e["key"]] = np.frombuffer(blob, dtype=dt, count=e["count"], offset=e["offset"]).copy()
    return meta, arrays


SPHERE_FIELDS = ["nx", "ny", "nz", "cx", "cy", "cz", "r", "vx", "vy", "vz", "wx", "wy", "wz",
                 "density", "volume", "mass", "I", "dragging", "dragTargetX", "dragTargetY",
                 "dragTargetZ", "Fx", "Fy", "Fz", "Tx", "Ty", "Tz"]
