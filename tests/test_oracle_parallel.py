"""The OpenMP builds of the oracle (float32 storage: a multi-core CPU baseline;
FP64 storage: precision checks at full size) compute exactly what the
sequential builds compute: every field and the sphere state after
40 coupled steps of a 40x28x40 scene whose sphere is moved by the momentum it
exchanges (below the acceleration cap's reach, so the order of the force sum
shows in the trajectory: replaying the planes' impulses in reverse order
changes it), for several thread counts."""
import os
import subprocess
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))

SCRIPT = r"""
import sys, json
import numpy as np
sys.path.insert(0, sys.argv[1])
import oracle_py as O
bits = sys.argv[2]
bits = {"32": 32, "64": 64}.get(bits, bits)
params = dict(tau=0.6, tauPhi=0.58, sigma=0.01, gravity=1e-4, rhoL=1.0, rhoG=0.1, W=4.0)
s = O.OSim(40, 28, 40, bits=bits)
s.set_params(**params)
s.init_flat(14)
b = O.sphere(40, 28, 40, 20.3, 12, 19.7, 5, 3.0, bits=bits)
b.vy = 0.02
s.coupled(b, 40)
out = {k: getattr(s, k).tobytes().hex() for k in ("f", "h", "rho", "ux", "uy", "uz", "phi", "tag")}
out["sphere"] = [b.cx, b.cy, b.cz, b.vx, b.vy, b.vz, b.wx, b.wy, b.wz]
print(json.dumps(out))
"""


def run(bits, threads):
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    r = subprocess.run([sys.executable, "-c", SCRIPT, HERE, bits], env=env, capture_output=True,
                       text=True, check=True, timeout=300)
    import json
    return json.loads(r.stdout.strip().splitlines()[-1])


@pytest.mark.parametrize("bits", ["32", "64"])
def test_parallel_oracle_equals_sequential(bits):
    ref = run(bits, 1)
    for threads in (2, 5):
        got = run(bits + "_omp", threads)
        for k in ref:
            assert got[k] == ref[k], (threads, k)
