"""Measurement probes of the library: the plain copy and the stream kernels'
population pattern report finite positive rates below the device's peak, and
reject arguments outside their range."""
import math

import pytest

from gpu_util import announce
from lbm_py import lbm

pytestmark = pytest.mark.gpu

PEAK_GBS = 8000.0


def test_copy_probe_rate():
    announce("copy probe")
    g = lbm.copy_bandwidth(1 << 26, 3)
    assert math.isfinite(g) and 0 < g < PEAK_GBS, g


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
@pytest.mark.parametrize("dims", [(64, 48, 40), (33, 47, 29)])
def test_pull_pattern_probe_rate(storage, dims):
    announce("pull pattern probe", storage, dims)
    g = lbm.pull_bandwidth(*dims, storage, 3)
    assert math.isfinite(g) and 0 < g < PEAK_GBS, g


def test_pull_pattern_probe_arguments():
    for bad in ((2, 8, 8, "f32", 1), (8, 8, 8, "f32", 0)):
        with pytest.raises(lbm.LbmError):
            lbm.pull_bandwidth(*bad)
    L = lbm.lib()
    import ctypes as C
    g = C.c_double(0)
    assert L.lbm_probe_pull_bandwidth(8, 8, 8, 7, 1, C.byref(g)) != 0  # unknown storage
