"""The phase-separation probe (su2_phase_probe) on the GPU: both ways of
separating two dependent phases run, every phase's writes reach the next
workgroup, and the times come back positive and finite.  After the probe the
buffer holds, in every element, the number of phases run over both ways:
two warm-up rounds and two timed rounds, each of `phases` phases, once by
launches and once in the cooperative kernel."""
import ctypes as C
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("blocks", [8, 64])
def test_phase_probe_runs_and_counts(gpu, blocks):
    from qcd_2d_gpu_amd import _lib as L

    lib = gpu.load_library()
    words, phases = 4, 25
    buf = gpu.DeviceBuffer(blocks * 256 * words * 8)
    buf.zero()
    lu, bu = C.c_double(), C.c_double()
    rc = lib.su2_phase_probe(buf.ptr, blocks, words, phases, None, C.byref(lu), C.byref(bu))
    assert rc == L.SU2_OK, lib.su2_last_error()
    vals = buf.download(np.float64)
    buf.free()
    # each element gets +1 per phase: 2 rounds x (launches + cooperative) x phases
    assert np.all(vals == 4.0 * phases)
    assert 0.0 < lu.value < 1e4 and math.isfinite(lu.value)
    assert 0.0 < bu.value < 1e5 and math.isfinite(bu.value)
