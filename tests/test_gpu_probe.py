"""The phase-separation probe (su2_phase_probe) on the GPU: the three ways of
separating two dependent phases run, every phase's writes reach the next
workgroup, and the times come back positive and finite.  After the probe the
buffer holds, in every element, the number of phases run over all ways:
a warm-up round and a timed round, each of `phases` phases, by launches, in
the cooperative kernel with the grid barrier and (unless skipped) in the one
with the XCD-local barrier, which itself checks every read after its
barrier against the writer's value and fails the call on a stale one."""
import ctypes as C
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("blocks,xcd", [(8, True), (64, True), (64, False), (3, True)])
def test_phase_probe_runs_and_counts(gpu, blocks, xcd):
    from qcd_2d_gpu_amd import _lib as L

    lib = gpu.load_library()
    words, phases = 4, 25
    buf = gpu.DeviceBuffer(blocks * 256 * words * 8)
    buf.zero()
    lu, bu, xu = C.c_double(), C.c_double(), C.c_double(-1.0)
    rc = lib.su2_phase_probe(buf.ptr, blocks, words, phases, None, C.byref(lu), C.byref(bu),
                             C.byref(xu) if xcd else None)
    assert rc == L.SU2_OK, lib.su2_last_error()
    vals = buf.download(np.float64)
    buf.free()
    # each element gets +1 per phase: 2 rounds x (launches + cooperative [+ XCD-local]) x phases
    assert np.all(vals == (6.0 if xcd else 4.0) * phases)
    assert 0.0 < lu.value < 1e4 and math.isfinite(lu.value)
    assert 0.0 < bu.value < 1e5 and math.isfinite(bu.value)
    if xcd:
        assert 0.0 < xu.value < 1e5 and math.isfinite(xu.value)
    else:
        assert xu.value == -1.0
