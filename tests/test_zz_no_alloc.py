"""The allocation-free time step (knc_block_time_step_noalloc) against the
context's knc_time_step: the same dt, bit for bit, on two calls in a row
(the second on a smaller block) for SRHD and Euler, on the null stream and
on a stream of the caller's; the same error for a failing zone; sizes beyond
the scratch array's limit rejected before any GPU work.  The block's
conserved planes sit in device memory the test allocates through the HIP
runtime the library itself links."""
import ctypes

import numpy as np
import pytest

import knc
from knc import capi

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    capi.lib()   # loads libamdhip64.so.7 first; the name below then resolves to that copy
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    hip.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


class DevicePlanes:
    """The four conserved planes of a block, SoA, in device memory."""

    def __init__(self, hip, conserved):
        nr, nq, _ = conserved.shape
        soa = np.ascontiguousarray(np.moveaxis(conserved, 2, 0))
        self.hip = hip
        self.ptr = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.ptr), soa.nbytes) == 0
        assert hip.hipMemcpy(self.ptr, soa.ctypes.data, soa.nbytes, H2D) == 0
        self.shape = (nr, nq)

    def free(self):
        self.hip.hipFree(self.ptr)


def _context_dt(hydro, b, rv, cons, sm):
    ctx = knc.Context(hydro)
    ctx.insert_block(b.index, rv, b.qv, cons, sm)
    try:
        return ctx.time_step()
    finally:
        ctx.close()


def _two_calls(setup, block, stream=None):
    hip = _hip()
    h = knc.hydro_from_spec(setup.hydro)
    b = setup.blocks[block]
    half = b.nr // 2
    cases = [(b.rv, b.conserved, b.scalar_mass), (b.rv[:half + 1], b.conserved[:half], b.scalar_mass[:half])]
    out = []
    for rv, cons, sm in cases:
        ref = _context_dt(h, b, rv, cons, sm)
        planes = DevicePlanes(hip, cons)
        try:
            dt = capi.block_time_step_noalloc(h, b.index, rv, b.qv, planes.ptr.value, stream)
        finally:
            planes.free()
        print(f"noalloc {setup.name} block {b.index} {cons.shape[0]}x{cons.shape[1]}: dt {dt!r} context {ref!r}",
              flush=True)
        assert dt == ref
        out.append(dt)
    return out


def test_srhd_adaptive_dt_matches_context_twice():
    s = knc.setups.bomb_jet(nr=512, nq=128, nblocks=4, nshell=8)
    _two_calls(s, 1)


def test_euler_dt_matches_context_twice_on_a_caller_stream():
    hip = _hip()
    st = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(st)) == 0
    try:
        _two_calls(knc.setups.euler_blast(), 0, st)
    finally:
        hip.hipStreamDestroy(st)


def test_srhd_fixed_dt_matches_context():
    s = knc.setups.jet_in_cloud(nr=256, nq=64, nblocks=2, corr=4.0)
    assert not s.hydro.adaptive_time_step
    _two_calls(s, 0)


def test_failing_zone_reported_like_the_context():
    hip = _hip()
    s = knc.setups.bomb_jet(nr=256, nq=64, nblocks=2, nshell=4)
    h = knc.hydro_from_spec(s.hydro)
    b = s.blocks[0]
    cons = b.conserved.copy()
    cons[5, 7, 0] = -1.0
    ctx = knc.Context(h)
    ctx.insert_block(b.index, b.rv, b.qv, cons, b.scalar_mass)
    with pytest.raises(knc.HydroError) as ref:
        ctx.time_step()
    ctx.close()
    planes = DevicePlanes(hip, cons)
    try:
        with pytest.raises(knc.HydroError) as got:
            capi.block_time_step_noalloc(h, b.index, b.rv, b.qv, planes.ptr.value)
    finally:
        planes.free()
    r, g = ref.value, got.value
    assert (g.code, g.block, g.i, g.j) == (r.code, r.block, r.i, r.j) == (1, b.index, 5, 7)
    assert (g.r, g.theta, g.value, g.conserved) == (r.r, r.theta, r.value, r.conserved)
    # the next call starts from a clear error log
    planes = DevicePlanes(hip, b.conserved)
    try:
        dt = capi.block_time_step_noalloc(h, b.index, b.rv, b.qv, planes.ptr.value)
    finally:
        planes.free()
    assert dt == _context_dt(h, b, b.rv, b.conserved, b.scalar_mass)


def test_sizes_beyond_the_scratch_limit_are_rejected():
    h = knc.relativistic_hydro()
    nr = capi.NOALLOC_MAX_NR + 1
    rv = np.linspace(1.0, 2.0, nr + 1)
    qv = np.linspace(0.0, np.pi, 9)
    with pytest.raises(RuntimeError, match="no-allocation scratch"):
        capi.block_time_step_noalloc(h, 0, rv, qv, 1)
    qv = np.linspace(0.0, np.pi, capi.NOALLOC_MAX_NQ + 2)
    with pytest.raises(RuntimeError, match="no-allocation scratch"):
        capi.block_time_step_noalloc(h, 0, rv[:9], qv, 1)
