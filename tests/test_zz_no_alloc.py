"""pa_stage_sweep_noalloc (sweeps and energies on caller-owned device memory, allocating
nothing) gives the same bits as pa_stage_sweep (host arrays, scratch run), on two calls in a
row of which the second has the smaller input.

The caller's device buffers come from the HIP runtime the library itself is linked against
(loaded by the library, reached here through ctypes), so the test does not bring a second
HIP runtime into the process."""
import ctypes

import numpy as np
import pytest

import paising_amd as pa

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def hip():
    pa.lib()  # loads libamdhip64.so.7 as the library's dependency; the name below finds it
    h = ctypes.CDLL("libamdhip64.so.7")
    vp = ctypes.c_void_p
    h.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
    h.hipFree.argtypes = [vp]
    h.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
    h.hipStreamCreate.argtypes = [ctypes.POINTER(vp)]
    h.hipStreamSynchronize.argtypes = [vp]
    h.hipStreamDestroy.argtypes = [vp]
    for f in ("hipMalloc", "hipFree", "hipMemcpy", "hipStreamCreate", "hipStreamSynchronize",
              "hipStreamDestroy", "hipDeviceSynchronize"):
        getattr(h, f).restype = ctypes.c_int
    return h


class DevBuf:
    """Device copy of a numpy array (caller-owned memory for the entry point)."""

    def __init__(self, h, arr):
        self.h, self.nbytes = h, arr.nbytes
        self.ptr = ctypes.c_void_p()
        assert h.hipMalloc(ctypes.byref(self.ptr), max(1, arr.nbytes)) == 0
        assert h.hipMemcpy(self.ptr, arr.ctypes.data, arr.nbytes, H2D) == 0

    def get(self, like):
        out = np.empty_like(like)
        assert self.h.hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H) == 0
        return out

    def free(self):
        self.h.hipFree(self.ptr)

# (first call, second smaller call): (L, coding, msbits, R, EqSw, slice_bits)
PAIRS = [
    ((64, pa.SSC, 32, 400, 5, 0), (48, pa.SSC, 32, 150, 3, 0)),
    ((64, pa.MSC, 32, 3000, 6, 0), (24, pa.MSC, 32, 700, 4, 0)),
    ((96, pa.MSC, 64, 900, 4, 0), (32, pa.MSC, 64, 200, 7, 16)),
    ((256, pa.MSC, 64, 300, 2, 32), (128, pa.MSC, 32, 100, 3, 0)),  # one sublattice in LDS
    # packed-byte SSC: the one-part walk (L=48), then the two-sweep step (L=24)
    ((48, pa.SSC, 32, 300, 3, 2), (24, pa.SSC, 32, 100, 5, 2)),
    # MSC with one site per stream and half sweep (L=16), odd sweep count
    ((32, pa.MSC, 32, 900, 3, 0), (16, pa.MSC, 32, 300, 5, 0)),
]


def population(p, R, seed):
    rng = np.random.default_rng(seed)
    shape, dt = pa._pop_shape(p, R)
    if p.coding == pa.SSC:
        return (rng.integers(0, 2, size=shape, dtype=np.int8) * 2 - 1).astype(np.int8)
    return rng.integers(0, np.iinfo(dt).max, size=shape, dtype=dt, endpoint=True)


def sweep_noalloc(h, p, pop, R, IS, beta, use_stream):
    """One call on device copies of the inputs: on a stream of the caller's (use_stream) or on
    the null stream."""
    pop = np.ascontiguousarray(pop)
    zeros = np.zeros(pa._members(p, R), np.int32)
    bufs = [DevBuf(h, pop), DevBuf(h, zeros), DevBuf(h, zeros)]
    stream = ctypes.c_void_p()
    if use_stream:
        assert h.hipStreamCreate(ctypes.byref(stream)) == 0
    try:
        pa.stage_sweep_noalloc(p, R, IS, beta, bufs[0].ptr.value, bufs[1].ptr.value,
                               bufs[2].ptr.value, stream.value or 0)
        assert (h.hipStreamSynchronize(stream) if use_stream else h.hipDeviceSynchronize()) == 0
        return bufs[0].get(pop), bufs[1].get(zeros), bufs[2].get(zeros)
    finally:
        if use_stream:
            h.hipStreamDestroy(stream)
        for b in bufs:
            b.free()


@pytest.mark.parametrize("pair", PAIRS, ids=lambda pr: "L%d_%s_then_L%d" % (
    pr[0][0], "SSC" if pr[0][1] == pa.SSC else "MSC%d" % pr[0][2], pr[1][0]))
def test_noalloc_sweep_matches_allocating_stage(pair):
    h = hip()
    for k, (L, coding, w, R, eqs, sw) in enumerate(pair):
        print(f"noalloc call {k}: L={L} coding={coding} w={w} R={R} EqSw={eqs} sw={sw}", flush=True)
        p = pa.make_params(L=L, coding=coding, msbits=w, Rinit=R, eqsweeps=eqs, seed=77 + k,
                           slice_bits=sw)
        pop = population(p, R, 5 + k)
        IS, beta = 1000 + 17 * k, 0.3 + 0.1 * k
        ref_pop, ref_IE, ref_M = pa.stage_sweep(p, pop, R, IS, beta)
        got_pop, got_IE, got_M = sweep_noalloc(h, p, pop, R, IS, beta, use_stream=k == 0)
        assert np.array_equal(ref_pop, got_pop)
        assert np.array_equal(ref_IE, got_IE) and np.array_equal(ref_M, got_M)
        assert not np.array_equal(ref_pop, pop)  # the sweeps did change the population


def test_noalloc_rejects_options_it_does_not_take():
    p = pa.make_params(L=16, coding=pa.MSC, msbits=32, Rinit=64, eqsweeps=1, adaptive=1)
    with pytest.raises(pa.PAError):
        pa.stage_sweep_noalloc(p, 64, 0, 0.1, 1, 1, 1, 0)
    with pytest.raises(pa.PAError):
        pa.stage_sweep_noalloc(pa.make_params(L=16, coding=pa.SSC, Rinit=64), 0, 0, 0.1, 1, 1, 1, 0)
