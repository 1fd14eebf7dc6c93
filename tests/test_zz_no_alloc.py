"""hps_emittance_noalloc (reduction partials in a static device array, no
allocation) against hps_beam_get_emittance on the same particles, through
the C ABI: two calls in a row, the second on a smaller beam, so the second
call's padding of the partial-sum trees must have been cleared of the
first call's partials."""
import ctypes as C

import numpy as np
import pytest

from hpsim_amd import HPSim
from hpsim_amd import _native as N
from hpsim_amd import configs as cfg

pytestmark = pytest.mark.gpu

COORDS = ("x", "xp", "y", "yp", "phi", "w")


def _beam_and_arrays(num, seed):
    spec = cfg.config(0, num=num, periods=1).beam
    spec.seed = seed
    dist = [np.ascontiguousarray(a, dtype=np.float64) for a in cfg.make_distribution(spec)]
    rng = np.random.default_rng(seed)
    loss = np.zeros(num, np.uint32)
    loss[rng.choice(np.arange(1, num), size=num // 37, replace=False)] = 7
    lloss = np.zeros(num, np.uint32)
    lloss[rng.choice(np.arange(1, num), size=num // 53, replace=False)] = 1
    b = HPSim.Beam(mass=spec.mass, charge=spec.charge, current=spec.current, num=num)
    b.set_distribution(*dist, loss=loss, lloss=lloss)
    b.set_frequency(spec.freq)
    return spec, b, dist, loss, lloss


def _getter(b):
    L = N.lib()
    out = []
    for axis in range(3):
        v = C.c_double()
        N.check(L.hps_beam_get_emittance(b._h, axis, C.byref(v)))
        out.append(v.value)
    return np.array(out)


class _Hip:
    """The HIP runtime the engine library itself links (a caller's own
    device buffers; torch bundles a second runtime that must not be mixed
    into the same process)."""

    def __init__(self):
        N.lib()
        self.rt = C.CDLL("libamdhip64.so.7")
        self.bufs = []

    def upload(self, a):
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), C.c_size_t(a.nbytes)) == 0
        self.bufs.append(p)
        assert self.rt.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0
        return p

    def free(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


def _noalloc(spec, dist, loss, lloss):
    hip = _Hip()
    try:
        dev = [hip.upload(a) for a in dist] + [hip.upload(loss), hip.upload(lloss)]
        dp = C.POINTER(C.c_double)
        up = C.POINTER(C.c_uint32)
        out = np.zeros(3)
        N.check(N.lib().hps_emittance_noalloc(*[C.cast(p, dp) for p in dev[:6]],
                                              C.cast(dev[6], up), C.cast(dev[7], up), len(loss),
                                              spec.mass, N.dptr(out)))
        return out
    finally:
        hip.free()


@pytest.mark.parametrize("nums", [(200_000, 3_000), (1_048_576, 70_001)])
def test_noalloc_emittance_matches_getter_twice(nums):
    for k, num in enumerate(nums):
        print(f"noalloc emittance: call {k + 1}, {num} particles", flush=True)
        spec, b, dist, loss, lloss = _beam_and_arrays(num, seed=11 + k)
        want = _getter(b)
        got = _noalloc(spec, dist, loss, lloss)
        assert np.all(np.isfinite(want)) and np.all(want > 0), want
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (num, got, want)


def test_noalloc_rejects_beams_beyond_its_limit():
    L = N.lib()
    dp = C.POINTER(C.c_double)
    up = C.POINTER(C.c_uint32)
    out = np.zeros(3)
    fake = C.cast(C.c_void_p(256), dp)   # never dereferenced: the size check comes first
    rc = L.hps_emittance_noalloc(fake, fake, fake, fake, fake, fake, C.cast(C.c_void_p(256), up),
                                 C.cast(C.c_void_p(256), up), 4194304 + 1, 939.294, N.dptr(out))
    assert rc != 0
    assert "exceed its limit of 4194304" in L.hps_last_error().decode()
