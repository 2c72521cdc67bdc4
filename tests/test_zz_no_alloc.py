"""sdca_reweight_noalloc (through the C ABI) against sdca_reweight: the same weights and
neighbour counts bit for bit, on device buffers the caller owns (allocated here with
hipMalloc on the HIP runtime libsdca.so itself uses, and a stream of the caller's), on two
calls in a row -- the second with a smaller alignment -- and the size limit of the
library's fixed scratch array."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import superdca_amd as sd
from superdca_amd import synth

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def hip():
    sd.lib()  # loads the HIP runtime the library links; the same one serves the caller
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    h.hipStreamSynchronize.argtypes = [C.c_void_p]
    h.hipStreamDestroy.argtypes = [C.c_void_p]
    return h


class DeviceBuffers:
    """The caller's device memory and stream (hipMalloc / hipStreamCreate)."""

    def __init__(self):
        self.h = hip()
        self.ptrs = []
        self.stream = C.c_void_p()
        assert self.h.hipStreamCreate(C.byref(self.stream)) == 0

    def put(self, a: np.ndarray) -> int:
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), max(a.nbytes, 1)) == 0
        self.ptrs.append(p.value)
        assert self.h.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D) == 0
        return p.value

    def get(self, ptr: int, out: np.ndarray) -> np.ndarray:
        assert self.h.hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), out.nbytes, D2H) == 0
        return out

    def close(self):
        self.h.hipStreamSynchronize(self.stream)
        for p in self.ptrs:
            self.h.hipFree(C.c_void_p(p))
        self.h.hipStreamDestroy(self.stream)


def noalloc(states: np.ndarray, q: int, theta: float, mult: np.ndarray | None = None):
    B, L = states.shape
    db = DeviceBuffers()
    try:
        d_states = db.put(np.ascontiguousarray(states, dtype=np.uint8))
        d_mult = None if mult is None else db.put(np.ascontiguousarray(mult, dtype=np.uint32))
        # poisoned outputs: every entry must be written
        d_w = db.put(np.full(B, np.nan))
        d_iw = db.put(np.full(B, np.iinfo(np.uint64).max, dtype=np.uint64))
        sd.reweight_noalloc(d_states, B, L, q, theta, d_mult, d_iw, d_w, db.stream.value)
        assert db.h.hipStreamSynchronize(db.stream) == 0
        return db.get(d_w, np.empty(B)), db.get(d_iw, np.empty(B, dtype=np.uint64))
    finally:
        db.close()


def reference(states: np.ndarray, q: int, theta: float, mult: np.ndarray | None = None):
    with sd.Context(states, q=q) as ctx:
        return ctx.reweight(theta, mult)


def test_noalloc_matches_reweight_on_two_calls():
    d = synth.workload("C1")
    st, _ = sd.prepare_alignment(d["chars"])
    print(f"call 1: B={st.shape[0]} L={st.shape[1]} q=4 theta=0.9", flush=True)
    w1, iw1 = noalloc(st, 4, 0.9)
    rw1, riw1 = reference(st, 4, 0.9)
    np.testing.assert_array_equal(iw1, riw1)
    np.testing.assert_array_equal(w1, rw1)
    # a smaller alignment next, with multiplicities and another threshold: the scratch keeps
    # the first call's planes and counters, which must not leak into the result
    small = np.ascontiguousarray(st[:120, :500])
    mult = np.random.default_rng(3).integers(1, 5, small.shape[0]).astype(np.uint32)
    print(f"call 2: B={small.shape[0]} L={small.shape[1]} q=4 theta=0.95 with multiplicities", flush=True)
    w2, iw2 = noalloc(small, 4, 0.95, mult)
    rw2, riw2 = reference(small, 4, 0.95, mult)
    np.testing.assert_array_equal(iw2, riw2)
    np.testing.assert_array_equal(w2, rw2)


def test_noalloc_q5_alignment():
    cfg = sorted(synth.c3_workloads(), key=lambda c: c["B"] * c["N"])[2]
    st, _ = sd.prepare_alignment(synth.generate(**cfg)["chars"], q=5)
    print(f"B={st.shape[0]} L={st.shape[1]} q=5 theta=0.9", flush=True)
    w, iw = noalloc(st, 5, 0.9)
    rw, riw = reference(st, 5, 0.9)
    np.testing.assert_array_equal(iw, riw)
    np.testing.assert_array_equal(w, rw)


def test_noalloc_refuses_an_alignment_beyond_its_scratch():
    # B = 3000, L = 94,000, q = 5 needs 176 MB of bit-planes: refused before any launch
    db = DeviceBuffers()
    try:
        buf = db.put(np.zeros(8))
        with pytest.raises(sd.SdcaError) as e:
            sd.reweight_noalloc(buf, 3000, 94000, 5, 0.9, None, None, buf, db.stream.value)
    finally:
        db.close()
    assert e.value.code == sd.SDCA_E_ARG
    assert "scratch" in str(e.value)
