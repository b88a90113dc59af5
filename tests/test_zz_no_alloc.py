"""The no-allocation level product (vckss_mi355x_spmv_noalloc_v1) against
the allocating one (vckss_mi355x_spmv_v1), through the C ABI.

Both run the same segment kernels in the same order, so the bar is bit
identity.  Two calls in a row share the library's fixed device scratch, the
second on a smaller input (a coarser level, fewer columns); a level that does
not fit the scratch returns 40 with a clear message; argument errors return
20.  Device memory for x and y is the caller's: the test allocates it with
hipMalloc from the HIP runtime the library itself links (one runtime in the
process, so the pointers are that runtime's).
"""
import ctypes as C
import functools

import numpy as np
import pytest

import fevc_amd as F
from fevc_amd import synth

pytestmark = pytest.mark.gpu

COLS = ("worker", "firm", "deletion", "outcome", "frequency", "target")
SCRATCH_DOUBLES = 4 << 20


@functools.lru_cache(maxsize=None)
def panel(workers, firms, seed):
    spec = synth.PanelSpec(workers=workers, firms=firms, periods=5, mobility=0.2, seed=seed, name="noalloc")
    d = synth.generate(spec)
    return tuple(np.asarray(d[k]) for k in COLS)


@functools.lru_cache(maxsize=None)
def hip():
    """The HIP runtime the engine library links (already loaded by it)."""
    F.library()
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


def noalloc(e, level, columns, x_host):
    """Copy x to caller-owned device memory, run the no-allocation product
    into caller-owned y (filled with NaN first), copy y back."""
    h = hip()
    x_host = np.ascontiguousarray(x_host, np.float64)
    nbytes = x_host.nbytes
    x, y = C.c_void_p(), C.c_void_p()
    assert h.hipMalloc(C.byref(x), nbytes) == 0 and h.hipMalloc(C.byref(y), nbytes) == 0
    try:
        nan = np.full_like(x_host, np.nan)
        assert h.hipMemcpy(x, x_host.ctypes.data, nbytes, 1) == 0  # host to device
        assert h.hipMemcpy(y, nan.ctypes.data, nbytes, 1) == 0
        rc = F.library().vckss_mi355x_spmv_noalloc_v1(e.generation, level, columns, x, y)
        out = np.empty_like(x_host)
        assert h.hipMemcpy(out.ctypes.data, y, nbytes, 2) == 0  # device to host
    finally:
        h.hipFree(x)
        h.hipFree(y)
    return rc, out


def test_noalloc_matches_allocating_entry_point_twice():
    e = F.Engine(*panel(20000, 1500, 11))
    try:
        levels = e.dims["levels"]
        assert levels >= 2
        n = e.level_vertices()
        rng = np.random.default_rng(1)
        for level, columns in ((0, 8), (1, 4)):  # the second call is smaller
            x = rng.normal(size=n[level] * columns)
            want = e.spmv(level, x, columns)
            print(f"noalloc: level {level}, {n[level]} vertices, {columns} columns", flush=True)
            rc, got = noalloc(e, level, columns, x)
            assert rc == 0, F.library().vckss_rust_engine_last_error()
            assert np.array_equal(got, want), np.max(np.abs(got - want))
        # argument errors: columns not a power of two, level out of range
        x = rng.normal(size=n[0] * 3)
        rc, _ = noalloc(e, 0, 3, x)
        assert rc == 20
        rc, _ = noalloc(e, levels, 1, x[: n[0]])
        assert rc == 20
    finally:
        e.release()


def test_noalloc_rejects_a_level_beyond_the_scratch():
    e = F.Engine(*panel(600000, 30000, 3))
    try:
        n0 = e.level_vertices()[0]
        rng = np.random.default_rng(2)
        x = rng.normal(size=n0 * 16)
        print(f"noalloc: level 0, {n0} vertices, 16 columns", flush=True)
        rc, got = noalloc(e, 0, 16, x)
        assert rc == 0, F.library().vckss_rust_engine_last_error()
        assert np.array_equal(got, e.spmv(0, x, 16))
        # x, y and at least one chunk partial per row: 3 n0 columns doubles
        assert 3 * n0 * 64 > SCRATCH_DOUBLES
        rc, _ = noalloc(e, 0, 64, rng.normal(size=n0 * 64))
        assert rc == 40
        assert b"no-allocation scratch holds" in F.library().vckss_rust_engine_last_error()
    finally:
        e.release()
