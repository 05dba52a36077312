"""swe_fct_positivity_csr_noalloc (FCT_positivity, SOURCES/fct.f90:5-42, with R- in the library's
fixed device array) against swe_fct_positivity_csr on the same inputs: bitwise, on two calls in a
row, the second smaller than the first, and the rows limit refused before any device work."""
import ctypes as C

import numpy as np
import pytest

from conftest import golden, same_bits

pytestmark = pytest.mark.gpu


def _random_case(nx, ny, seed):
    from swe_amd import mesh as M
    from swe_amd.solver import assemble
    rng = np.random.default_rng(seed)
    m = M.renumber_first_touch(M.structured_mesh(nx, ny, 0.0, 1.0, 0.0, 1.0, jitter=0.2, seed=seed))
    ia, ja, *_ = assemble(m.jj, m.rr)
    npn, nnz = ia.size - 1, ja.size
    f = rng.normal(0, 1.0, nnz)
    f[rng.random(nnz) < 0.1] = 0.0
    us = rng.uniform(-1e-4, 2.0, npn)
    us[rng.random(npn) < 0.1] = 0.0
    return ia, ja, f, us, 1e-3


def _with_alloc(ia, ja, f, us, htiny):
    from swe_amd.solver import _check, _ptr, load_library
    ia1 = np.ascontiguousarray(ia + 1, np.int32)
    ja1 = np.ascontiguousarray(ja + 1, np.int32)
    lij = np.zeros(ja.size)
    _check(load_library().swe_fct_positivity_csr(ia.size - 1, ja.size, _ptr(ia1), _ptr(ja1),
                                                 _ptr(np.ascontiguousarray(us)), _ptr(np.ascontiguousarray(f)),
                                                 float(htiny), _ptr(lij)))
    return lij


def _hip():
    """The HIP runtime the library itself links (found in this process's mappings after the library
    is loaded), for the test's own device buffers: a second HIP runtime in the process (torch's
    bundled one) would not see the GPU."""
    import os
    from swe_amd.solver import load_library
    load_library()
    path = None
    with open("/proc/self/maps") as f:
        for line in f:
            if "libamdhip64.so" in line:
                path = line.split()[-1]
                break
    hip = C.CDLL(path or "libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    assert os.path.basename(path or "libamdhip64.so").startswith("libamdhip64.so")
    return hip


def _no_alloc(ia, ja, f, us, htiny):
    from swe_amd.solver import _check, load_library
    hip = _hip()
    host = [np.ascontiguousarray(ia + 1, np.int32), np.ascontiguousarray(ja + 1, np.int32),
            np.ascontiguousarray(us, np.float64), np.ascontiguousarray(f, np.float64),
            np.full(ja.size, -7.0)]
    dev = []
    try:
        for a in host:
            p = C.c_void_p()
            assert hip.hipMalloc(C.byref(p), a.nbytes) == 0
            dev.append(p)
            assert hip.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0     # host to device
        d_ia, d_ja, d_us, d_f, d_l = dev
        _check(load_library().swe_fct_positivity_csr_noalloc(ia.size - 1, ja.size, d_ia, d_ja, d_us, d_f,
                                                             float(htiny), d_l, None))
        assert hip.hipDeviceSynchronize() == 0
        out = np.zeros(ja.size)
        assert hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), d_l, out.nbytes, 2) == 0   # device to host
    finally:
        for p in dev:
            hip.hipFree(p)
    return out


def test_no_alloc_equals_allocating_entry_point_twice_in_a_row():
    z = golden("fct.npz")
    big = (z["ia"], z["ja"], z["f"], z["us"], float(z["htiny"]))
    small = _random_case(9, 7, seed=4)
    assert small[0].size < big[0].size and small[1].size < big[1].size
    print("no-alloc FCT_positivity: rows", big[0].size - 1, "then", small[0].size - 1, flush=True)
    for case in (big, small):
        ref = _with_alloc(*case)
        got = _no_alloc(*case)
        assert same_bits(got, ref)
    # the reference's own output for the first case
    assert same_bits(_no_alloc(*big), z["lij"])


def test_no_alloc_refuses_rows_beyond_its_scratch():
    from swe_amd.solver import SWE_EINVAL, load_library
    lib = load_library()
    dummy = C.c_void_p(8)     # never dereferenced: the size check comes first
    rc = lib.swe_fct_positivity_csr_noalloc(1 << 23, 10, dummy, dummy, dummy, dummy, 1e-3, dummy, None)
    assert rc == SWE_EINVAL
    assert b"scratch" in lib.swe_last_error()
