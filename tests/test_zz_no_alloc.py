"""The allocation-free dot product (cghip_dot_noalloc) against the
context's tree-ordered dot (cghip_vector_dot, the same kernels): bit for
bit, on two calls in a row, the second with a smaller input, through the C
ABI on device pointers of a context's vectors."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hip():
    # the HIP runtime libcghip.so itself is linked against (same soname)
    return C.CDLL("libamdhip64.so.7")


def test_dot_noalloc_matches_vector_dot():
    from cgxx_amd import cghip
    from cgxx_amd.cghip import Context

    L = cghip.load()
    hip = _hip()
    n1, n2 = 1_000_003, 12_345
    rng = np.random.default_rng(7)
    a = rng.standard_normal(n1)
    b = rng.standard_normal(n1)
    with Context(0) as big, Context(0) as small:
        big.set_size(n1)
        big.upload("P", a)
        big.upload("R", b)
        small.set_size(n2)
        small.upload("P", a[:n2])
        small.upload("R", b[:n2])
        ref = {n1: big.dot("P", "R"), n2: small.dot("P", "R")}
        pa, pb, pz = big.pointer("P"), big.pointer("R"), big.pointer("Z")
        for n in (n1, n2):
            print(f"cghip_dot_noalloc n={n}", flush=True)
            assert L.cghip_dot_noalloc(pa, pb, n, pz, None) == 0, L.cghip_last_error(None)
            assert hip.hipDeviceSynchronize() == 0
            got = big.download("Z")[0]
            assert np.float64(got).view(np.uint64) == np.float64(ref[n]).view(np.uint64), (n, got, ref[n])
        # invalid arguments are refused before any launch
        assert L.cghip_dot_noalloc(pa, pb, -1, pz, None) == 1
        assert b"negative" in L.cghip_last_error(None)
        assert L.cghip_dot_noalloc(pa, pb, 10, None, None) == 1
