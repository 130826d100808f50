"""The no-allocation variant of four_to_phys (cfd_four_to_phys_planes)
against the allocating entry point cfd_four_to_phys, through the C ABI.

Both run the same z-inverse and paired x-inverse kernels on the same planes,
so the physical values must agree bit for bit.  Two calls in a row, the
second with a smaller input (fewer planes of another field of a smaller
channel), check that the shared per-GPU scratch keeps nothing that matters
from the previous call; a request larger than the scratch is refused with a
message giving the limit, and out-of-range plane ranges are refused."""
import ctypes

import numpy as np
import pytest

from cfd_amd import Channel, workloads
from cfd_amd._lib import CfdError

from .helpers import SMALLPOSTS, random_field

pytestmark = pytest.mark.gpu


def _fill(ch, p, seed):
    rng = np.random.default_rng(seed)
    for name in ("u1", "u2", "u3", "p"):
        ch.set_field(name, random_field(p, name, rng))


def _hip():
    """The HIP runtime the library is linked to (the mapping of
    libamdhip64 in this process once the library is loaded): device buffers
    for the output come from the same runtime as the library's streams."""
    from cfd_amd import _lib
    _lib.load()
    with open("/proc/self/maps") as fh:
        paths = {ln.split()[-1] for ln in fh if "libamdhip64" in ln}
    assert len(paths) == 1, paths
    hip = ctypes.CDLL(paths.pop())
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    return hip


def _planes(ch, name, plane0, nplanes):
    hip = _hip()
    p = ch.params
    out = np.full((nplanes, p.ngalz, p.ngalx), np.nan)
    dev = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(dev), out.nbytes) == 0
    try:
        # NaN in the device buffer first, so an unwritten element shows
        assert hip.hipMemcpy(dev, out.ctypes.data, out.nbytes, 1) == 0  # host to device
        ch.four_to_phys_planes(name, plane0, nplanes, dev.value)
        ch.sync()
        assert hip.hipMemcpy(out.ctypes.data, dev, out.nbytes, 2) == 0  # device to host
    finally:
        hip.hipFree(dev)
    return out


def test_no_alloc_four_to_phys_matches_allocating_entry_point():
    p, _ = workloads.get("rt395_384x192x192")
    print("case rt395_384x192x192 no-alloc four_to_phys", flush=True)
    big = Channel(p)
    _fill(big, p, 7)
    ref = big.four_to_phys("u1")
    # all 209 planes need 83 MB of half-transformed planes: refused
    with pytest.raises(CfdError, match="scratch"):
        big.four_to_phys_planes("u1", 0, big.nyu, 1)
    # first call: 40 planes from plane 100
    got = _planes(big, "u1", 100, 40)
    assert np.array_equal(got, ref[100:140])
    big.close()
    # second call, smaller input: 7 planes of p (odd count: the last plane
    # pairs with itself) of a smaller channel
    print("case smallposts no-alloc four_to_phys", flush=True)
    small = Channel(SMALLPOSTS)
    _fill(small, SMALLPOSTS, 8)
    ref = small.four_to_phys("p")
    got = _planes(small, "p", 3, 7)
    assert np.array_equal(got, ref[3:10])
    # every plane of u2 in one call, then a single plane
    ref = small.four_to_phys("u2")
    assert np.array_equal(_planes(small, "u2", 0, small.nyv), ref)
    assert np.array_equal(_planes(small, "u2", small.nyv - 1, 1), ref[-1:])
    for bad in ((-1, 2), (0, 0), (small.nyv - 1, 2)):
        with pytest.raises(CfdError, match="outside"):
            small.four_to_phys_planes("u2", bad[0], bad[1], 1)
    small.close()
