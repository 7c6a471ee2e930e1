"""The encode entry point that allocates nothing
(lig_rs_encode_non_systematic_cols_dev_noalloc), through the C ABI: its
output equals the allocating entry point's on the same columns, on two calls
in a row (the second smaller, so stale scratch contents would show), device
memory use does not move across its calls, and sizes beyond its 64 MiB
scratch are refused."""
import sys

import numpy as np
import pytest

import ligerito_hip as lh
from hipmem import DeviceBuffer, hip

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def used_bytes():
    import ctypes as C

    free, total = C.c_size_t(0), C.c_size_t(0)
    assert hip().hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return total.value - free.value


@pytest.fixture(scope="module")
def ctx():
    if lh.lib().lig_device_count() == 0:
        pytest.fail("no GPU visible: the HIP library cannot run (no CPU fallback exists)")
    c = lh.Context(0)
    yield c
    c.close()


def encode_both(O, ctx, bits, logm, n, seed):
    m = 1 << logm
    M = 4 * m
    rs = lh.ReedSolomonEncoding(ctx, m, M, bits)
    cols = O.splitmix(seed, M * n, bits)
    a, b = DeviceBuffer.of(cols), DeviceBuffer.of(cols)
    rs.encode_non_systematic_cols_dev(a.value, n)
    before = used_bytes()
    rs.encode_non_systematic_cols_dev_noalloc(b.value, n)
    after = used_bytes()
    want, got = a.download(cols), b.download(cols)
    a.free()
    b.free()
    return want, got, after - before


@pytest.mark.parametrize("bits", [32, 128])
def test_noalloc_matches_allocating_entry_point(O, ctx, bits):
    # the larger call first, then a smaller one on the same scratch
    for logm, n, seed in ((12, 64, 1), (6, 8, 2)):
        say("noalloc encode", bits, logm, n)
        want, got, grew = encode_both(O, ctx, bits, logm, n, seed)
        assert np.array_equal(got, want)
        assert grew == 0, f"device memory in use grew by {grew} bytes across the no-allocation call"
    # one column takes no scratch
    want, got, _ = encode_both(O, ctx, bits, 5, 1, 3)
    assert np.array_equal(got, want)


def test_noalloc_matches_oracle(O, ctx):
    m, n, bits = 1 << 7, 16, 32
    say("noalloc encode vs oracle")
    rs = lh.ReedSolomonEncoding(ctx, m, 4 * m, bits)
    cols = O.splitmix(9, 4 * m * n, bits)
    d = DeviceBuffer.of(cols)
    rs.encode_non_systematic_cols_dev_noalloc(d.value, n)
    got = d.download(cols).reshape(n, 4 * m)
    d.free()
    for j in range(n):
        assert np.array_equal(got[j], O.rs_encode_non_systematic(cols.reshape(n, 4 * m)[j].copy(), m, bits))


def test_noalloc_refuses_sizes_beyond_its_scratch(ctx):
    say("noalloc size limit")
    rs = lh.ReedSolomonEncoding(ctx, 1 << 12, 1 << 14, 32)
    d = DeviceBuffer(4)
    # 2^14 x 1025 x 4 bytes > 64 MiB: refused before any device work
    with pytest.raises(lh.LigeritoError) as e:
        rs.encode_non_systematic_cols_dev_noalloc(d.value, 1025)
    assert e.value.code == -1
    d.free()
