"""The roofline microbenchmarks (lig_bench_rate) run and give positive,
finite rates; the table formulation the FFT passes use outruns the
formulations it was chosen over (register bit masks, general products)."""
import ctypes as C
import math
import sys

import pytest

import ligerito_hip as lh

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def test_bench_rates():
    if lh.lib().lig_device_count() == 0:
        pytest.fail("no GPU visible: the HIP library cannot run (no CPU fallback exists)")
    ctx = lh.Context(0)
    rates = {}
    for k in ("mul_tab", "mul_clmul", "valu", "mul_tab5", "lds128", "mul_mask", "gf8_logexp"):
        say("bench rate", k)
        rates[k] = ctx.bench_rate(k)
        assert math.isfinite(rates[k]) and rates[k] > 0, (k, rates[k])
    assert rates["mul_tab"] > 2 * rates["mul_mask"]
    assert rates["mul_tab"] > 2 * rates["mul_clmul"]
    # a tower GF(2^32) product takes at least nine GF(2^8) products
    assert rates["mul_tab"] > rates["gf8_logexp"] / 9
    r = C.c_double(0)
    assert lh.lib().lig_bench_rate(ctx._h, 7, C.byref(r)) == -1
    ctx.close()
