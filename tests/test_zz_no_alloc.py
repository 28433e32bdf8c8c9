"""sphx_export_pairs_noalloc (the entry point that allocates nothing) against
sphx_export_pairs, through the C ABI: equal, bit for bit, on two calls in a
row with the second on a smaller input (the shared scratch keeps the first
call's pairs and counter), and a clear SPHX_E_CAPACITY above its maximum."""
import ctypes as C

import numpy as np
import pytest

import sphx
from sphx import scenarios

pytestmark = pytest.mark.gpu


def noalloc_call(ctx, cap):
    out = np.zeros((max(cap, 1), 2), dtype=np.uint32)
    count = C.c_uint64(0)
    rc = ctx._lib.sphx_export_pairs_noalloc(ctx._h, out.ctypes.data_as(C.POINTER(C.c_uint32)), cap,
                                            C.byref(count))
    return rc, count.value, out


def test_noalloc_export_matches_export(require_gpu):
    # a larger then a smaller input, each a context of its own (same device)
    for dims, seed in (((24, 20, 16), 1), ((10, 8, 6), 2)):
        st, p = scenarios.dam_break(*dims, v_sigma=0.2, seed=seed)
        st["id"] = np.argsort(scenarios.SplitMix64(seed).uniform(len(st["x"]))).astype(np.uint32)
        with sphx.Context(len(st["x"]), p) as ctx:
            ctx.upload(st)
            ctx.step(3)
            ref = ctx.export_pairs()
            rc, count, _ = noalloc_call(ctx, 0)     # count query
            assert rc == 3 and count == len(ref)
            rc, count, got = noalloc_call(ctx, len(ref))
            assert rc == 0 and count == len(ref)
            assert np.array_equal(got[:count], ref)
            assert np.array_equal(ctx.export_pairs(noalloc=True), ref)


def test_noalloc_export_rejects_more_than_its_maximum(require_gpu):
    # 96 x 96 x 64 lattice: ~23.6 M pairs > 2^21
    st, p = scenarios.dam_break(96, 96, 64, spacing=0.01, box=(1.2, 1.2, 0.8))
    with sphx.Context(len(st["x"]), p) as ctx:
        ctx.upload(st)
        ctx.density()
        rc, count, _ = noalloc_call(ctx, 0)
        assert rc == 3 and count > (1 << 21)
        assert "sphx_export_pairs" in ctx._lib.sphx_last_error(ctx._h).decode()
        assert count == len(ctx.export_pairs())
