"""GPU test of the entry point that allocates nothing
(hj_self_join_count_device_noalloc): through the C ABI, its counts and
statistics equal those of hj_self_join_count_device on the same device
points, on two calls in a row (the second with a smaller input, so it reuses
the scratch array that the first one left behind), for the FP64 kernel and the
FP16 filter; a join too large for the scratch array fails with a capacity
error before anything runs, and the entry point still works afterwards."""
import numpy as np
import pytest

import hilbertjoin_amd as hj
import oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kernel", ["fp64", "f16"])
def test_noalloc_equals_allocating_entry_point(ctx, kernel):
    eps = 0.06
    big = hj.gen_mixture(91, 60000, 16, 32, 0.03)
    small = hj.gen_mixture(92, 5000, 16, 32, 0.03)
    for X in (big, small):
        print(f"noalloc kernel={kernel} n={len(X)}", flush=True)
        dev = ctx.to_device(X)
        want, st_want = ctx.self_join_count(dev, eps, kernel=kernel)
        got, st = ctx.self_join_count_noalloc(dev, eps, kernel=kernel)
        assert got == want > 0
        assert st.tile_pairs == st_want.tile_pairs
        assert st.candidate_pairs == st_want.candidate_pairs
        assert st.band_pairs == st_want.band_pairs
        if len(X) <= 5000:
            assert got == oracle.self_count(X, eps)
        dev.free()


def test_noalloc_rejects_joins_beyond_its_scratch(ctx):
    X = hj.gen_uniform(93, 1_000_000, 64)
    dev = ctx.to_device(X)
    with pytest.raises(hj.HJError) as e:
        ctx.self_join_count_noalloc(dev, 0.2, kernel="fp64")
    assert e.value.code == 4 and "scratch" in str(e.value)
    dev.free()
    Y = hj.gen_uniform(94, 3000, 8)
    dY = ctx.to_device(Y)
    got, _ = ctx.self_join_count_noalloc(dY, 0.3)
    assert got == oracle.self_count(Y, 0.3)
    dY.free()
