"""twisty_hip_run_noalloc (no device allocation) against twisty_hip_run.

The no-allocation entry point runs the same forward and finalize kernels with
the same schedule as twisty_hip_run; only the per-block partials and the FP64
result live in the library's static device scratch instead of context
buffers.  Its result must therefore equal twisty_hip_run's bit for bit, on
consecutive calls whose sizes shrink (the scratch keeps the previous call's
contents), and its limits must be reported before anything is launched.
"""
import numpy as np
import pytest

from tests import cases
from twisty_amd import capi

pytestmark = pytest.mark.gpu

FIELDS = ("log10_sum_w", "log10_sum_w2", "max_log10_w", "num_paths", "num_valid", "num_invalid", "num_clamped")


def _same(a, b, tag):
    for f in FIELDS:
        assert a[f] == b[f] or (np.isnan(a[f]) and np.isnan(b[f])), (tag, f, a[f], b[f])


def test_noalloc_matches_run_on_shrinking_calls(hip_device, oracle):
    case = cases.C2
    with cases.hip_context(case, oracle) as ctx:
        shapes = [(1 << 20, 0), (5000, 12345)]
        ref = [ctx.run(n, case.K, seed=case.seed, path_offset=off) for n, off in shapes]
        got = [ctx.run_noalloc(n, case.K, seed=case.seed, path_offset=off) for n, off in shapes]
        for (n, off), r, g in zip(shapes, ref, got):
            print(f"[noalloc C2] paths={n} offset={off}: log10 sum w {g['log10_sum_w']:.12f} "
                  f"(run {r['log10_sum_w']:.12f}), valid {g['num_valid']}", flush=True)
            assert g["num_paths"] == n and g["num_valid"] > 0
            _same(g, r, f"C2 paths={n}")


def test_noalloc_chains_and_burn_in(hip_device, oracle):
    case = cases.MID
    with cases.hip_context(case, oracle) as ctx:
        ctx.set_burn_in(3)
        shapes = [(4096, 16), (1000, 16)]
        ref = [ctx.run(n, case.K, seed=case.seed, chain_length=L) for n, L in shapes]
        got = [ctx.run_noalloc(n, case.K, seed=case.seed, chain_length=L) for n, L in shapes]
        ctx.set_burn_in(0)
        for (n, L), r, g in zip(shapes, ref, got):
            _same(g, r, f"N33 chains paths={n} chain={L}")


def test_noalloc_limits_reported_before_launch(hip_device, oracle):
    case = cases.MID
    with cases.hip_context(case, oracle) as ctx:
        before = ctx.run_noalloc(2048, case.K, seed=case.seed)
        with pytest.raises(capi.TwistyHipError) as e:
            ctx.run_noalloc(0, case.K, seed=case.seed)
        assert e.value.code == capi.ERR_ARG
        # 256 CUs x 32 blocks per CU exceeds the 4096 static partials
        ctx.set_schedule(0, 0, 32)
        with pytest.raises(capi.TwistyHipError) as e:
            ctx.run_noalloc(2048, case.K, seed=case.seed)
        assert e.value.code == capi.ERR_UNSUPPORTED and "static scratch" in str(e.value)
        ctx.set_schedule(0, 0, 0)
        ctx.set_exit_histogram(16, 16)
        with pytest.raises(capi.TwistyHipError) as e:
            ctx.run_noalloc(2048, case.K, seed=case.seed)
        assert e.value.code == capi.ERR_UNSUPPORTED
        ctx.set_exit_histogram(0, 0)
        _same(ctx.run_noalloc(2048, case.K, seed=case.seed), before, "after rejected calls")
