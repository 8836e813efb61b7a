"""mrpop_check_edges_noalloc: the host-buffer edge check staged through the
library's fixed per-device array instead of context allocations.  Its
results must equal mrpop_check_edges' bit for bit (edge verdicts and
per-sample masks) on consecutive calls, the second smaller than the first,
so that nothing of an earlier call's staging survives into a later one; a
call larger than the array is refused before anything runs."""
import numpy as np
import pytest

from mrpop import _lib, scenes

pytestmark = pytest.mark.gpu


def _edges(rng, p, n, lo=0.01, hi=2.0):
    starts = np.concatenate([p.start[None], p.goals])
    idx = rng.integers(len(starts), size=n)
    frm = starts[idx].astype(np.float32)
    d = rng.normal(size=(n, p.dim))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    to = (frm + d * rng.uniform(lo, hi, size=(n, 1))).astype(np.float32)
    return frm, to


def test_noalloc_matches_check_edges(ctx):
    rng = np.random.default_rng(7)
    p5 = scenes.make_problem("shelf5", 3, _lib.gpu_checker(ctx), n_goals=4)
    p4 = scenes.make_problem("bin4", 5, _lib.gpu_checker(ctx), n_goals=4)
    w5 = ctx.world(p5.scene, p5.n_arms, p5.bases)
    w4 = ctx.world(p4.scene, p4.n_arms, p4.bases)
    # first call: 5 arms, 40000 edges of 64 samples with masks (about 12 MB of staging)
    frm, to = _edges(rng, p5, 40000)
    print("[noalloc] call 1: shelf5 FULL G=64 40000 edges with masks", flush=True)
    ref_e, ref_m = w5.check_edges(_lib.MODE_FULL, 64, frm, to, with_mask=True)
    got_e, got_m = w5.check_edges(_lib.MODE_FULL, 64, frm, to, with_mask=True, noalloc=True)
    assert np.array_equal(got_e, ref_e) and np.array_equal(got_m, ref_m)
    assert 0 < ref_e.sum() < len(ref_e)
    # second call, smaller: 4 arms, 3000 edges, verdicts only; then with masks
    frm, to = _edges(rng, p4, 3000)
    print("[noalloc] call 2: bin4 FULL G=32 3000 edges", flush=True)
    ref_e = w4.check_edges(_lib.MODE_FULL, 32, frm, to)
    got_e = w4.check_edges(_lib.MODE_FULL, 32, frm, to, noalloc=True)
    assert np.array_equal(got_e, ref_e)
    assert 0 < ref_e.sum() < len(ref_e)
    ref_e, ref_m = w4.check_edges(_lib.MODE_DRRT, 16, frm, to, with_mask=True)
    got_e, got_m = w4.check_edges(_lib.MODE_DRRT, 16, frm, to, with_mask=True, noalloc=True)
    assert np.array_equal(got_e, ref_e) and np.array_equal(got_m, ref_m)
    # one arm with its base pose
    f1, t1 = frm[:500, 14:21], to[:500, 14:21]
    print("[noalloc] call 3: bin4 SINGLE arm 2 G=32 500 edges with masks", flush=True)
    ref_e, ref_m = w4.check_edges(_lib.MODE_SINGLE, 32, f1, t1, arm=2, with_mask=True)
    got_e, got_m = w4.check_edges(_lib.MODE_SINGLE, 32, f1, t1, arm=2, with_mask=True, noalloc=True)
    assert np.array_equal(got_e, ref_e) and np.array_equal(got_m, ref_m)


def test_noalloc_refuses_calls_beyond_the_array(ctx):
    p5 = scenes.make_problem("shelf5", 3, _lib.gpu_checker(ctx))
    w5 = ctx.world(p5.scene, p5.n_arms, p5.bases)
    # 5 arms: 280 bytes of endpoints per edge, so 130000 edges exceed 32 MiB
    n = 130000
    frm = np.repeat(p5.start[None], n, axis=0)
    with pytest.raises(_lib.MrpopError) as e:
        w5.check_edges(_lib.MODE_FULL, 32, frm, frm, noalloc=True)
    assert e.value.code == _lib.ERR_BUDGET
    # a call that fits still works afterwards
    got = w5.check_edges(_lib.MODE_FULL, 32, frm[:1000], frm[:1000], noalloc=True)
    assert not got.any()  # the start configuration is free
