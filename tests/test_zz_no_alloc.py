"""wdm_select_box_noalloc (the selection of a BoxOutlet, Outlet.cpp:12-40, without device
allocation) against wdm_select_box through the C ABI: the same ids in the same (ascending)
order and the same count, on two calls in a row, the second on a smaller scene (its memory
keeps the first call's ids beyond the second's count), with caps below the count, with the
inside/outside and mask variants, and on the NONE holes of removed particles."""
import numpy as np
import pytest

from woodem_amd import runner, scenes, soa

pytestmark = pytest.mark.gpu


def _scene(ctx, nx, ny, nz):
    a = scenes.box_lattice(nx=nx, ny=ny, nz=nz, box=((0., 0., 0.), (nx * .011 + .02, ny * .011 + .02, nz * .011 + .2)))
    r = runner.HipRun(a, runner.params_from_meta(a), runner.cell_from_meta(a), ctx=ctx)
    r.step(30)
    return r


def _check(ctx, sid, lo, hi, **kw):
    want = ctx.select_box(sid, lo, hi, **kw)
    got, n = ctx.select_box_noalloc(sid, lo, hi, **kw)
    assert n == len(want) and np.array_equal(got, want)
    if n > 3:  # a cap below the count: the first ids, and the full count
        got, n2 = ctx.select_box_noalloc(sid, lo, hi, cap=3, **kw)
        assert n2 == n and np.array_equal(got, want[:3])
    return n


def test_noalloc_select_equals_select_box(hip_ctx):
    ctx = hip_ctx
    big = _scene(ctx, 120, 100, 40)  # 480,000 spheres + 6 walls: 1,876 blocks
    n1 = _check(ctx, big.sid, (0., 0., 0.), (10., 10., .25), inside=True)
    assert n1 > 200000
    small = _scene(ctx, 9, 7, 5)  # 315 spheres, 2 blocks
    n2 = _check(ctx, small.sid, (0., 0., 0.), (10., 10., .03), inside=True)
    assert 0 < n2 < 315
    _check(ctx, small.sid, (0., 0., 0.), (10., 10., .03), inside=False)
    _check(ctx, small.sid, (0., 0., 0.), (10., 10., .03), inside=True, mask=soa.DEFAULT_MOVABLE_MASK, markMask=0b1000)
    assert _check(ctx, small.sid, (5., 5., 5.), (6., 6., 6.), inside=True) == 0
    # marked particles (markMask) are left out: mark a few, then select again
    sel = ctx.select_box(small.sid, (0., 0., 0.), (10., 10., .03), inside=True)
    ctx.set_masks(small.sid, sel[:10], soa.DEFAULT_MOVABLE_MASK | 0b1000)
    assert _check(ctx, small.sid, (0., 0., 0.), (10., 10., .03), inside=True, markMask=0b1000) == len(sel) - 10
    # holes of removed particles are never selected
    ctx.remove_particles(small.sid, sel[10:20])
    _check(ctx, small.sid, (0., 0., 0.), (10., 10., .03), inside=True)
    big.close()
    small.close()
