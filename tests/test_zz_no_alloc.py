"""The allocation-free potential getter, mpid_get_electrostatic_potential_noalloc,
through the C ABI: its values equal those of mpid_get_electrostatic_potential
for the same evaluation, on two calls in a row (the second with fewer
points, so stale staging data past its end must not leak into the result),
and it rejects a point count above its staging limit and a context without
a completed evaluation.  Same kernel, same device state: equal to rounding
(relative 1e-12)."""
import ctypes

import numpy as np
import pytest

from mpidopenmmplugin_amd import _capi
from mpidopenmmplugin_amd.kernel import MPIDContext
from mpidopenmmplugin_amd.waterbox import make_waterbox

pytestmark = pytest.mark.gpu


def _noalloc(lib, ctx, points):
    pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1)
    out = np.full(pts.size // 3, np.nan)
    rc = lib.mpid_get_electrostatic_potential_noalloc(ctx, int(out.size), _capi.dptr(pts), _capi.dptr(out))
    return rc, out


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_noalloc_potential_matches_getter(precision):
    print(f"[gpu] noalloc potential water216 {precision}", flush=True)
    lib = _capi.load()
    force, pos, box = make_waterbox("water216")
    ctx = MPIDContext(force, box, precision=precision)
    kctx = ctx.kernel._ctx
    L = np.asarray(box, dtype=np.float64)  # the three box edges
    rng = np.random.default_rng(11)

    # before any evaluation: an error, not a result
    rc, _ = _noalloc(lib, kctx, rng.uniform(0.0, 1.0, (4, 3)) * L)
    assert rc == _capi.MPID_ERR_INVALID
    assert b"no completed evaluation" in lib.mpid_last_error(kctx)

    ctx.setPositions(pos)
    ctx.getState()
    for npts in (1500, 300):  # the second call is smaller than the first
        pts = rng.uniform(0.0, 1.0, (npts, 3)) * L
        ref = ctx.kernel.getElectrostaticPotential(pts, pos, box)
        rc, got = _noalloc(lib, kctx, pts)
        assert rc == _capi.MPID_OK, lib.mpid_last_error(kctx)
        assert got.shape == ref.shape
        assert np.all(np.isfinite(got))
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-9)

    # beyond the staging limit: rejected before any device work
    big = np.zeros((_capi.MPID_NOALLOC_MAX_POINTS + 1, 3))
    rc, _ = _noalloc(lib, kctx, big)
    assert rc == _capi.MPID_ERR_INVALID
    assert b"exceed the limit" in lib.mpid_last_error(kctx)
    # the context still works after the rejected call
    pts = rng.uniform(0.0, 1.0, (10, 3)) * L
    rc, got = _noalloc(lib, kctx, pts)
    assert rc == _capi.MPID_OK
    np.testing.assert_allclose(got, ctx.kernel.getElectrostaticPotential(pts, pos, box), rtol=1e-12, atol=1e-9)
