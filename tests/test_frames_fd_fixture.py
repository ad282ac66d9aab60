"""The committed finite differences of the frames cluster
(tests/golden/frames_gradients_fd.npz, read by the GPU gradient tests)
against the oracle: entries recomputed here with the same steps and stencils
(the ThreeFold and NoAxisType sites' charges, the ZThenX site's position,
without cutoff; the ThreeFold site's charge with PME) equal the fixture's to
1e-12 relative (the oracle is deterministic), and the fixture holds every
site and class the GPU test compares."""
import numpy as np
import pytest

from tests.helpers import frames_cluster
from tests.test_gpu_gradients import FRAME_CLASSES, FRAME_SITES, WIDTH, fd_gradients, frames_fd


@pytest.mark.parametrize("periodic,checks", [
    (False, [(0, "charges"), (15, "charges"), (2, "positions")]),
    (True, [(0, "charges")]),
], ids=["nocutoff", "pme"])
def test_fixture_matches_oracle(periodic, checks):
    force, pos, box = frames_cluster(periodic=periodic, target_eps=1e-10)
    flat = force.flatten()
    fd = frames_fd("pme" if periodic else "nocutoff", flat, pos, box)
    for cls in FRAME_CLASSES:
        w = 3 if cls == "positions" else WIDTH[cls]
        assert fd[cls].shape == (len(FRAME_SITES), w), cls
        assert np.all(np.isfinite(fd[cls])), cls
    for site, cls in checks:
        got = fd_gradients(flat, pos, box, [site], [cls])[cls][0]
        np.testing.assert_allclose(fd[cls][FRAME_SITES.index(site)], got, rtol=1e-12, atol=0, err_msg=(site, cls))
