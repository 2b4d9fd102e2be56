"""Allocation-free Courant step (rh_courant_noalloc) against rh_courant.

The variant keeps its block minima and result in a static device array of the
library instead of buffers the context allocates.  Two calls in a row on two
contexts, the second on a smaller level (fewer blocks, larger time step): if
the second call read any block minimum left over from the first, it would
return the first level's smaller time step.  Both must equal rh_courant's
value bit for bit."""
import numpy as np
import pytest

import rhydro

pytestmark = pytest.mark.gpu


def _lognormal(level, seed):
    m = rhydro.uniform_mesh(level)
    rhydro.init_lognormal(m, level, seed, 5.0 / 3.0, 0.5, 0.5)
    return m


def test_noalloc_courant_equals_courant_two_calls_smaller_second():
    gamma, cf = 5.0 / 3.0, 0.8
    big, small = 7, 5  # 32768 octs (128 blocks), then 512 octs (2 blocks)
    ma, mb = _lognormal(big, 11), _lognormal(small, 12)
    sa = rhydro.Solver(gamma=gamma, riemann="hllc")
    sb = rhydro.Solver(gamma=gamma, riemann="hllc")
    try:
        sa.bind(ma, rhydro.RH_DEVICE_RESIDENT)
        sb.bind(mb, rhydro.RH_HOST_COHERENT)
        dxa, dxb = 0.5 ** big, 0.5 ** small
        print(f"case noalloc courant levels {big},{small}", flush=True)
        # Building a level's tables allocates: before any call on the level
        # the allocation-free variant refuses.
        with pytest.raises(rhydro.RhError) as err:
            sa.courant_noalloc(big, dxa, gamma, cf)
        assert err.value.code == rhydro.RH_ERR_ARG
        want_a = sa.courant(big, dxa, gamma, cf)
        want_b = sb.courant(small, dxb, gamma, cf)
        assert want_b > want_a  # a stale block minimum would show
        got_a = sa.courant_noalloc(big, dxa, gamma, cf)
        got_b = sb.courant_noalloc(small, dxb, gamma, cf)
        assert np.float64(got_a).tobytes() == np.float64(want_a).tobytes(), (got_a, want_a)
        assert np.float64(got_b).tobytes() == np.float64(want_b).tobytes(), (got_b, want_b)
        # Once more on the first level: the result does not depend on what
        # the previous call left in the array.
        assert sa.courant_noalloc(big, dxa, gamma, cf) == want_a
    finally:
        sa.close()
        sb.close()
