"""b2r_boundary_updating_noalloc (the exchange with its message buffers in the
library's fixed device array) against b2r_boundary_updating, through the C
ABI: two calls in a row on different contexts, the second smaller, each
bitwise equal to the allocating path; and a geometry whose messages exceed
the array is refused before any launch."""
import ctypes as C
import sys

import numpy as np
import pytest

from _util import id_field

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _case(A, env, blocks, R, radius, periodic, esize):
    say("noalloc exchange", env, blocks, R, radius, periodic, esize)
    with A.Context(env, blocks, R, radius, periodic, esize) as c:
        lay = c.layout
        raw = id_field(lay, blocks, np.float64, -1)
        raw = (np.where(raw == -1, 255, raw % 251).astype(np.uint8) if esize == 1 else
               raw.astype({4: np.float32, 8: np.float64}[esize]))
        f, g = c.field(), c.field()
        c.raw_upload(f, raw)
        c.raw_upload(g, raw)
        c.boundary_updating(f, 0)
        c.boundary_updating_noalloc(g, 0)
        want, got = c.raw(f), c.raw(g)
        assert not np.array_equal(want, raw)  # the exchange wrote halos
        assert np.array_equal(got, want)


def test_noalloc_equals_allocating_exchange_twice_smaller_second():
    from b2r_amd import _abi as A
    # first call: 8 blocks of 4-byte cells, the larger messages
    _case(A, (96, 80, 64), (2, 2, 2), 2, 2, 1, 4)
    # second call on another context: smaller, frozen faces, byte cells; the
    # array keeps the first call's contents, which this call must not depend on
    _case(A, (40, 24, 32), (2, 1, 2), 1, 1, 0, 1)


def test_noalloc_refuses_messages_beyond_the_array():
    from b2r_amd import _abi as A
    # one Z message set: 2 * S_x * S_y * h * 4 bytes = 2 * 4098^2 * 4 > 64 MiB
    with A.Context((4096, 4096, 4), (1, 1, 1), 1, 1, 0, 4) as c:
        f = c.field()
        with pytest.raises(A.B2RError) as e:
            c.boundary_updating_noalloc(f, 0)
        assert e.value.code == A.B2R_E_ARG and "scratch" in str(e.value)
