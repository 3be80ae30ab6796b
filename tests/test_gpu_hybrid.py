"""Hybrid CPU+GPU heat runs (b2r_heat_run_hybrid): the host cores own the top
z planes, the GPU the rest, joint layers of R*D planes exchanged every R steps
(the CPU+GPU scheme of the B2R paper, heat-3D.cpp:47, 55; heat-3D.hpp:15).
The result equals the GPU-only run bit for bit."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

A = None


def setup_module(module):
    global A
    from b2r_amd import _abi
    A = _abi


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


CASES = [
    # env, blocks, R, D, mode, nsteps, cpu_planes
    ((24, 20, 32), (1, 1, 1), 1, 1, 0, 7, 1),
    ((24, 20, 32), (1, 1, 1), 2, 1, 1, 9, 5),
    ((32, 24, 40), (2, 1, 2), 3, 1, 0, 10, 3),
    ((32, 24, 40), (2, 2, 2), 2, 2, 1, 8, 13),
    ((24, 24, 48), (1, 2, 3), 1, 4, 0, 6, 20),
    ((40, 32, 64), (2, 1, 4), 2, 4, 0, 11, 8),
    ((30, 26, 36), (1, 1, 2), 4, 1, 1, 13, 18),
]


@pytest.mark.parametrize("env,blocks,R,D,mode,nsteps,cpu_planes", CASES)
def test_hybrid_equals_gpu_only(env, blocks, R, D, mode, nsteps, cpu_planes):
    say("hybrid", env, blocks, R, D, mode, nsteps, cpu_planes)
    p = A.heat_params(D, mode, (0.021, 0.033, 0.017))
    with A.Context(env, blocks, R, D, A.BOUNDARY_FROZEN, 4) as c:
        a, b = c.field(), c.field()
        c.init_uniform(a, 9)
        res, _ = c.heat_run(a, b, p, nsteps)
        want = c.download(res)
        c.init_uniform(a, 9)
        c.init_uniform(b, 9)
        res, ms = c.heat_run_hybrid(a, b, p, nsteps, cpu_planes, nthreads=4, timed=True)
        got = c.download(res)
    assert ms > 0
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_hybrid_argument_checks():
    p = A.heat_params(2, A.HEAT_F32)
    with A.Context((24, 24, 24), (1, 1, 1), 2, 2, A.BOUNDARY_FROZEN, 4) as c:
        a, b = c.field(), c.field()
        for planes in (3, 21):  # outside [R*D, env_z - R*D] = [4, 20]
            with pytest.raises(A.B2RError) as e:
                c.heat_run_hybrid(a, b, p, 4, planes)
            assert e.value.code == A.B2R_E_ARG
    with A.Context((24, 24, 24), (1, 1, 1), 1, 2, A.BOUNDARY_PERIODIC, 4) as c:
        a, b = c.field(), c.field()
        with pytest.raises(A.B2RError) as e:
            c.heat_run_hybrid(a, b, p, 4, 4)
        assert e.value.code == A.B2R_E_UNSUPPORTED
