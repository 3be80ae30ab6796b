"""Checks at the workloads' full sizes, through properties that hold at any
size (the oracle's own full-size run would take minutes), and, for the elastic
step, the oracle itself for a few steps:
- heat delta1 768^3, one block, frozen faces: FP32 against the exact FP64
  eigenmode of the discrete operator (rel-L2 <= 1e-5); the fused R = 2 run
  equal to the single-pass R = 1 run bit for bit;
- heat delta4 768^3 on 2x2x2 blocks, periodic: against the exact FP64 Fourier
  mode (rel-L2 <= 1e-5); the 2x2x2 decomposition equal to one block bit for
  bit;
- elastic 256^3: the forward bitwise against the FP32 oracle for 4 steps, with
  the fused step kernel and with the stage kernels.
The eigenvalue uses the FP32-rounded taps and diffusion numbers the kernel
uses, so the closed form is exact for the discrete scheme; what remains is
FP32 rounding."""
import numpy as np
import pytest

from b2r_amd import _abi as A
import _elastic as E
import _oracle as O

pytestmark = pytest.mark.gpu

TAPS = {1: (-2.0, 1.0), 4: (-205 / 72, 1.6, -0.2, 8 / 315, -1 / 560)}


def say(*a):
    print(*a, flush=True)


def eigenmode(env, D, k, periodic):
    """The separable eigenmode (x, y, z factors) and its eigenvalue: frozen
    zero frame sin(pi i/(N+1)) (delta1), periodic cos(2 pi i/N)."""
    c = [float(np.float32(v)) for v in TAPS[D]]
    if periodic:
        th = [2.0 * np.pi / n for n in env]
        modes = [np.cos(t * np.arange(n)) for t, n in zip(th, env)]
        lam = 1.0 + sum(ka * (c[0] + 2.0 * sum(c[j] * np.cos(j * t) for j in range(1, D + 1)))
                        for ka, t in zip(k, th))
    else:
        modes = [np.sin(np.pi * np.arange(1, n + 1) / (n + 1)) for n in env]
        lam = 1.0 + sum(ka * (2.0 * np.cos(np.pi / (n + 1)) - 2.0) for ka, n in zip(k, env))
    return modes, lam


def run_heat(env, blocks, D, k, steps, periodic, u0, R=1, fused=0):
    p = A.heat_params(D, A.HEAT_F32, k)
    bnd = A.BOUNDARY_PERIODIC if periodic else A.BOUNDARY_FROZEN
    with A.Context(env, blocks, R, D, bnd, 4) as c:
        a, b = c.field(), c.field()
        c.upload(a, u0)
        res, _ = c.heat_run(a, b, p, steps, fused=fused)
        return c.download(res)


def mode_field(modes, env):
    sx, sy, sz = modes
    plane = np.outer(sy, sx)
    u = np.empty((env[2], env[1], env[0]), np.float32)
    for z in range(env[2]):
        u[z] = (sz[z] * plane).astype(np.float32)
    return u


def rel_l2_vs_mode(u, modes, lamn):
    sx, sy, sz = modes
    plane = np.outer(sy, sx)
    num = den = 0.0
    for z in range(u.shape[0]):
        ex = (lamn * sz[z]) * plane
        d = u[z].astype(np.float64) - ex
        num += float((d * d).sum())
        den += float((ex * ex).sum())
    return (num / den) ** 0.5


def test_heat_delta1_768_eigenmode_and_fusion():
    env, D, steps = (768, 768, 768), 1, 20
    k = [float(np.float32(0.1))] * 3
    modes, lam = eigenmode(env, D, k, False)
    u0 = mode_field(modes, env)
    say("heat delta1 768^3 single pass")
    u1 = run_heat(env, (1, 1, 1), D, k, steps, False, u0)
    err = rel_l2_vs_mode(u1, modes, lam ** steps)
    say("rel-L2 vs FP64", err)
    assert err <= 1e-5
    say("heat delta1 768^3 fused R=2")
    u2 = run_heat(env, (1, 1, 1), D, k, steps, False, u0, R=2, fused=1)
    assert np.array_equal(u1.view(np.uint32), u2.view(np.uint32))


def test_heat_delta4_768_fourier_mode_and_decomposition():
    env, D, steps = (768, 768, 768), 4, 10
    k = [float(np.float32(0.05))] * 3
    modes, lam = eigenmode(env, D, k, True)
    u0 = mode_field(modes, env)
    say("heat delta4 768^3 2x2x2")
    u8 = run_heat(env, (2, 2, 2), D, k, steps, True, u0)
    err = rel_l2_vs_mode(u8, modes, lam ** steps)
    say("rel-L2 vs FP64", err)
    assert err <= 1e-5
    say("heat delta4 768^3 one block")
    u1 = run_heat(env, (1, 1, 1), D, k, steps, True, u0)
    assert np.array_equal(u1.view(np.uint32), u8.view(np.uint32))


@pytest.mark.parametrize("kern", [1, 2])
def test_elastic_256_forward_bitwise_vs_oracle(kern):
    P = E.small_problem(env=(256, 256, 256), nsteps=4)
    say("elastic 256^3 forward, el_kernel", kern)
    with A.Context(P["env"], (1, 1, 1), 1, 4, A.BOUNDARY_FROZEN, 4) as c:
        c.set_tuning(el_kernel=kern)
        s = A.ElasticShot(c, float(P["dt"]), P["ih"])
        s.set_medium(P["lam"], P["mu"], P["b"])
        s.set_sources(P["src"], P["comp"], P["wav"])
        s.set_receivers(P["rec"], 1, P["nsteps"])
        s.forward(P["nsteps"], 2)
        rec = s.receivers()
        st = [s.state(f) for f in range(9)]
        s.close()
    st32, rec32 = E.forward(P["env"], P["lam"], P["mu"], P["b"], P["dt"], P["ih"], P["nsteps"],
                            P["src"], P["comp"], P["wav"], P["rec"], 1)
    assert np.array_equal(rec.view(np.uint32), rec32.view(np.uint32))
    for f in range(9):
        assert np.array_equal(st[f].view(np.uint32), st32[f].view(np.uint32)), f
