"""GPU parity of the elastic path through the C ABI: forward bit-exact against
the FP32 oracle (plain, batched and graph modes, multi-block B2R windows,
uneven batches), FP64 error report, and the hand-written adjoint against the
FP64 CPU adjoint, finite differences and a dot-product test."""
import sys

import numpy as np
import pytest

import _elastic as E
from _util import rel_l2

pytestmark = pytest.mark.gpu

A = None


def setup_module(module):
    global A
    from b2r_amd import _abi
    A = _abi


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bits_equal(a, b):
    return np.array_equal(np.ascontiguousarray(a, np.float32).view(np.uint32),
                          np.ascontiguousarray(b, np.float32).view(np.uint32))


def _shot(ctx, P, rec_comp=1, nt=None):
    s = A.ElasticShot(ctx, float(P["dt"]), P["ih"])
    s.set_medium(P["lam"], P["mu"], P["b"])
    s.set_sources(P["src"], P["comp"], P["wav"])
    s.set_receivers(P["rec"], rec_comp, nt or P["nsteps"])
    return s


# (mode, el_kernel): the plain loop, and the batched modes with the fused
# step kernel (1) and with the separate stress and velocity kernels (2)
MODE_KERNELS = [(0, 0), (1, 1), (1, 2), (2, 1), (2, 2)]


@pytest.mark.parametrize("mode,kern", MODE_KERNELS)
@pytest.mark.parametrize("env,blocks,R", [((48, 40, 48), (1, 1, 1), 1),
                                          ((48, 40, 48), (2, 1, 2), 1),
                                          ((48, 40, 48), (2, 2, 1), 2),
                                          ((48, 56, 56), (1, 2, 2), 3),
                                          ((130, 50, 36), (1, 1, 1), 1),
                                          ((136, 56, 40), (2, 2, 1), 2)])
def test_forward_bitwise_vs_oracle(env, blocks, R, mode, kern):
    P = E.small_problem(env=env, nsteps=13)
    say("elastic forward", env, blocks, "R", R, "mode", mode, "el_kernel", kern)
    with A.Context(P["env"], blocks, R, 4, A.BOUNDARY_FROZEN, 4) as c:
        c.set_tuning(el_kernel=kern)
        with _shot(c, P) as s:
            s.forward(P["nsteps"], mode)
            got = [s.state(k) for k in range(9)]
            rec = s.receivers()
    st, want_rec = E.forward(P["env"], P["lam"], P["mu"], P["b"], P["dt"], P["ih"], P["nsteps"],
                             P["src"], P["comp"], P["wav"], P["rec"], 1)
    for k in range(9):
        assert bits_equal(got[k], st[k]), A.EL_FIELDS[k]
    assert bits_equal(rec, want_rec)


@pytest.mark.parametrize("mode,kern", [(0, 0), (2, 1), (2, 2)])
@pytest.mark.parametrize("env,blocks,R", [((32, 24, 40), (1, 1, 1), 1),
                                          ((48, 40, 32), (2, 1, 2), 1),
                                          ((48, 48, 48), (1, 2, 2), 2)])
def test_periodic_forward_bitwise_vs_oracle(env, blocks, R, mode, kern):
    # _PERIODIC_ (b2r_config.h:12) for eqwp: a source two cells from the x face,
    # so the wave wraps within the run
    nsteps = 12
    P = E.small_problem(env=env, seed=7, nsteps=nsteps)
    src = np.array([[2, env[1] // 2, env[2] // 3]], np.int32)
    rec = np.array([[env[0] - 1, env[1] // 2, env[2] // 3], [0, 0, 0], [5, 3, env[2] - 1]],
                   np.int32)
    say("elastic periodic forward", env, blocks, "R", R, "mode", mode, "el_kernel", kern)
    with A.Context(env, blocks, R, 4, A.BOUNDARY_PERIODIC, 4) as c:
        c.set_tuning(el_kernel=kern)
        with A.ElasticShot(c, float(P["dt"]), P["ih"]) as s:
            s.set_medium(P["lam"], P["mu"], P["b"])
            s.set_sources(src, [1], P["wav"])
            s.set_receivers(rec, 1, nsteps)
            s.forward(nsteps, mode)
            got = [s.state(k) for k in range(9)]
            got_rec = s.receivers()
    st, want_rec = E.forward_periodic(env, P["lam"], P["mu"], P["b"], P["dt"], P["ih"], nsteps,
                                      src, [1], P["wav"], rec, 1)
    assert np.abs(st[1][:, :, -4:]).max() > 0  # wrapped across the x face
    for k in range(9):
        assert bits_equal(got[k], st[k]), A.EL_FIELDS[k]
    assert bits_equal(got_rec, want_rec)


@pytest.mark.parametrize("kern", [1, 2])
def test_forward_displacement_and_split_runs(kern):
    P = E.small_problem(env=(24, 20, 22), nsteps=10)
    with A.Context(P["env"], (1, 1, 1), 1, 4, A.BOUNDARY_FROZEN, 4) as c:
        c.set_tuning(el_kernel=kern)
        with _shot(c, P) as s:
            s.enable_displacement()
            s.forward(4, 0)
            s.forward(3, 1)   # continues the clock (sources, receivers)
            s.forward(3, 2)   # an odd count: a fused run ends in the second set
            disp = s.state("disp")
            vy = s.state("vy")
            rec = s.receivers()
    st, want_rec, d = E.forward(P["env"], P["lam"], P["mu"], P["b"], P["dt"], P["ih"], 10,
                                P["src"], P["comp"], P["wav"], P["rec"], 1, disp=True)
    assert bits_equal(vy, st[1]) and bits_equal(disp, d) and bits_equal(rec, want_rec)


def test_uneven_batch_matches_single_shots(tmp_path):
    envs = [(32, 24, 40), (48, 32, 24), (24, 40, 32)]
    Ps = [E.small_problem(env=e, seed=i, nsteps=9) for i, e in enumerate(envs)]
    ctxs = [A.Context(P["env"], (1, 1, 1), 1, 4, A.BOUNDARY_FROZEN, 4) for P in Ps]
    shots = [_shot(c, P) for c, P in zip(ctxs, Ps)]
    try:
        # the batch autotuner sets one choice in every context and resets the shots
        rec_file = str(tmp_path / "tune.txt")
        shots[0].forward(2, 1)
        t = A.el_autotune_batch(shots, rec_file)
        assert t["el_kernel"] in (1, 2)
        assert all(c.get_tuning()["el_kernel"] == t["el_kernel"] for c in ctxs)
        assert not np.any(shots[0].state(1))
        line = open(rec_file).read()
        assert "kind=el_batch shots=3" in line and "fused_ms_per_step=" in line
        A.el_forward_batch(shots, 9, mode=2)
        for P, s in zip(Ps, shots):
            _, want = E.forward(P["env"], P["lam"], P["mu"], P["b"], P["dt"], P["ih"], 9,
                                P["src"], P["comp"], P["wav"], P["rec"], 1)
            assert bits_equal(s.receivers(), want)
    finally:
        for s in shots:
            s.close()
        for c in ctxs:
            c.close()


def test_forward_error_vs_fp64():
    P = E.small_problem(env=(40, 36, 44), nsteps=60)
    with A.Context(P["env"], (1, 1, 1), 1, 4, A.BOUNDARY_FROZEN, 4) as c:
        with _shot(c, P) as s:
            s.forward(P["nsteps"], 2)
            rec = s.receivers()
    _, r64 = E.forward(P["env"], P["lam"], P["mu"], P["b"], float(P["dt"]), P["ih"], P["nsteps"],
                       P["src"], P["comp"], P["wav"].astype(np.float64), P["rec"], 1, fp64=True)
    err = rel_l2(rec, r64)
    say("elastic receivers rel-L2 vs FP64:", err, "max abs:", float(np.abs(rec - r64).max()))
    assert err < 1e-5


# ---------------------------------------------------------------- adjoint
def _obs(P):
    _, obs = E.forward(P["env"], P["lam"] * np.float32(1.03), P["mu"] * np.float32(0.98), P["b"],
                       P["dt"], P["ih"], P["nsteps"], P["src"], P["comp"], P["wav"], P["rec"], 1)
    return obs


@pytest.mark.parametrize("plain", [False, True])
@pytest.mark.parametrize("ckpt", [1, 7, 100])
def test_gradient_vs_fp64_cpu_adjoint(ckpt, plain):
    # both algorithms: the reverse sweep (default) and the plain
    # store-and-recompute loop
    P = E.small_problem(env=(20, 18, 16), nsteps=30)
    obs = _obs(P)
    with A.Context(P["env"], (1, 1, 1), 1, 4, A.BOUNDARY_FROZEN, 4) as c:
        with _shot(c, P) as s:
            J, g = s.gradient(obs, P["nsteps"], ckpt, plain=plain)
    lam, mu, b = (P[k].astype(np.float64) for k in ("lam", "mu", "b"))
    _, rec = E.forward(P["env"], lam, mu, b, float(P["dt"]), P["ih"], P["nsteps"], P["src"],
                       P["comp"], P["wav"].astype(np.float64), P["rec"], 1, fp64=True)
    J64 = 0.5 * float(((rec - obs) ** 2).sum())
    g64 = E.adjoint(P["env"], lam, mu, b, float(P["dt"]), P["ih"], P["nsteps"], P["src"],
                    P["comp"], P["wav"].astype(np.float64), P["rec"], 1, rec - obs)
    say("ckpt", ckpt, "misfit", J, J64, "grad rel:", [rel_l2(g[k], g64[k]) for k in range(3)])
    assert abs(J - J64) <= 1e-4 * J64
    for k in range(3):
        assert rel_l2(g[k], g64[k]) < 1e-3


def test_gradient_checkpoint_invariance():
    # plain: recomputation replays the forward bit for bit, so the checkpoint
    # interval does not change the gradient at all; reverse sweep: the
    # reconstructed states differ from the forward ones by rounding only
    P = E.small_problem(env=(20, 18, 16), nsteps=20)
    obs = _obs(P)
    out, rev = [], []
    for ckpt in (1, 6, 20):
        with A.Context(P["env"], (1, 1, 1), 1, 4, A.BOUNDARY_FROZEN, 4) as c:
            with _shot(c, P) as s:
                out.append(s.gradient(obs, 20, ckpt, plain=True)[1])
                rev.append(s.gradient(obs, 20, ckpt)[1])
    for g in out[1:]:
        for k in range(3):
            assert bits_equal(g[k], out[0][k])
    for g in rev:
        for k in range(3):
            assert rel_l2(g[k], out[0][k]) < 1e-5, (k, rel_l2(g[k], out[0][k]))


def test_gradient_reverse_sweep_long_segment():
    # 240 steps reconstructed backwards without a checkpoint reset against the
    # plain algorithm's stored states: the drift of the reconstruction stays
    # at rounding level
    P = E.small_problem(env=(22, 20, 18), nsteps=240)
    obs = _obs(P)
    with A.Context(P["env"], (1, 1, 1), 1, 4, A.BOUNDARY_FROZEN, 4) as c:
        with _shot(c, P) as s:
            J0, g0 = s.gradient(obs, P["nsteps"], 48, plain=True)
            J1, g1 = s.gradient(obs, P["nsteps"], 240)
    say("long segment", J0, J1, [rel_l2(g1[k], g0[k]) for k in range(3)])
    assert J1 == J0
    for k in range(3):
        assert rel_l2(g1[k], g0[k]) < 1e-4


@pytest.mark.parametrize("mode", [1, 2])
def test_gradient_batch_uneven_shots(mode):
    # three shots of different sizes in one batched gradient: each equals its
    # single-shot reverse-sweep gradient bit for bit; the plain per-shot loop
    # (mode 0) agrees to rounding
    sizes = [(20, 18, 16), (24, 16, 20), (16, 22, 18)]
    probs = [E.small_problem(env=e, nsteps=26, seed=3 + i) for i, e in enumerate(sizes)]
    obs = [_obs(P) for P in probs]
    ctxs = [A.Context(P["env"], (1, 1, 1), 1, 4, A.BOUNDARY_FROZEN, 4) for P in probs]
    shots = [_shot(c, P) for c, P in zip(ctxs, probs)]
    try:
        single = [s.gradient(o, 26, 7) for s, o in zip(shots, obs)]
        mis, grads, _ = A.el_gradient_batch(shots, obs, 26, 7, mode)
        mis0, grads0, _ = A.el_gradient_batch(shots, obs, 26, 7, 0)
        for k in range(3):
            say("batch shot", k, mis[k], single[k][0], mis0[k])
            assert mis[k] == single[k][0]
            assert abs(mis0[k] - mis[k]) <= 1e-6 * mis[k]
            for i in range(3):
                assert bits_equal(grads[k][i], single[k][1][i])
                assert rel_l2(grads0[k][i], grads[k][i]) < 1e-5
    finally:
        for s in shots:
            s.close()
        for c in ctxs:
            c.close()


@pytest.mark.parametrize("mb", [2, 4, 6, 0])
def test_gradient_batch_memory_groups(mb):
    # a memory budget smaller than the batch runs the shots in groups: the
    # results equal the one-group batch bit for bit; a budget below one shot
    # is an error
    sizes = [(20, 18, 16), (24, 16, 20), (16, 22, 18), (18, 18, 18)]
    probs = [E.small_problem(env=e, nsteps=26, seed=13 + i) for i, e in enumerate(sizes)]
    obs = [_obs(P) for P in probs]
    ctxs = [A.Context(P["env"], (1, 1, 1), 1, 4, A.BOUNDARY_FROZEN, 4) for P in probs]
    shots = [_shot(c, P) for c, P in zip(ctxs, probs)]
    try:
        mis, grads, _ = A.el_gradient_batch(shots, obs, 26, 7, 2)
        # 2.6-3.2 MB per shot here: 2 MB fits none, 4 MB one shot per group,
        # 6 MB two, 0 the free memory (one group)
        ctxs[0].set_tuning(el_grad_mb=mb)
        say("gradient memory budget", mb, "MB")
        if mb == 2:
            with pytest.raises(A.B2RError):
                A.el_gradient_batch(shots, obs, 26, 7, 2)
            return
        mis2, grads2, ms = A.el_gradient_batch(shots, obs, 26, 7, 2, timed=True)
        assert ms > 0
        for k in range(len(shots)):
            assert mis2[k] == mis[k]
            for i in range(3):
                assert bits_equal(grads2[k][i], grads[k][i])
    finally:
        for s in shots:
            s.close()
        for c in ctxs:
            c.close()


def test_gradient_finite_differences():
    P = E.small_problem(env=(18, 16, 14), nsteps=24)
    obs = _obs(P)
    with A.Context(P["env"], (1, 1, 1), 1, 4, A.BOUNDARY_FROZEN, 4) as c:
        with _shot(c, P) as s:
            _, g = s.gradient(obs, P["nsteps"], 5)
    m = [P[k].astype(np.float64) for k in ("lam", "mu", "b")]

    def J(mm):
        _, rec = E.forward(P["env"], *mm, float(P["dt"]), P["ih"], P["nsteps"], P["src"],
                           P["comp"], P["wav"].astype(np.float64), P["rec"], 1, fp64=True)
        return 0.5 * float(((rec - obs) ** 2).sum())

    rng = np.random.default_rng(7)
    for trial in range(6):
        which = trial % 3
        dirn = rng.standard_normal(m[which].shape) * m[which]
        eps = 1e-5
        mp = [a.copy() for a in m]
        mn = [a.copy() for a in m]
        mp[which] += eps * dirn
        mn[which] -= eps * dirn
        fd = (J(mp) - J(mn)) / (2 * eps)
        an = float((g[which].astype(np.float64) * dirn).sum())
        say("FD", which, fd, an)
        assert abs(fd - an) <= 1e-3 * max(abs(fd), abs(an))


def test_autograd_wrapper_vp_vs_rho():
    # the torch wrapper's (vp, vs, rho) gradient against FP64 central
    # differences of the oracle misfit, on random directions
    import torch
    from b2r_amd.autograd import elastic_misfit
    from b2r_amd import workloads as W
    P = E.small_problem(env=(16, 14, 18), nsteps=20)
    obs = _obs(P)
    vp, vs, rho = (torch.tensor(P[k], requires_grad=True) for k in ("vp", "vs", "rho"))
    with A.Context(P["env"], (1, 1, 1), 1, 4, A.BOUNDARY_FROZEN, 4) as c:
        with _shot(c, P) as s:
            J = elastic_misfit(vp, vs, rho, s, obs, P["nsteps"], 6)
            J.backward()

    def J64(v, w, r):
        # medium in FP64 from FP64 inputs
        v, w, r = (np.asarray(a, np.float64) for a in (v, w, r))
        lam, mu, b = r * (v * v - 2 * w * w), r * w * w, 1.0 / r
        _, rec = E.forward(P["env"], lam, mu, b, float(P["dt"]), P["ih"], P["nsteps"], P["src"],
                           P["comp"], P["wav"].astype(np.float64), P["rec"], 1, fp64=True)
        return 0.5 * float(((rec - obs) ** 2).sum())

    m = [P[k].astype(np.float64) for k in ("vp", "vs", "rho")]
    grads = [vp.grad.numpy(), vs.grad.numpy(), rho.grad.numpy()]
    rng = np.random.default_rng(3)
    for which in range(3):
        d = rng.standard_normal(m[which].shape) * m[which]
        eps = 1e-6
        mp = [a.copy() for a in m]
        mn = [a.copy() for a in m]
        mp[which] += eps * d
        mn[which] -= eps * d
        fd = (J64(*mp) - J64(*mn)) / (2 * eps)
        an = float((grads[which] * d).sum())
        say("autograd", which, fd, an)
        assert abs(fd - an) <= 2e-3 * max(abs(fd), abs(an))


def test_adjoint_dot_product():
    # P: initial state -> receiver data (linear).  With obs = 0 the adjoint
    # state at t = 0 is P^T P x0, so <x0, P^T P x0> = |P x0|^2 = 2 J.
    P = E.small_problem(env=(18, 16, 20), nsteps=25)
    gx, gy, gz = P["env"]
    rng = np.random.default_rng(11)
    x0 = [rng.standard_normal((gz, gy, gx)).astype(np.float32) * np.float32(1e-3)
          for _ in range(9)]
    with A.Context(P["env"], (1, 1, 1), 1, 4, A.BOUNDARY_FROZEN, 4) as c:
        s = A.ElasticShot(c, float(P["dt"]), P["ih"])
        s.set_medium(P["lam"], P["mu"], P["b"])
        s.set_receivers(P["rec"], 1, P["nsteps"])
        s.reset()
        for k in range(9):
            s.set_state(k, x0[k])
        J, _, adj0 = s.gradient(np.zeros((P["nsteps"], len(P["rec"])), np.float32),
                                P["nsteps"], 6, from_current=True, adjoint_state=True)
        s.close()
    lhs = sum(float((x0[k].astype(np.float64) * adj0[k].astype(np.float64)).sum())
              for k in range(9))
    say("dot product", lhs, 2 * J)
    assert abs(lhs - 2 * J) <= 1e-4 * abs(2 * J)


def test_elastic_zchunk_tuning_and_autotune(tmp_path):
    # the z chunk changes the launch shape only: results are bit-identical;
    # the autotuner records its choice and leaves the shot reset
    P = E.small_problem(env=(40, 36, 44), nsteps=12)
    outs = []
    with A.Context(P["env"], (1, 1, 1), 1, 4, A.BOUNDARY_FROZEN, 4) as c:
        for kern in (1, 2):
            for zc in (0, 5, 16, 64):
                c.set_tuning(el_kernel=kern, el_zc=zc, el_step_zc=zc)
                with _shot(c, P) as s:
                    s.forward(P["nsteps"], 1)
                    outs.append((s.receivers(), s.state(1)))
        for rec, vy in outs[1:]:
            assert bits_equal(rec, outs[0][0]) and bits_equal(vy, outs[0][1])
        rec_file = str(tmp_path / "tune.txt")
        with _shot(c, P) as s:
            s.forward(3, 1)
            t = s.autotune(rec_file)
            assert t["el_zc"] in (0, 16, 24, 32, 48, 64, 96, 128, 256)
            assert t["el_step_zc"] in (0, 16, 24, 32, 48, 64, 96, 128, 256)
            assert t["el_kernel"] in (1, 2)
            assert not np.any(s.state(1))
            s.forward(P["nsteps"], 2)
            assert bits_equal(s.receivers(), outs[0][0])
        line = open(rec_file).read()
        assert "kind=el_step" in line and "best=el_kernel:" in line and ",el_step_zc:" in line


def test_fused_kernel_needs_closed_faces():
    # an open face keeps cells the fused step does not carry over: asking for
    # it there is an error, the automatic choice takes the separate kernels
    P = E.small_problem(env=(24, 20, 22), nsteps=4)
    with A.Context(P["env"], (1, 1, 1), 1, 4, A.BOUNDARY_FROZEN, 4) as c:
        c.set_frozen_faces(0x3F & ~1)
        with _shot(c, P) as s:
            c.set_tuning(el_kernel=1)
            with pytest.raises(A.B2RError):
                s.forward(2, 1)
            c.set_tuning(el_kernel=0)
            s.reset()
            s.forward(4, 2)
            got = s.receivers()
            c.set_tuning(el_kernel=2)
            s.reset()
            s.forward(4, 2)
            assert bits_equal(got, s.receivers())


def test_gradient_batch_larger_grids_agree_with_plain():
    # two 96^3-class shots with the bench's medium generator and a surface
    # receiver grid: the batched reverse sweep (graphs) and the plain loop
    # agree to rounding, with identical misfits
    from b2r_amd import workloads as W
    envs = [(96, 96, 96), (80, 96, 112)]
    nsteps = 40
    ctxs, shots, obs = [], [], []
    try:
        for i, env in enumerate(envs):
            vp, vs, rho = W.layered_medium(env, 20 + i)
            dt = W.cfl_dt(4500.0, 10.0, 0.4)
            c = A.Context(env, (1, 1, 1), 1, 4, A.BOUNDARY_FROZEN, 4)
            s = A.ElasticShot(c, dt, (0.1, 0.1, 0.1))
            s.set_medium(*W.lame(*W.perturbed_layers((vp, vs, rho), env, 30 + i)))
            s.set_sources([[env[0] // 2, 12, env[2] // 2]], [1],
                          W.ricker(10.0, dt, nsteps)[:, None] * np.float32(1e6))
            s.set_receivers(W.surface_receivers(env, n=16), 1, nsteps)
            s.forward(nsteps, 2)
            obs.append(s.receivers())
            s.set_medium(*W.lame(vp, vs, rho))
            ctxs.append(c)
            shots.append(s)
        mis_b, g_b, _ = A.el_gradient_batch(shots, obs, nsteps, 10, 2)
        mis_p, g_p, _ = A.el_gradient_batch(shots, obs, nsteps, 10, 0)
        for k in range(len(envs)):
            say("shot", k, mis_b[k], mis_p[k], [rel_l2(g_b[k][i], g_p[k][i]) for i in range(3)])
            assert mis_b[k] == mis_p[k] and mis_b[k] > 0
            for i in range(3):
                assert np.isfinite(g_b[k][i]).all()
                assert rel_l2(g_b[k][i], g_p[k][i]) < 1e-5
    finally:
        for s in shots:
            s.close()
        for c in ctxs:
            c.close()
