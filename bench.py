"""Benchmark of the B2R stencil hot path on one MI355X.

Headline (value): 3D Game of Life, 512^3 u8 cells Bernoulli(0.3) seed 1,
periodic, 2x2x2 subdomain blocks with R-deep halos exchanged on the device every
R generations, R in {1,2,4} (best reported), the R generations of a window in one
fused launch; --steps generations timed.  A step is one generation of the whole
512^3 grid; GLUP/s = cells * steps / time.

Also reported in the same JSON line ("also"): measured STREAM copy bandwidth;
heat delta1 at 768^3 (single pass as % of measured STREAM, and the fused R=2
window); heat delta4 768^3 with 8 Gaussian blobs on 2x2x2 blocks; heat delta1
128^3 next to the reference's own CPU code; the 16-shot elastic batch (plain
per-shot loop vs batched graph); the misfit gradient of 8 elastic 256^3 shots
(batched reverse sweep vs the plain per-shot loop); CPU baselines on the host
cores.  Launch shapes are autotuned per size and the
choices are appended to --tune-record.

Usage: python bench.py [--gpus 1] [--steps 1000] [--warmup 20] [--quick]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# the CPU baselines' OpenMP threads stay next to each other (read when the
# oracle's OpenMP runtime starts, so set before it is loaded)
os.environ.setdefault("OMP_PROC_BIND", "close")
# the CPUs this process may run on, read before an OpenMP runtime binds the
# main thread to one of them
try:
    CPUS_AVAILABLE = len(os.sched_getaffinity(0))
except (AttributeError, OSError):
    CPUS_AVAILABLE = None

METRIC = "GLUP/s & % HBM roofline, 1 MI355X; batched+adjoint vs plain at stated FP64 err"
HBM_PEAK_GBS = 8000.0
# VALU issue: one wave64 vector instruction per cycle per CU (4 SIMDs, 4 cycles
# per wave64 instruction each), 256 CUs at 2.4 GHz, in G wave instructions/s
VALU_PEAK_GINST = 256 * 2.4


def host_cpu():
    """The host CPU as the baselines see it: model, physical cores (distinct
    (physical id, core id) pairs of /proc/cpuinfo), logical CPUs, the CPUs this
    process may run on, and the OpenMP binding."""
    model, cores, logical = None, set(), 0
    phys = core = None
    try:
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "processor":
                logical += 1
            elif k == "model name" and model is None:
                model = v
            elif k == "physical id":
                phys = v
            elif k == "core id":
                core = v
                cores.add((phys, core))
    except OSError:
        pass
    return {"cpu_model": model, "physical_cores_host": len(cores) or None,
            "logical_cpus_host": logical or None, "cpus_available": CPUS_AVAILABLE,
            "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
            "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS")}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def read_traffic():
    """Per-launch HBM bytes of the dominant kernel from the committed rocprofv3
    --pmc summary (profiles/traffic.txt: 'key: value' lines), or None."""
    p = os.path.join(ROOT, "profiles", "traffic.txt")
    if not os.path.exists(p):
        return None
    d = {}
    for line in open(p):
        if ":" in line and not line.lstrip().startswith("#"):
            k, v = line.split(":", 1)
            d[k.strip()] = v.strip()
    return d


def record_choice(path, what, best, tried):
    """Append the measured choice of R (and of fusion) for a size to the
    autotuner's record file, in its line format."""
    if not path:
        return
    with open(path, "a") as f:
        f.write("%s best=R:%s glups=%s\n" % (
            what, best, ",".join("%s:%.1f" % (k, v) for k, v in tried.items())))


# ------------------------------------------------------------------ Game of Life
def gol_bench(A, R, steps, warmup, fused=1, zc=0, record=None, exchange=False):
    """steps generations of 512^3; kernel_ms = one launch of the dominant kernel
    (the fused R-generation kernel, or the single-generation kernel).  exchange:
    every window runs the pack/unpack boundary_updating on byte fields and the
    kernel reads its own halo (gol_exchange = 1)."""
    env = (512, 512, 512)
    with A.Context(env, (2, 2, 2), R, 1, A.BOUNDARY_PERIODIC, 1) as c:
        a, b = c.field(), c.field()
        c.init_bernoulli(a, 1, int(0.3 * 16777216))
        c.boundary_updating_direct([a])
        if exchange:
            c.set_tuning(gol_exchange=1)
        kind = 2 if fused else 1
        if zc:
            c.set_tuning(gol_zc=zc)
        else:
            c.autotune(kind, a, b, None, record)
        c.init_bernoulli(a, 1, int(0.3 * 16777216))
        if warmup:
            c.gol_run(a, b, warmup, fused=fused)
        c.init_bernoulli(a, 1, int(0.3 * 16777216))
        c.sync()
        t0 = time.perf_counter()
        _, ms = c.gol_run(a, b, steps, fused=fused, timed=True)
        c.sync()
        wall = time.perf_counter() - t0
        kiters = 50 if fused else 200
        kms = c.time_steps(kind, a, b, kiters) / kiters
        tune = c.get_tuning()
    cells = 512 ** 3
    gens = R if fused else 1
    # 256-cell block rows are whole words: the bit-sliced kernel reads its halo
    # from the neighbours, and runs on packed fields unless gol_bytes is set
    packed = bool(fused and tune["gol_tile"] in GOL_BIT_TILES and not tune["gol_bytes"] and
                  not tune["gol_exchange"])
    return dict(R=R, fused=fused, ms=ms, wall_s=wall, glups=cells * steps / (ms * 1e-3) / 1e9,
                kernel_ms=kms, gens_per_launch=gens,
                kernel_gbs_alg=2.0 * cells * gens / (kms * 1e-3) / 1e9,
                zc=tune["gol_zc"] if fused else tune["gol_step_zc"],
                tile=tune["gol_tile"] if fused else None, packed=packed)


# (output quads of 16 cells per tile row, rows incl. halo) of the gol_tile values
GOL_TILES = {1: (8, 32), 2: (8, 48), 3: (16, 56)}
# (output words of 32 cells per row, rows incl. halo, rows per thread, waves of
# whole block rows 1 or of whole periodic 512-cell grid rows 2) of the
# bit-sliced kernel
GOL_BIT_TILES = {0: (8, 64, 1, 0), 4: (8, 32, 1, 0), 5: (8, 64, 1, 0), 6: (8, 96, 1, 0),
                 7: (8, 40, 1, 0), 8: (8, 72, 1, 0), 9: (8, 72, 2, 0), 10: (8, 136, 2, 0),
                 11: (8, 72, 1, 1), 12: (8, 48, 1, 1), 13: (16, 72, 2, 2), 14: (16, 40, 1, 2),
                 15: (16, 64, 1, 2)}


# (gol_tile, gol_zc) profiled beside the tuner's choice at 512^3 on 2x2x2 blocks,
# and (heat_ry, heat_wy, heat_wx, heat_zc) beside it for heat delta1 768^3 and
# delta4 768^3 on 2x2x2 blocks
GOL_PROFILED = ((14, 16), (13, 8), (11, 16))
HEAT_PROFILED = {1: ((8, 4, 4, 256), (8, 3, 3, 384)), 4: ((8, 3, 3, 384),)}
# (el_step_zc, el_zc) of the elastic 256^3 profile
EL_PROFILED = (128, 32)


def gol_kernel_name(R, tile, packed=True):
    """The template of the run's middle launches: IO 3 packed to packed (the
    first launch of a run packs as it loads, IO 2, the last unpacks as it
    stores, IO 1), IO 0 on byte fields."""
    if tile in GOL_BIT_TILES:
        nw, nr, rt, xd = GOL_BIT_TILES[tile]
        return "gol_bits_kernel<%d, %d, %d, %d, %d, %d>" % (R, nw, nr, 3 if packed else 0, rt, xd)
    return "gol_fused_kernel<%d, %d, %d>" % ((R,) + GOL_TILES[tile])


# ------------------------------------------------------------------ heat
def stream_bw(A):
    nbytes = 4 << 30
    with A.Context((16, 16, 16), (1, 1, 1), 1, 1, 0, 4) as c:
        src = c.dev_alloc(nbytes)
        dst = c.dev_alloc(nbytes)
        c.stream_copy(dst, src, nbytes, 2)
        iters = 10
        ms = c.stream_copy(dst, src, nbytes, iters)
    return 2.0 * nbytes * iters / (ms * 1e-3) / 1e9


def heat_bench(A, env, blocks, R, D, k, steps, init, fused=0, record=None, warmup=2):
    p = A.heat_params(D, A.HEAT_F32, k)
    with A.Context(env, blocks, R, D, A.BOUNDARY_FROZEN, 4) as c:
        a, b = c.field(), c.field()
        init(c, a)
        c.heat_run(a, b, p, warmup * R, fused=fused)
        kind = 3 if fused else 0
        c.autotune(kind, a, b, p, record)
        init(c, a)
        c.sync()
        _, ms = c.heat_run(a, b, p, steps, fused=fused, timed=True)
        kiters = 10
        kms = c.time_steps(kind, a, b, kiters, p) / kiters
        tune = c.get_tuning()
    cells = env[0] * env[1] * env[2]
    per_launch = R if fused else 1
    return dict(R=R, fused=fused, steps=steps, ms=round(ms, 3),
                glups=round(cells * steps / (ms * 1e-3) / 1e9, 2),
                kernel_ms=round(kms, 4), steps_per_launch=per_launch,
                kernel_gbs_alg=round(8.0 * cells * per_launch / (kms * 1e-3) / 1e9, 1),
                tuning={"heat_wy": tune["heat_wy"], "heat_wx": tune["heat_wx"],
                        "heat_ry": tune["heat_ry"], "heat_zc": tune["heat_zc"]})


def blob_tables(env, nb=8, sigma=20.0, seed=2):
    import numpy as np
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0, [env[0], env[1], env[2]], size=(nb, 3))
    amp = rng.uniform(0.5, 2.0, nb)
    tabs = []
    for a in range(3):
        x = np.arange(env[a], dtype=np.float64)
        tabs.append(np.exp(-((x[None, :] - centres[:, a:a + 1]) ** 2) / (2 * sigma * sigma)).ravel())
    return amp, tabs


# ------------------------------------------------------------------ elastic
def elastic_shots(A, W, sizes, nsteps, seed=4):
    import numpy as np
    rng = np.random.default_rng(seed)
    ctxs, shots = [], []
    for i, env in enumerate(sizes):
        vp, vs, rho = W.layered_medium(env, seed + i)
        dt = W.cfl_dt(4500.0, 10.0, 0.4)
        c = A.Context(env, (1, 1, 1), 1, 4, A.BOUNDARY_FROZEN, 4)
        s = A.ElasticShot(c, dt, (0.1, 0.1, 0.1))
        s.set_medium(*W.lame(vp, vs, rho))
        src = [[int(rng.integers(env[0] // 4, 3 * env[0] // 4)), int(rng.integers(8, env[1] // 4)),
                int(rng.integers(env[2] // 4, 3 * env[2] // 4))]]
        s.set_sources(src, [1], W.ricker(10.0, dt, nsteps)[:, None] * np.float32(1e6))
        s.set_receivers(W.surface_receivers(env), 1, nsteps)
        ctxs.append(c)
        shots.append(s)
    return ctxs, shots


def elastic_batch_bench(A, W, nshots=16, nsteps=100, record=None):
    sizes = W.shot_sizes(nshots, seed=4)
    cells = sum(e[0] * e[1] * e[2] for e in sizes)
    ctxs, shots = elastic_shots(A, W, sizes, nsteps)
    try:
        t = A.el_autotune_batch(shots, record)
        out = {"shots": nshots, "steps": nsteps, "cells_per_step": cells,
               "sizes": ["%dx%dx%d" % e for e in sizes],
               "el_kernel": t["el_kernel"],
               "el_zc": t["el_step_zc" if t["el_kernel"] == 1 else "el_zc"]}
        for s in shots:
            s.reset()
        A.el_forward_batch(shots, 2, mode=2)
        for s in shots:
            s.reset()
        ms_b = A.el_forward_batch(shots, nsteps, mode=2, timed=True)
        t0 = time.perf_counter()
        ms_p = 0.0
        for s in shots:  # the plain per-shot loop in the reference's structure
            s.reset()
            ms_p += s.forward(nsteps, 0, timed=True)
        wall_p = time.perf_counter() - t0
        out.update(batched_graph_ms=round(ms_b, 2), plain_loop_ms=round(ms_p, 2),
                   plain_loop_wall_s=round(wall_p, 2), stated_steps=2000,
                   batched_ms_per_step=round(ms_b / nsteps, 4),
                   plain_ms_per_step=round(ms_p / nsteps, 4),
                   batched_glups=round(cells * nsteps / (ms_b * 1e-3) / 1e9, 2),
                   plain_glups=round(cells * nsteps / (ms_p * 1e-3) / 1e9, 2),
                   speedup_batched_vs_plain=round(ms_p / ms_b, 3),
                   batched_gbs_84=round(84.0 * cells * nsteps / (ms_b * 1e-3) / 1e9, 1))
        return out
    finally:
        for s in shots:
            s.close()
        for c in ctxs:
            c.close()


def elastic_gradient_bench(A, W, nshots=8, nsteps=300, ckpt=50, record=None,
                           modes=((2, "batched_graph"), (1, "batched"), (0, "plain_loop"))):
    """Misfit gradients of nshots 256^3 shots over one layered medium: 64x64
    surface receivers of v_y, observed data from the per-layer 5%-perturbed
    model, checkpoints every ckpt steps.  The batched reverse sweep (all shots
    per launch, HIP graphs) against the plain per-shot loop."""
    import numpy as np
    env = (256, 256, 256)
    cells = env[0] * env[1] * env[2]
    vp, vs, rho = W.layered_medium(env, 5)
    true_m = W.lame(vp, vs, rho)
    obs_m = W.lame(*W.perturbed_layers((vp, vs, rho), env, 6))
    ctxs, shots = elastic_shots(A, W, [env] * nshots, nsteps, seed=9)
    try:
        for s in shots:
            s.autotune(record)
            s.set_medium(*obs_m)
            s.reset()
        A.el_forward_batch(shots, nsteps, mode=2)
        obs = [s.receivers() for s in shots]
        for s in shots:
            s.set_medium(*true_m)
        groups, gbytes = A.el_gradient_plan(shots, nsteps, ckpt)
        g1500, g1500b = A.el_gradient_plan(shots, 1500, ckpt)
        out = {"shots": nshots, "grid": "256^3", "steps": nsteps, "ckpt": ckpt,
               "receivers": "64x64 surface, v_y", "stated_steps": 1500,
               "groups": groups, "group_max_GB": round(gbytes / 1e9, 1),
               "groups_at_stated_steps": g1500, "group_max_GB_at_stated_steps":
                   round(g1500b / 1e9, 1)}
        res = {}
        for mode, name in modes:
            t0 = time.perf_counter()
            mis, grads, ms = A.el_gradient_batch(shots, obs, nsteps, ckpt, mode, timed=True)
            wall = time.perf_counter() - t0
            res[name] = (mis, grads)
            out[name + "_ms"] = round(ms, 1)
            out[name + "_wall_s"] = round(wall, 2)
            out[name + "_shot_glups"] = round(nshots * cells * nsteps / (ms * 1e-3) / 1e9, 3)
            out[name + "_ms_per_shot_step"] = round(ms / (nshots * nsteps), 4)
            log(f"  gradient {name}: {ms:.1f} ms")
        if "plain_loop" not in res:  # the batched path alone (a profiling run)
            out["misfit"] = [round(m, 6) for m in res["batched_graph"][0]]
            return out
        out["speedup_batched_vs_plain"] = round(out["plain_loop_ms"] / out["batched_graph_ms"], 3)
        mb, gb_ = res["batched_graph"]
        mp, gp = res["plain_loop"]
        out["misfit"] = [round(m, 6) for m in mb]
        out["misfit_max_rel_diff_batched_vs_plain"] = float(max(
            abs(a - b) / abs(b) for a, b in zip(mb, mp)))
        out["grad_max_rel_l2_batched_vs_plain"] = float(max(
            np.linalg.norm((a.astype(np.float64) - b).ravel()) / np.linalg.norm(b.ravel().astype(np.float64))
            for ga, gp_ in zip(gb_, gp) for a, b in zip(ga, gp_)))
        out["plain_definition"] = ("per shot: stress + vx + vy + vz launches and a device sync per "
                                   "step, checkpoints, each segment recomputed storing its states, "
                                   "adjoint kernels and a sync per step")
        return out
    finally:
        for s in shots:
            s.close()
        for c in ctxs:
            c.close()


# ------------------------------------------------------------------ CPU baselines
def cpu_gol_baseline(budget_s=10.0):
    """The oracle's Game of Life (C, OpenMP over z) on the host cores: 512^3
    periodic, as many generations as fit the time budget."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O
    u = O.bernoulli((512, 512, 512), 1, 0.3)
    gens = 0
    t0 = time.perf_counter()
    while True:
        u = O.gol(u, 1, True)
        gens += 1
        el = time.perf_counter() - t0
        if el >= budget_s or gens >= 1000:
            break
    return dict(value=round(512 ** 3 * gens / el / 1e9, 4), unit="GLUP/s",
                cores=O.lib.orc_num_threads(), kind="port",
                sample=f"512^3 periodic Game of Life, {gens} generations ({el:.1f} s), "
                       "oracle C code, OpenMP over z"), u, gens


def cpu_heat_reference(steps=100):
    """The reference's own cpu_fdm_heat_diffuse1 (heat-3D.cpp:48-62, built from
    its sources into oracle/_ref) on 128^3 + zero frame, 100 steps."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O
    if not O.ref_available():
        return None
    u = O.frame(O.uniform((128, 128, 128), 0), 1)
    v = u.copy()
    t0 = time.perf_counter()
    r = O.ref_lib()
    pu, pv = u.ctypes.data, v.ctypes.data
    for _ in range(steps):
        r.ref_cpu_fdm_heat_diffuse1(pu, pv, 130, 130, 130, 0)
        pu, pv = pv, pu
    el = time.perf_counter() - t0
    return dict(value=round(128 ** 3 * steps / el / 1e9, 4), unit="GLUP/s",
                cores=O.lib.orc_num_threads(), kind="reference",
                sample=f"128^3 heat delta1, {steps} steps ({el:.2f} s)"), u


def heat_reference_check(A, u_ref, steps=100):
    """The GPU REF-mode path on the same 128^3 problem against the reference's
    own CPU result (bitwise) and against the FP64 oracle."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O
    p = A.heat_params(1, A.HEAT_REF)
    with A.Context((128, 128, 128), (1, 1, 1), 1, 1, A.BOUNDARY_FROZEN, 4) as c:
        a, b = c.field(), c.field()
        c.init_uniform(a, 0)
        res, _ = c.heat_run(a, b, p, steps)
        g = c.download(res)
    ref = O.unframe(u_ref, 1)
    u64 = O.unframe(O.heat_f64(O.frame(O.uniform((128, 128, 128), 0).astype(np.float64), 1), 1,
                               [0.01, 0.01, 0.04], steps, False), 1)

    def rel(x):
        x = x.astype(np.float64)
        return float(np.linalg.norm((x - u64).ravel()) / np.linalg.norm(u64.ravel()))
    return {"bit_exact_vs_reference_cpu": bool(np.array_equal(g.view(np.uint32),
                                                              ref.view(np.uint32))),
            "gpu_rel_l2_vs_fp64": rel(g), "gpu_max_abs_vs_fp64": float(np.abs(g - u64).max()),
            "reference_cpu_rel_l2_vs_fp64": rel(ref)}


def cpu_elastic_baseline(steps=10):
    """The oracle's FP32 elastic step (C, OpenMP over z) on a 128^3 layered
    medium, a bounded sample of the per-shot work."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np
    import _elastic as E
    from b2r_amd import workloads as W
    env = (128, 128, 128)
    vp, vs, rho = W.layered_medium(env, 0)
    lam, mu, b = W.lame(vp, vs, rho)
    dt = np.float32(W.cfl_dt(4500.0, 10.0))
    t0 = time.perf_counter()
    E.forward(env, lam, mu, b, dt, (np.float32(0.1),) * 3, steps, [[64, 20, 64]], [1],
              np.ones((steps, 1), np.float32))
    el = time.perf_counter() - t0
    import _oracle as O
    return dict(value=round(128 ** 3 * steps / el / 1e9, 5), unit="GLUP/s",
                cores=O.lib.orc_num_threads(), kind="port",
                sample=f"128^3 elastic FP32, {steps} steps ({el:.2f} s), oracle C code, OpenMP")


# ------------------------------------------------------------------ accuracy
def heat_eigenmode_error(A, env, D, k, steps, blocks, R, fused, periodic):
    """FP32 GPU heat against the closed form of an exact eigenmode (FP64):
    frozen zero frame, delta1: u0 = prod sin(pi m_a i_a / (N_a + 1)); periodic,
    any D: u0 = prod cos(2 pi m_a i_a / N_a).  Then u_n = lam^n u0 with lam
    from the FP32 taps and diffusion numbers the kernel uses."""
    import numpy as np
    gx, gy, gz = env
    k32 = [float(np.float32(k))] * 3
    (sx, sy, sz), lam = heat_eigenmode(env, D, k32, periodic)
    plane = np.outer(sy, sx)
    u0 = np.empty((gz, gy, gx), np.float32)
    for z in range(gz):
        u0[z] = (sz[z] * plane).astype(np.float32)
    p = A.heat_params(D, A.HEAT_F32, k32)
    bnd = A.BOUNDARY_PERIODIC if periodic else A.BOUNDARY_FROZEN
    with A.Context(env, blocks, R, D, bnd, 4) as c:
        a, b = c.field(), c.field()
        c.upload(a, u0)
        res, _ = c.heat_run(a, b, p, steps, fused=fused)
        u = c.download(res)
    del u0
    lamn = lam ** steps
    num = den = mx = 0.0
    for z in range(gz):
        ex = (lamn * sz[z]) * plane
        d = u[z].astype(np.float64) - ex
        num += float((d * d).sum())
        den += float((ex * ex).sum())
        mx = max(mx, float(np.abs(d).max()))
    return {"rel_l2_vs_fp64": num ** 0.5 / den ** 0.5, "max_abs_vs_fp64": mx,
            "max_abs_over_peak": mx / abs(lamn)}


def heat_eigenmode(env, D, k, periodic):
    """Separable exact eigenmode of the heat step on the owned grid (x, y, z
    factors) and its eigenvalue, with the FP32 taps and diffusion numbers k
    the kernel uses: frozen zero frame (delta1 only) sin(pi i/(N+1)),
    periodic cos(2 pi i/N) for any D."""
    import numpy as np
    c32 = [float(np.float32(v)) for v in _heat_taps(D)]
    if periodic:
        th = [2.0 * np.pi / n for n in env]
        modes = [np.cos(t * np.arange(n)) for t, n in zip(th, env)]
        lam = 1.0 + sum(ka * (c32[0] + 2.0 * sum(c32[j] * np.cos(j * t) for j in range(1, D + 1)))
                        for ka, t in zip(k, th))
    else:
        assert D == 1
        modes = [np.sin(np.pi * np.arange(1, n + 1) / (n + 1)) for n in env]
        lam = 1.0 + sum(ka * (2.0 * np.cos(np.pi / (n + 1)) - 2.0) for ka, n in zip(k, env))
    return modes, lam


def _heat_taps(D):
    return {1: (-2.0, 1.0), 2: (-2.5, 4 / 3, -1 / 12), 3: (-49 / 18, 1.5, -3 / 20, 1 / 90),
            4: (-205 / 72, 1.6, -0.2, 8 / 315, -1 / 560)}[D]


def gol_bitexact_check(A, grid_cpu, gens, R=4):
    """The GPU path (512^3, 2x2x2, R-generation fused launches) against the
    oracle's result after the same generations."""
    import numpy as np
    env = (512, 512, 512)
    with A.Context(env, (2, 2, 2), R, 1, A.BOUNDARY_PERIODIC, 1) as c:
        a, b = c.field(), c.field()
        c.init_bernoulli(a, 1, int(0.3 * 16777216))
        res, _ = c.gol_run(a, b, gens, fused=1)
        g = c.download(res)
    diff = int(np.count_nonzero(g != grid_cpu))
    return {"generations": gens, "R": R, "cells_differing": diff, "bit_exact": diff == 0,
            "population": int(g.sum(dtype=np.int64))}


def elastic_checks(A, W):
    """Elastic forward FP32 (GPU) against the oracle's FP32 (bitwise) and FP64
    forward, and the batched gradient against the FP64 adjoint, on small grids."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np
    import _elastic as E
    out = {}
    P = E.small_problem(env=(64, 56, 48), nsteps=200)
    with A.Context(P["env"], (1, 1, 1), 1, 4, A.BOUNDARY_FROZEN, 4) as c:
        s = A.ElasticShot(c, float(P["dt"]), P["ih"])
        s.set_medium(P["lam"], P["mu"], P["b"])
        s.set_sources(P["src"], P["comp"], P["wav"])
        s.set_receivers(P["rec"], 1, P["nsteps"])
        s.forward(P["nsteps"], 2)
        rec = s.receivers()
        vy = s.state(1)
        s.close()
    st32, rec32 = E.forward(P["env"], P["lam"], P["mu"], P["b"], P["dt"], P["ih"], P["nsteps"],
                            P["src"], P["comp"], P["wav"], P["rec"], 1)
    lam, mu, b = (P[k].astype(np.float64) for k in ("lam", "mu", "b"))
    st64, rec64 = E.forward(P["env"], lam, mu, b, float(P["dt"]), P["ih"], P["nsteps"], P["src"],
                            P["comp"], P["wav"].astype(np.float64), P["rec"], 1, fp64=True)

    def rel(a, r):
        a, r = np.asarray(a, np.float64), np.asarray(r, np.float64)
        return float(np.linalg.norm((a - r).ravel()) / np.linalg.norm(r.ravel()))

    out["forward_64x56x48_200_steps"] = {
        "bitwise_vs_fp32_oracle": bool(np.array_equal(rec.view(np.uint32), rec32.view(np.uint32))
                                       and np.array_equal(vy.view(np.uint32),
                                                          st32[1].view(np.uint32))),
        "receivers_rel_l2_vs_fp64": rel(rec, rec64),
        "receivers_max_abs_vs_fp64": float(np.abs(rec.astype(np.float64) - rec64).max()),
        "vy_rel_l2_vs_fp64": rel(vy, st64[1]),
        "vy_max_abs_vs_fp64": float(np.abs(vy.astype(np.float64) - st64[1]).max())}
    P = E.small_problem(env=(24, 22, 20), nsteps=60)
    _, obs = E.forward(P["env"], P["lam"] * np.float32(1.03), P["mu"] * np.float32(0.98), P["b"],
                       P["dt"], P["ih"], P["nsteps"], P["src"], P["comp"], P["wav"], P["rec"], 1)
    with A.Context(P["env"], (1, 1, 1), 1, 4, A.BOUNDARY_FROZEN, 4) as c:
        s = A.ElasticShot(c, float(P["dt"]), P["ih"])
        s.set_medium(P["lam"], P["mu"], P["b"])
        s.set_sources(P["src"], P["comp"], P["wav"])
        s.set_receivers(P["rec"], 1, P["nsteps"])
        mis, grads, _ = A.el_gradient_batch([s], [obs], P["nsteps"], 10, 2)
        s.close()
    lam, mu, b = (P[k].astype(np.float64) for k in ("lam", "mu", "b"))
    _, r64 = E.forward(P["env"], lam, mu, b, float(P["dt"]), P["ih"], P["nsteps"], P["src"],
                       P["comp"], P["wav"].astype(np.float64), P["rec"], 1, fp64=True)
    g64 = E.adjoint(P["env"], lam, mu, b, float(P["dt"]), P["ih"], P["nsteps"], P["src"],
                    P["comp"], P["wav"].astype(np.float64), P["rec"], 1, r64 - obs)
    J64 = 0.5 * float(((r64 - obs) ** 2).sum())
    out["gradient_24x22x20_60_steps_ckpt10"] = {
        "misfit_rel_vs_fp64": abs(mis[0] - J64) / J64,
        "g_lambda_rel_l2_vs_fp64_adjoint": rel(grads[0][0], g64[0]),
        "g_mu_rel_l2_vs_fp64_adjoint": rel(grads[0][1], g64[1]),
        "g_b_rel_l2_vs_fp64_adjoint": rel(grads[0][2], g64[2])}
    return out


def profile_kernels(A):
    """A short fixed sequence for rocprofv3 counter runs, each kernel at the
    launch shapes the bench runs it with: a copy kernel of known size (1 GiB
    read + 1 GiB written per launch, the scale for FETCH_SIZE / WRITE_SIZE),
    then, after the same autotuning the bench does, 5 launches each of the
    tuner's choice and of the other shapes it picks between from run to run
    (GOL_PROFILED, HEAT_PROFILED) for the fused Game of Life kernel (512^3,
    2x2x2, periodic, R = 1, 2, 4, on the bit-packed fields the runs use), the
    heat delta1 step kernel (768^3) and the heat delta4 step kernel (768^3 on
    2x2x2 blocks), then 5 elastic steps (256^3) with the fused step kernel and
    with the stage kernels, at the z chunks EL_PROFILED.  The tuner's own
    launches come first, so the last 5 dispatches of each kernel template and
    grid are the profiled ones; each launch set is logged as a JSON line on
    stderr with the kernel's name."""
    from b2r_amd import workloads as W
    nbytes = 1 << 30
    with A.Context((16, 16, 16), (1, 1, 1), 1, 1, 0, 4) as c:
        src, dst = c.dev_alloc(nbytes), c.dev_alloc(nbytes)
        log("stream_copy 1 GiB x3")
        c.stream_copy(dst, src, nbytes, 3)
    for R in (1, 2, 4):
        with A.Context((512, 512, 512), (2, 2, 2), R, 1, A.BOUNDARY_PERIODIC, 1) as c:
            a, b = c.field(), c.field()
            c.init_bernoulli(a, 1, int(0.3 * 16777216))
            c.boundary_updating_direct([a])
            t = c.autotune(2, a, b, None, None)
            choice = (t["gol_tile"], t["gol_zc"])
            # the tuner's choice, then the shapes it picks between from run to
            # run (within a few per cent of each other), so that the counters
            # cover the bench's choice whichever it is
            for tile, zc in [choice] + [x for x in GOL_PROFILED if x != choice]:
                c.set_tuning(gol_tile=tile, gol_zc=zc)
                c.init_bernoulli(a, 1, int(0.3 * 16777216))
                packed = tile in GOL_BIT_TILES and not t["gol_bytes"]
                log(json.dumps({"profile": "gol_R%d" % R, "kernel": gol_kernel_name(R, tile, packed),
                                "gol_tile": tile, "gol_zc": zc, "chosen": (tile, zc) == choice,
                                "packed": packed, "launches": 5}))
                c.time_steps(2, a, b, 5)
    for name, env, blocks, D, k in (("heat_D1_768", (768, 768, 768), (1, 1, 1), 1, 0.1),
                                    ("heat_D4_768_2x2x2", (768, 768, 768), (2, 2, 2), 4, 0.05)):
        prm = A.heat_params(D, A.HEAT_F32, (k, k, k))
        with A.Context(env, blocks, 1, D, 0, 4) as c:
            a, b = c.field(), c.field()
            c.init_uniform(a, 0)
            c.heat_run(a, b, prm, 2)
            t = c.autotune(0, a, b, prm, None)
            choice = (t["heat_ry"], t["heat_wy"] or 4, t["heat_wx"] or 1, t["heat_zc"])
            # the tuner's choice, then the shapes it picks between from run to run
            for ry, wy, wx, zc in [choice] + [x for x in HEAT_PROFILED[D] if x != choice]:
                c.set_tuning(heat_ry=ry, heat_wy=wy, heat_wx=wx, heat_zc=zc)
                log(json.dumps({"profile": name, "kernel": "heat_step_kernel<%d, 0, %d, %d, %d>" % (
                    D, ry, wy, wx), "heat_ry": ry, "heat_wy": wy, "heat_wx": wx, "heat_zc": zc,
                    "chosen": (ry, wy, wx, zc) == choice, "launches": 5}))
                c.time_steps(0, a, b, 5, prm)
    env = (256, 256, 256)
    vp, vs, rho = W.layered_medium(env, 0)
    with A.Context(env, (1, 1, 1), 1, 4, 0, 4) as c:
        s = A.ElasticShot(c, W.cfl_dt(float(vp.max()), 10.0), (0.1, 0.1, 0.1))
        s.set_medium(*W.lame(vp, vs, rho))
        t = s.autotune(None)
        first = t["el_kernel"]
        # both kernels at fixed z chunks (the tuner's picks vary between runs
        # by a few per cent), so every counter pass profiles the same launches
        c.set_tuning(el_step_zc=EL_PROFILED[0], el_zc=EL_PROFILED[1])
        for kern in (first, 3 - first):
            c.set_tuning(el_kernel=kern)
            log(json.dumps({"profile": "elastic_256", "kernel": "el_step_kernel" if kern == 1
                            else "el_stress_kernel + el_velocity_kernel<7>",
                            "el_step_zc": EL_PROFILED[0], "el_zc": EL_PROFILED[1],
                            "tuner": {"el_kernel": first, "el_step_zc": t["el_step_zc"],
                                      "el_zc": t["el_zc"]}, "steps": 5}))
            s.forward(5, 1)
        s.close()
    print(json.dumps({"profile_kernels": "done"}), flush=True)


def heat_sections(A, bw, rec=None):
    """Heat delta1 768^3 (single pass and fused R=2) and delta4 768^3 on 2x2x2
    blocks (R and fusion tried, best kept), as percent of the STREAM copy bw."""
    also = {}
    env = (768, 768, 768)
    log("heat delta1 768^3 single pass")
    h1 = heat_bench(A, env, (1, 1, 1), 1, 1, (0.1, 0.1, 0.1), 100,
                    lambda c, f: c.init_uniform(f, 0), fused=0, record=rec)
    h1["kernel_pct_of_stream"] = round(100 * h1["kernel_gbs_alg"] / bw, 1)
    h1["run_pct_of_stream"] = round(100 * h1["glups"] * 8 / bw, 1)
    also["heat_delta1_768_single_pass"] = h1
    log(json.dumps(h1))
    log("heat delta1 768^3 fused R=2")
    h1f = heat_bench(A, env, (1, 1, 1), 2, 1, (0.1, 0.1, 0.1), 100,
                     lambda c, f: c.init_uniform(f, 0), fused=1, record=rec)
    h1f["effective_pct_of_stream"] = round(100 * h1f["glups"] * 8 / bw, 1)
    also["heat_delta1_768_fused_R2"] = h1f
    log(json.dumps(h1f))
    amp, (ex, ey, ez) = blob_tables(env)
    h4runs = []
    for R4, f4 in ((1, 0), (1, 1), (2, 0), (2, 1)):
        log(f"heat delta4 768^3 blobs 2x2x2, 500 steps, R={R4} fused={f4}")
        h4 = heat_bench(A, env, (2, 2, 2), R4, 4, (0.05, 0.05, 0.05), 500,
                        lambda c, f: c.init_blobs(f, amp, ex, ey, ez), fused=f4, record=rec)
        h4["run_pct_of_stream"] = round(100 * h4["glups"] * 8 / bw, 1)
        h4["kernel_pct_of_stream"] = round(100 * h4["kernel_gbs_alg"] / bw, 1)
        h4runs.append(h4)
        log(json.dumps(h4))
    also["heat_delta4_768_2x2x2"] = max(h4runs, key=lambda r: r["glups"])
    record_choice(rec, "kind=heat_R S=768x768x768 blocks=2x2x2 heat_D=4",
                  "%d%s" % (also["heat_delta4_768_2x2x2"]["R"],
                            "_fused" if also["heat_delta4_768_2x2x2"]["fused"] else ""),
                  {"%d%s" % (r["R"], "_fused" if r["fused"] else ""): r["glups"]
                   for r in h4runs})
    also["heat_delta4_768_2x2x2_tried"] = [
        {k: r[k] for k in ("R", "fused", "glups", "tuning")} for r in h4runs]
    return also


def hybrid_bench(A, nsteps=40):
    """Hybrid CPU+GPU heat (the CPU owns the top cpu_planes z planes, joint
    layers exchanged over PCIe every R steps) against the GPU-only run: heat
    delta1 256^3, one block, frozen faces, R = 1 and 4, CPU shares of h and 4h
    planes (h = R*D).  Wall time of whole runs (the hybrid run synchronises with
    the host); the results are compared bit for bit."""
    import numpy as np
    env, k = (256, 256, 256), (0.1, 0.1, 0.1)
    p = A.heat_params(1, A.HEAT_F32, k)
    nthreads = CPUS_AVAILABLE or 0
    out = {"workload": "heat delta1 256^3 F32, one block, frozen faces, %d steps" % nsteps,
           "cpu_threads": nthreads, "runs": []}
    for R in (1, 4):
        with A.Context(env, (1, 1, 1), R, 1, A.BOUNDARY_FROZEN, 4) as c:
            a, b = c.field(), c.field()
            c.init_uniform(a, 0)
            c.heat_run(a, b, p, 2 * R)  # warm-up, tunes nothing (the default shape)
            c.init_uniform(a, 0)
            c.sync()
            t0 = time.perf_counter()
            res, _ = c.heat_run(a, b, p, nsteps)
            c.sync()
            gpu_ms = (time.perf_counter() - t0) * 1e3
            want = c.download(res)
            for planes in (R, 4 * R):
                c.init_uniform(a, 0)
                c.heat_run_hybrid(a, b, p, 2 * R, planes, nthreads)  # warm-up
                c.init_uniform(a, 0)
                c.sync()
                t0 = time.perf_counter()
                res, _ = c.heat_run_hybrid(a, b, p, nsteps, planes, nthreads)
                c.sync()
                hyb_ms = (time.perf_counter() - t0) * 1e3
                got = c.download(res)
                out["runs"].append({
                    "R": R, "cpu_planes": planes, "cpu_share": round(planes / env[2], 4),
                    "gpu_only_ms_per_step": round(gpu_ms / nsteps, 4),
                    "hybrid_ms_per_step": round(hyb_ms / nsteps, 4),
                    "hybrid_over_gpu_only": round(hyb_ms / gpu_ms, 2),
                    "bitwise_equal_to_gpu_only": bool(np.array_equal(got.view(np.uint32),
                                                                     want.view(np.uint32)))})
                log(json.dumps(out["runs"][-1]))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--quick", action="store_true", help="headline only, no extras")
    ap.add_argument("--cpu-budget", type=float, default=10.0)
    ap.add_argument("--elastic-steps", type=int, default=100)
    ap.add_argument("--gradient-steps", type=int, default=300)
    ap.add_argument("--tune-record", default="",
                    help="append the autotuner's choices to this file")
    ap.add_argument("--profile-kernels", action="store_true",
                    help="run the fixed kernel sequence used by the rocprofv3 counter runs")
    ap.add_argument("--only", choices=["elastic-gradient", "elastic-gradient-batched",
                                       "elastic-batch", "heat", "elastic-full", "hybrid"],
                    help="run one section alone and print its JSON (for profiling); "
                         "elastic-gradient-batched: the batched gradient alone, 4 shots of "
                         "256^3; elastic-full: the 16-shot forward over 2000 steps and the "
                         "16-shot 256^3 gradient over 1500 steps (checkpoints every 50), "
                         "batched against the plain per-shot loop")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the halo "
              "exchange between GPUs is designed but not built", file=sys.stderr)
        sys.exit(2)

    from b2r_amd import _abi as A  # raises if the HIP library is missing
    from b2r_amd import workloads as W

    if args.profile_kernels:
        profile_kernels(A)
        return
    if args.only == "elastic-gradient":
        print(json.dumps(elastic_gradient_bench(A, W, 8, args.gradient_steps)), flush=True)
        return
    if args.only == "elastic-gradient-batched":
        print(json.dumps(elastic_gradient_bench(A, W, 4, args.gradient_steps,
                                                modes=((2, "batched_graph"),))), flush=True)
        return
    if args.only == "elastic-batch":
        print(json.dumps(elastic_batch_bench(A, W, 16, args.elastic_steps)), flush=True)
        return
    if args.only == "elastic-full":
        log("elastic 16-shot batch, 2000 steps")
        fwd = elastic_batch_bench(A, W, 16, 2000)
        log(json.dumps(fwd))
        log("elastic gradient, 16 shots of 256^3, 1500 steps")
        grad = elastic_gradient_bench(A, W, 16, 1500, 50,
                                      modes=((2, "batched_graph"), (0, "plain_loop")))
        print(json.dumps({"elastic_16_shots_2000_steps": fwd,
                          "elastic_gradient_16x256_1500_steps": grad}), flush=True)
        return
    if args.only == "hybrid":
        print(json.dumps(hybrid_bench(A)), flush=True)
        return
    if args.only == "heat":
        bw = stream_bw(A)
        out = heat_sections(A, bw)
        out["stream_copy_GBs"] = round(bw, 1)
        print(json.dumps(out), flush=True)
        return
    rec = args.tune_record or None

    runs = []
    for R in (1, 2, 4):
        log(f"gol 512^3 R={R} fused")
        runs.append(gol_bench(A, R, args.steps, args.warmup, fused=1, record=rec))
        log(json.dumps(runs[-1]))
    best = max(runs, key=lambda r: r["glups"])
    r1 = next(r for r in runs if r["R"] == 1)
    record_choice(rec, "kind=gol_R S=512x512x512 blocks=2x2x2 periodic", best["R"],
                  {r["R"]: r["glups"] for r in runs})
    log("gol 512^3 R=3 fused (outside the configured R set, reported only)")
    r3 = gol_bench(A, 3, args.steps, args.warmup, fused=1, record=rec)
    log(json.dumps(r3))
    cells = 512 ** 3
    log("gol 512^3 R=1 step-by-step (plain)")
    plain = gol_bench(A, 1, args.steps, args.warmup, fused=0)
    log(json.dumps(plain))
    tr = read_traffic() or {}
    # counters of the same R, tile shape and field form (the profile sequence
    # runs the autotuner's choices); null where they were taken at another one
    form = "packed" if best["packed"] else "bytes"
    tpre = f"gol_fused_R{best['R']}_tile{best['tile']}_zc{best['zc']}_{form}"
    traffic = float(tr[tpre + "_bytes_per_launch"]) if tpre + "_bytes_per_launch" in tr else None
    valu = (float(tr[tpre + "_valu_insts_per_launch"])
            if tpre + "_valu_insts_per_launch" in tr else None)
    kern_s = best["kernel_ms"] * 1e-3
    valu_ach = valu / kern_s / 1e9 if valu else None
    hbm_ach = traffic / kern_s / 1e9 if traffic else None
    log("gol 512^3 R=1,2,4 fused with the pack/unpack exchange every window")
    xruns = []
    for R in (1, 2, 4):
        xruns.append(gol_bench(A, R, args.steps, args.warmup, fused=1, record=rec, exchange=True))
        log(json.dumps(xruns[-1]))

    also = {"gol_512_R3_fused": {k: r3[k] for k in ("glups", "kernel_ms", "zc")}}
    heat_ref_grid = None
    if not args.quick:
        log("stream")
        bw = stream_bw(A)
        also["stream_copy_GBs"] = round(bw, 1)
        also.update(heat_sections(A, bw, rec))
        h0 = heat_bench(A, (128, 128, 128), (1, 1, 1), 1, 1, (0.1, 0.1, 0.1), 100,
                        lambda c, f: c.init_uniform(f, 0))
        also["heat_delta1_128"] = h0
        ref = cpu_heat_reference()
        if ref:
            also["heat_delta1_128_cpu_reference"], heat_ref_grid = ref
        log("hybrid CPU+GPU heat")
        also["heat_hybrid_256"] = hybrid_bench(A)
        log("elastic 16-shot batch")
        also["elastic_16_shots"] = elastic_batch_bench(A, W, 16, args.elastic_steps, rec)
        log(json.dumps(also["elastic_16_shots"]))
        log("elastic gradient, 8 shots of 256^3")
        also["elastic_gradient_8x256"] = elastic_gradient_bench(A, W, 8, args.gradient_steps,
                                                                record=rec)
        log(json.dumps(also["elastic_gradient_8x256"]))
        log("elastic gradient, 16 shots of 256^3")
        also["elastic_gradient_16x256"] = elastic_gradient_bench(
            A, W, 16, args.gradient_steps, record=rec,
            modes=((2, "batched_graph"), (0, "plain_loop")))
        log(json.dumps(also["elastic_gradient_16x256"]))
        also["elastic_cpu_port"] = cpu_elastic_baseline()

    # cpu_baseline leg: the oracle timed on the host cores, and used as the
    # checker of the GPU results
    log("cpu baseline")
    cpu, grid_cpu, gens = cpu_gol_baseline(args.cpu_budget)
    cpu.update(host_cpu())
    acc = {"gol_512_vs_oracle": gol_bitexact_check(A, grid_cpu, gens, best["R"])}
    del grid_cpu
    log(json.dumps(acc))
    if not args.quick:
        log("accuracy: heat eigenmodes, elastic forward and gradient")
        e1 = heat_eigenmode_error(A, (768, 768, 768), 1, 0.1, 100, (1, 1, 1), 1, 0, False)
        e1.update(workload="heat delta1 768^3, 100 steps, zero Dirichlet, sine eigenmode")
        acc["heat_delta1_768"] = e1
        e2 = heat_eigenmode_error(A, (768, 768, 768), 1, 0.1, 100, (1, 1, 1), 2, 1, False)
        e2.update(workload="the same through the fused R=2 kernel")
        acc["heat_delta1_768_fused_R2"] = e2
        e4 = heat_eigenmode_error(A, (768, 768, 768), 4, 0.05, 500, (2, 2, 2), 1, 0, True)
        e4.update(workload="heat delta4 768^3, 2x2x2 blocks, 500 steps, periodic Fourier mode")
        acc["heat_delta4_768"] = e4
        if heat_ref_grid is not None:
            acc["heat_delta1_128_100_steps"] = heat_reference_check(A, heat_ref_grid)
        acc.update(elastic_checks(A, W))
        log(json.dumps(acc))
    out = {
        "metric": METRIC,
        "value": round(best["glups"], 3),
        "unit": "GLUP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(best["ms"] / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic (SplitMix64 Bernoulli(0.3), seed 1)",
        "config": {"workload": "gol3d_512^3_u8_periodic_2x2x2", "R": best["R"],
                   "generations": args.steps, "fused_R_generations_per_launch": True,
                   "gol_zc": best["zc"], "gol_tile": best["tile"],
                   "fields_bit_packed_between_launches": best["packed"],
                   "glups_by_R": {str(r["R"]): round(r["glups"], 3) for r in runs},
                   "plain_step_by_step_R1_glups": round(plain["glups"], 3),
                   # the same run with the reference's scheme: boundary_updating
                   # (pack/unpack, byte fields) before every R-generation launch
                   "with_pack_unpack_exchange": {
                       "glups_by_R": {str(r["R"]): round(r["glups"], 3) for r in xruns},
                       "kernel_ms_by_R": {str(r["R"]): round(r["kernel_ms"], 5) for r in xruns},
                       "window_ms_by_R": {str(r["R"]): round(r["ms"] / (args.steps / r["R"]), 5)
                                          for r in xruns},
                       "tile_by_R": {str(r["R"]): r["tile"] for r in xruns}}},
        # The dominant kernel issues VALU instructions at close to the CU rate
        # and moves far fewer bytes than HBM could: its bound is the VALU
        # issue rate.  achieved = SQ_INSTS_VALU of one launch (rocprofv3
        # counters of this template and grid, profiles/traffic.txt) / the
        # launch time measured here; traffic = FETCH_SIZE + WRITE_SIZE of
        # one launch, whose rate against HBM is hbm_frac.
        "roofline": {"bound": "valu",
                     "achieved": round(valu_ach, 1) if valu_ach else None,
                     "peak": VALU_PEAK_GINST, "unit": "Ginst/s",
                     "frac": round(valu_ach / VALU_PEAK_GINST, 4) if valu_ach else None,
                     "traffic": traffic,
                     "kernel": gol_kernel_name(best["R"], best["tile"], best["packed"]),
                     "kernel_ms": round(best["kernel_ms"], 5),
                     "valu_insts_per_launch": valu,
                     "hbm_achieved_GBs": round(hbm_ach, 1) if hbm_ach else None,
                     "hbm_frac": round(hbm_ach / HBM_PEAK_GBS, 4) if hbm_ach else None,
                     "bytes_per_launch_algorithmic_u8": 2 * cells * best["gens_per_launch"],
                     "note": "VALU issue peak = 1 wave64 instruction per cycle per CU x "
                             "256 CUs x 2.4 GHz; the packed fields hold 1 bit per cell, so "
                             "the u8 algorithmic bytes (2 B per cell-generation) exceed what "
                             "the launch moves (traffic)",
                     # one generation per launch: the R=1 fused launch, and the
                     # plain step kernel of the reference's structure, as
                     # algorithmic u8 bytes against HBM
                     "single_pass_kernel": gol_kernel_name(1, r1["tile"], r1["packed"]),
                     "single_pass_u8_GBs": round(r1["kernel_gbs_alg"], 1),
                     "plain_step_kernel": "gol_step_kernel",
                     "plain_step_u8_GBs": round(plain["kernel_gbs_alg"], 1),
                     "plain_step_hbm_frac": round(plain["kernel_gbs_alg"] / HBM_PEAK_GBS, 4)},
        "cpu_baseline": cpu,
        "accuracy": acc,
        "also": also,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
