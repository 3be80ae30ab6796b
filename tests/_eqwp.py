"""The reference eqwp model as the host shims set it up (eqwp.cu:9-25,
515-582), restated for the checks: constant medium, and the y-force patch that
every block builds in its own padded coordinates, forced while the window
start is below FRC_TIME.  Checker only."""
import numpy as np

import _elastic as E

LAME1, LAME2, INV_RHO = 32.0, 174.87, 1.0 / 1800.0
DT, IH, FRC_TIME, INIT_Y, FORCE = 0.5, 0.1, 50, 7, -10000.0


def patch_sources(env, blocks, R, D_cfg):
    """Global cells of the per-block patches: padded block Ng = B + 2*D_cfg*R
    (what the driver passes to eqwp_cuda_init), x in [Ngx/2, Ngx/2 + Ngx/10),
    z likewise, row iy = Ngy - INIT_Y - D_cfg*R (eqwp.cu:135, 229, 561-566),
    as owned coordinates of each block."""
    B = [env[a] // blocks[a] for a in range(3)]
    href = D_cfg * R
    Ng = [B[a] + 2 * href for a in range(3)]
    oy = Ng[1] - INIT_Y - 2 * href
    out = []
    for bz in range(blocks[2]):
        for by in range(blocks[1]):
            for bx in range(blocks[0]):
                if not 0 <= oy < B[1]:
                    continue
                for x in range(Ng[0] // 2, Ng[0] // 2 + Ng[0] // 10):
                    for z in range(Ng[2] // 2, Ng[2] // 2 + Ng[2] // 10):
                        ox, oz = x - href, z - href
                        if 0 <= ox < B[0] and 0 <= oz < B[2]:
                            out.append([bx * B[0] + ox, by * B[1] + oy, bz * B[2] + oz])
    return np.array(out, np.int32).reshape(-1, 3)


def forced_steps(R):
    """Every inner step of a window whose start ts < FRC_TIME is forced."""
    return -(-FRC_TIME // R) * R


def forward(env, blocks, R, D_cfg, iters, periodic=False, disp=False):
    """The oracle's FP32 forward of the reference model: (state, receivers[, disp])."""
    gx, gy, gz = env
    shape = (gz, gy, gx)
    lam = np.full(shape, LAME1, np.float32)
    mu = np.full(shape, LAME2, np.float32)
    b = np.full(shape, np.float32(INV_RHO), np.float32)
    src = patch_sources(env, blocks, R, D_cfg)
    nt = min(forced_steps(R), iters)
    wav = np.zeros((iters, max(len(src), 1)), np.float32)
    wav[:nt, :len(src)] = FORCE
    ih = (np.float32(IH),) * 3
    if periodic:
        return E.forward_periodic(env, lam, mu, b, np.float32(DT), ih, iters, src,
                                  [1] * len(src), wav[:, :len(src)])
    return E.forward(env, lam, mu, b, np.float32(DT), ih, iters, src, [1] * len(src),
                     wav[:, :len(src)], np.array([[1, 1, 1]], np.int32), 1, disp=disp)
