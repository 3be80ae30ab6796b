"""Small helpers shared by the tests (no GPU access here)."""
from __future__ import annotations

import numpy as np


def expected_layout(env, blocks, R, radius, esize):
    """The field layout rule of the library (b2r_create), restated for CPU tests."""
    B = [env[a] // blocks[a] for a in range(3)]
    h = radius * R
    S = [b + 2 * h for b in B]
    if esize == 1:
        xoff = (16 - h % 16) % 16
        pitch = -(-(xoff + S[0] + 3) // 16) * 16
    else:
        xoff, pitch = 0, S[0]
    slice_ = pitch * S[1]
    q = 256 // esize
    bstride = -(-(slice_ * S[2]) // q) * q
    nblk = blocks[0] * blocks[1] * blocks[2]
    return dict(B=B, S=S, halo=h, nblocks=nblk, x_offset=xoff, pitch=pitch, slice=slice_,
                block_stride=bstride, field_elems=bstride * nblk)


class Lay:
    """Attribute view with the b2r_layout field names."""

    def __init__(self, d):
        self.__dict__.update(d)


def block_view(raw, lay, b):
    """(S_z, S_y, S_x) view of block b inside a raw field array."""
    base = b * lay.block_stride + lay.x_offset
    Sx, Sy, Sz = lay.S
    out = np.empty((Sz, Sy, Sx), raw.dtype)
    for z in range(Sz):
        for y in range(Sy):
            o = base + z * lay.slice + y * lay.pitch
            out[z, y] = raw[o:o + Sx]
    return out


def set_block(raw, lay, b, blk):
    base = b * lay.block_stride + lay.x_offset
    Sx, Sy, Sz = lay.S
    for z in range(Sz):
        for y in range(Sy):
            o = base + z * lay.slice + y * lay.pitch
            raw[o:o + Sx] = blk[z, y]


def id_field(lay, blocks, dtype, poison):
    """Raw field whose owned cells hold their global id (x + X*(y + Y*z)) and
    whose halos hold `poison`."""
    raw = np.full(lay.field_elems, poison, dtype)
    B, h = lay.B, lay.halo
    X, Y = B[0] * blocks[0], B[1] * blocks[1]
    for bz in range(blocks[2]):
        for by in range(blocks[1]):
            for bx in range(blocks[0]):
                b = (bz * blocks[1] + by) * blocks[0] + bx
                blk = np.full((lay.S[2], lay.S[1], lay.S[0]), poison, dtype)
                z, y, x = np.meshgrid(np.arange(B[2]) + bz * B[2], np.arange(B[1]) + by * B[1],
                                      np.arange(B[0]) + bx * B[0], indexing="ij")
                blk[h:h + B[2], h:h + B[1], h:h + B[0]] = (x + X * (y + Y * z)).astype(dtype)
                set_block(raw, lay, b, blk)
    return raw


def expected_halo_ids(lay, blocks, periodic, poison):
    """For each block, the (S_z,S_y,S_x) array a correct exchange produces from
    id_field: every cell whose global position (after periodic wrap) lies in the
    global grid holds that position's id, except halo cells on frozen global
    faces, which keep `poison`; cells that no exchange reaches (none, for a
    dimension-ordered exchange) would also keep poison."""
    B, h = lay.B, lay.halo
    G = [B[a] * blocks[a] for a in range(3)]
    # 0/1 (or bool): no axis or every axis wraps; otherwise a per-axis mask
    pm = 7 if int(periodic) == 1 else int(periodic)
    out = []
    for bz in range(blocks[2]):
        for by in range(blocks[1]):
            for bx in range(blocks[0]):
                c = (bx, by, bz)
                coords = []
                valid = []
                for a in range(3):
                    g = np.arange(lay.S[a]) - h + c[a] * B[a]
                    if (pm >> a) & 1:
                        valid.append(np.ones_like(g, bool))
                        g = g % G[a]
                    else:
                        valid.append((g >= 0) & (g < G[a]))
                    coords.append(g)
                z, y, x = np.meshgrid(coords[2], coords[1], coords[0], indexing="ij")
                vz, vy, vx = np.meshgrid(valid[2], valid[1], valid[0], indexing="ij")
                ids = (x + G[0] * (y + G[1] * z)).astype(np.float64)
                ok = vz & vy & vx
                ids[~ok] = poison
                out.append(ids)
    return out


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))
