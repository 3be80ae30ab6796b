"""Readers of the on-disk outputs the host driver writes in the reference's
formats (b2r_amd/host/b2r_output.cpp):

- ``displacement-<ts>.bin`` (eqwp.cu:616-627): three int32 (Ngx, Ngy, Ngz), then
  Ngx*Ngy*Ngz float32, x fastest;
- ``displacement-<ts>.txt`` (eqwp.cu:636-650): one line per (z, y) row, every
  value followed by ',';
- ``R-<R>-rank-<rank>-i-<step>.txt`` (heat-3D.hpp:105-127): per z a
  ``slice: [z, :, :]`` line, then one line per row of ``%0.2f,`` values.
"""
import numpy as np


def read_displacement_bin(path):
    """Returns the (z, y, x) float32 array of a displacement-<ts>.bin file."""
    with open(path, "rb") as f:
        hdr = np.frombuffer(f.read(12), np.int32)
        if hdr.size != 3 or np.any(hdr <= 0):
            raise ValueError("%s: bad header %s" % (path, hdr))
        nx, ny, nz = (int(v) for v in hdr)
        data = np.frombuffer(f.read(), np.float32)
    if data.size != nx * ny * nz:
        raise ValueError("%s: %d values for a %dx%dx%d grid" % (path, data.size, nx, ny, nz))
    return data.reshape(nz, ny, nx)


def _row(line):
    if not line.endswith(","):
        raise ValueError("row does not end with ','")
    return [float(v) for v in line[:-1].split(",")]


def read_displacement_txt(path, shape):
    """Returns the (z, y, x) float array of a displacement-<ts>.txt file of the
    given (nx, ny, nz) grid (the text form carries no header)."""
    nx, ny, nz = shape
    with open(path) as f:
        rows = [_row(line.rstrip("\n")) for line in f if line.strip()]
    a = np.array(rows, np.float64)
    if a.shape != (ny * nz, nx):
        raise ValueError("%s: %s values, expected %d rows of %d" % (path, a.shape, ny * nz, nx))
    return a.reshape(nz, ny, nx)


def read_block_step_txt(path):
    """Returns the (z, y, x) array of one R-<R>-rank-<rank>-i-<step>.txt block dump."""
    planes, cur, z = [], None, 0
    with open(path) as f:
        for line in f:
            line = line.rstrip("\n")
            if line.startswith("slice:"):
                if line != "slice: [%d, :, :]" % z:
                    raise ValueError("%s: unexpected %r" % (path, line))
                cur = []
                planes.append(cur)
                z += 1
            elif line:
                if cur is None:
                    raise ValueError("%s: row before the first slice line" % path)
                cur.append(_row(line))
    return np.array(planes, np.float64)
