"""Synthetic StagYY binary files for the reader tests.

Writes the layout StagReadBinaryFile parses (src/readers/stagYY/binary_loader.c:48-206):
little-endian int32 magic, ntot[4], float aspect[2], npar[4], zg[2*nz+1],
float rcmb, int32 step, float time, erupta, botT, float x[nx], y[ny], z[nz],
(float scale for velocity-pressure files), then per parallel subdomain
(block, z, y, x order) its values in (b, k, j, i, component) order with one
ghost column in x and y for velocity-pressure files of format >= 9.  Ghost
entries hold a sentinel the reader must skip.
"""
import numpy as np

GHOST = np.float32(-7.0e30)


def yin_axes(n0, n1, n2, rcmb=0.55):
    """Cell-centred Yin axes: x in (0, pi/2) (latitude pi/4 - x), y in (0, 3pi/2)
    (longitude y - 3pi/4), z in (0, 1 - rcmb)."""
    x = ((np.arange(n0) + 0.5) * (0.5 * np.pi / n0)).astype(np.float32)
    y = ((np.arange(n1) + 0.5) * (1.5 * np.pi / n1)).astype(np.float32)
    z = ((np.arange(n2) + 0.5) * ((1.0 - rcmb) / n2)).astype(np.float32)
    return x, y, z


def write_stag(path, values, x, y, z, rcmb=0.55, npar=(1, 1, 1, 1), magic=None, scale=1.0):
    """values: float32 [nblocks, n2, n1, n0, nval] (nval 1 or 4)."""
    values = np.asarray(values, dtype=np.float32)
    nb, n2, n1, n0, nval = values.shape
    if magic is None:
        magic = 9 if nval == 1 else 309
    xyp = 1 if (magic % 100 >= 9 and nval == 4) else 0
    ntot = (n0, n1, n2, nb)
    ncpu = [ntot[k] // npar[k] for k in range(4)]
    assert all(ncpu[k] * npar[k] == ntot[k] for k in range(4))
    out = [np.array([magic], "<i4").tobytes(), np.array(ntot, "<i4").tobytes(),
           np.array([1.0, 1.0], "<f4").tobytes(), np.array(npar, "<i4").tobytes(),
           np.linspace(0.0, 1.0, 2 * n2 + 1).astype("<f4").tobytes(),
           np.array([rcmb], "<f4").tobytes(), np.array([42], "<i4").tobytes(),
           np.array([3.5, 0.0, 1.0], "<f4").tobytes(),
           np.asarray(x, "<f4").tobytes(), np.asarray(y, "<f4").tobytes(), np.asarray(z, "<f4").tobytes()]
    if nval != 1:
        out.append(np.array([scale], "<f4").tobytes())
    for ib in range(npar[3]):
        for iz in range(npar[2]):
            for iy in range(npar[1]):
                for ix in range(npar[0]):
                    loc = np.full((ncpu[3], ncpu[2], ncpu[1] + xyp, ncpu[0] + xyp, nval), GHOST, dtype="<f4")
                    loc[:, :, :ncpu[1], :ncpu[0], :] = values[ib * ncpu[3]:(ib + 1) * ncpu[3],
                                                              iz * ncpu[2]:(iz + 1) * ncpu[2],
                                                              iy * ncpu[1]:(iy + 1) * ncpu[1],
                                                              ix * ncpu[0]:(ix + 1) * ncpu[0], :]
                    out.append(loc.tobytes())
    with open(path, "wb") as f:
        f.write(b"".join(out))
