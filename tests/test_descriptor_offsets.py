"""Offset algebra of the 16-bit descriptor path (lb_pull.hpp: fast_path_args,
pull_node, store_node), restated on the CPU.

For every tile / lane / link that may take the descriptor path, the address
base + scalar offset + lane offset must equal the per-link address
A + (i*ps + e - off_i + correction) * esz, both offsets must fit 32 bits,
their sum must stay inside the descriptor's size, and every node must stay
inside its population buffer.  Lanes with a link leaving through a wall
carry a bounce bit and are excluded (their wave takes the per-link path), as
are tiles on a halo or z-wrap plane.  This is the check that the host guard
(`fast`) admits only lattices whose offsets are exact.
"""
import numpy as np
import pytest

PERIODIC, OUTFLOW, INLET, WALL = "per", "out", "inl", "wall"


def velocities(Q):
    i = np.arange(Q)
    if Q == 9:
        cx = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1])
        cy = np.array([0, 0, -1, 0, 1, -1, -1, 1, 1])    # storage rows
        return cx, cy, np.zeros(9, int)
    cx = np.where((i == 1) | ((i >= 7) & (i <= 10)), 1, np.where((i == 2) | ((i >= 11) & (i <= 14)), -1, 0))
    cy = np.where(np.isin(i, (3, 7, 11, 15, 16)), 1, np.where(np.isin(i, (4, 8, 12, 17, 18)), -1, 0))
    cz = np.where(np.isin(i, (5, 9, 13, 15, 17)), 1, np.where(np.isin(i, (6, 10, 14, 16, 18)), -1, 0))
    return cx, cy, cz


def check(Q, shape, bc, layout, bx, es=2, halo=False):
    nx, ny, nz = shape
    by = 256 // bx
    N = nx * ny * nz
    ps, sy, sz = (nx, Q * nx, Q * nx * ny) if layout == 1 else (N, nx, nx * ny)
    n, stride = (nx, ny, nz), (1, sy, sz)

    def corr(a, face, sign):
        k = bc[face]
        if k == PERIODIC:
            return sign * n[a] * stride[a]
        return sign * stride[a] if k in (OUTFLOW, INLET) else 0

    lo = [corr(a, 2 * a, +1) for a in range(3)]
    hi = [corr(a, 2 * a + 1, -1) for a in range(3)]
    cx, cy, cz = velocities(Q)
    off = cx + cy * sy + cz * sz
    # host side (fast_path_args)
    bias = -(hi[0] + hi[1])
    offmax = int(off.max())
    lso = [(i * ps + offmax - int(off[i])) * es for i in range(Q)]
    sso = [i * ps * es for i in range(Q)]
    lane = bx + by * sy
    fast = max(lso) + (lane + lo[0] + lo[1] + bias) * es < 2 ** 32 and max(sso) + lane * es < 2 ** 32
    if not fast:
        return 0
    lx, ly = np.arange(256) % bx, np.arange(256) // bx
    checked = 0
    for z in range(nz):
        zedge = Q == 19 and ((z == 0 and (halo or lo[2])) or (z == nz - 1 and (halo or hi[2])))
        for ty in range(-(-ny // by)):
            for tx in range(-(-nx // bx)):
                x, y = tx * bx + lx, ty * by + ly
                v = (x < nx) & (y < ny)
                x, y = x[v], y[v]
                e0 = tx * bx + sy * ty * by + sz * z
                e = e0 + lx[v] + sy * ly[v]
                assert np.array_equal(e, x + y * sy + z * sz)
                vs = (e - e0) * es          # stores: base B + e0*es
                for s in range(Q):
                    assert sso[s] + vs.max() < 2 ** 32
                    assert np.array_equal(e0 * es + sso[s] + vs, (s * ps + e) * es)
                if zedge:
                    continue
                b0 = e0 - offmax - bias     # loads: base A + b0*es
                records = min((Q * N - b0) * es, 2 ** 32 - 1)
                vb = (e - e0 + bias) * es
                lox, hix = np.where(x == 0, lo[0], 0), np.where(x == nx - 1, hi[0], 0)
                loy, hiy = np.where(y == 0, lo[1], 0), np.where(y == ny - 1, hi[1], 0)
                for i in range(Q):
                    c = np.zeros_like(e)
                    c = c + (lox if cx[i] > 0 else hix if cx[i] < 0 else 0)
                    c = c + (loy if cy[i] > 0 else hiy if cy[i] < 0 else 0)
                    sx, syy, szz = x - cx[i], y - cy[i], z - cz[i]
                    wall = (((sx < 0) & (bc[0] == WALL)) | ((sx >= nx) & (bc[1] == WALL)) |
                            ((syy < 0) & (bc[2] == WALL)) | ((syy >= ny) & (bc[3] == WALL)) |
                            ((szz < 0) & (bc[4] == WALL)) | ((szz >= nz) & (bc[5] == WALL)))
                    ok = ~wall
                    vo = (vb + c * es)[ok]
                    src = (i * ps + e - off[i] + c)[ok]
                    assert (vo >= 0).all() and (vo + lso[i] < 2 ** 32).all() and (vo + lso[i] < records).all()
                    assert np.array_equal(b0 * es + lso[i] + vo, src * es)
                    assert ((src >= 0) & (src < Q * N)).all()
                    checked += int(ok.sum())
    return checked


FACES = [(INLET, OUTFLOW), (PERIODIC, PERIODIC), (WALL, WALL), (OUTFLOW, INLET)]


@pytest.mark.parametrize("layout", [0, 1])
@pytest.mark.parametrize("shape", [(32, 16, 6), (33, 17, 5), (2, 3, 4)])
def test_d3q19_descriptor_offsets(layout, shape):
    n = 0
    for bx_faces in FACES:
        for by_faces in FACES[1:]:
            for bz_faces in FACES[1:3]:
                for bx in (32, 256):
                    halo = bz_faces == (PERIODIC, PERIODIC) and bx == 32
                    n += check(19, shape, bx_faces + by_faces + bz_faces, layout, bx, halo=halo)
    assert n > 0


@pytest.mark.parametrize("layout", [0, 1])
@pytest.mark.parametrize("shape", [(256, 64, 1), (33, 17, 1), (2, 3, 1)])
def test_d2q9_descriptor_offsets(layout, shape):
    n = 0
    for bx_faces in FACES:
        for by_faces in FACES[1:]:
            for bx in (64, 128):
                n += check(9, shape, bx_faces + by_faces + (PERIODIC, PERIODIC), layout, bx)
    assert n > 0


def test_guard_rejects_offsets_past_32_bits():
    # A 16-bit SOA D3Q19 buffer past 4 GiB (slot 18 starts beyond 2^32
    # bytes) must be refused.  lb_create's limit of 2^31 population elements
    # keeps real 16-bit lattices below this, so the guard is a second line.
    assert check(19, (2048, 2048, 32), (PERIODIC,) * 6, 0, 256) == 0
    assert check(19, (2048, 2048, 32), (PERIODIC,) * 6, 0, 256, es=4) == 0
