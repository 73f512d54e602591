"""CPU restatement of the row-slab one-pass Shan-Chen sweep's indexing
(k_step_scr in csrc/lb_scr.hip): workgroups of whole x-rows over z-chunks,
waves handling 64-node row segments, 19 buffer loads per node with a scalar
row / plane offset and a per-lane x offset, stores at the node, and four psi
planes in LDS with one padding element per row end.

For every workgroup, wave, segment and lane of a set of lattice shapes,
launch shapes and both population layouts it checks that
  * every load reads population i of the upstream node (x - c_i, periodic),
    inside the population buffer;
  * the sweeps store every node of the lattice exactly once per step;
  * every psi the sweep reads from LDS is the psi of the neighbour x + c_i,
    written by the march for the right plane and not yet overwritten, and
    every LDS index stays inside the four slots the host allocates.
The kernel's parity is checked on the GPU; this test pins the address
arithmetic without one."""
import numpy as np
import pytest

C19 = [(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1), (1, 1, 0), (1, -1, 0),
       (1, 0, 1), (1, 0, -1), (-1, 1, 0), (-1, -1, 0), (-1, 0, 1), (-1, 0, -1), (0, 1, 1), (0, 1, -1),
       (0, -1, 1), (0, -1, -1)]
WAVES = (6, 8, 12, 16)


def wrap1(v, n):
    return np.where(v < 0, v + n, np.where(v >= n, v - n, v))


def layout(kind, nx, ny, nz):
    N = nx * ny * nz
    if kind == "rows":
        return nx, 19 * nx, 19 * nx * ny          # ps, sy, sz
    return N, nx, nx * ny


def max_spi(nw, bf16):
    return {6: 2 if bf16 else 1, 8: 3 if bf16 else 2, 12: 2 if bf16 else 1, 16: 1}[nw]


def max_rows(nw, nx, bf16=True):
    nseg = (nx + 63) // 64
    if 2 * nseg > nw:
        return 0
    by = nw * max_spi(nw, bf16) // nseg
    while by >= 1 and 4 * (by + 2) * (nx + 2) * 4 > 65536:
        by -= 1
    return by


def emulate(nx, ny, nz, nw, BY, zc, kind, rev=None):
    ps, sy, sz = layout(kind, nx, ny, nz)
    N = nx * ny * nz
    nseg = (nx + 63) // 64
    W = nx + 2
    slot = (BY + 2) * W
    nty = -(-ny // BY)
    nblocks = nty * -(-nz // zc)
    stored = np.zeros(N, np.int64)
    lane = np.arange(64)
    for b in range(nblocks):
        ty, zch = b % nty, b // nty
        y0 = ty * BY
        bye = min(BY, ny - y0)
        z0 = zch * zc
        z1 = min(nz, z0 + zc)
        # LDS: (node id, plane k) written at each position of the four slots
        lds_node = np.full(4 * slot, -1, np.int64)
        lds_k = np.full(4 * slot, -99, np.int64)
        # wave w: interior segments s = w + j nw (row 1 + s / nseg), frame
        # segment f (row 0 or bye + 1): w; with the frame rows on the last
        # waves nw - 1 - w ("last", 16-bit storages) or w - foff in the wave's
        # own column w % nseg, foff the largest multiple of nseg with
        # foff + 2 nseg <= nw ("offset", FP32)
        foff = (nw - 2 * nseg) // nseg * nseg if rev == "offset" else 0
        spi = -(-BY * nseg // nw)
        segs = []

        def seg(r, xs, on, interior):
            x = xs * 64 + lane
            return dict(r=r, active=on, interior=interior and on, y=int(wrap1(np.array(y0 - 1 + r), ny)), x=x,
                        ok=x < nx)
        for w in range(nw):
            for j in range(spi):
                s = w + j * nw
                segs.append(seg(1 + s // nseg, s % nseg, s < bye * nseg, True))
            f, col = (nw - 1 - w, (nw - 1 - w) % nseg) if rev == "last" else (w - foff, w % nseg)
            segs.append(seg(0 if f < nseg else bye + 1, col, 0 <= f < 2 * nseg, False))
        for k in range(z0 - 1, z1 + 2):
            zk = int(wrap1(np.array(k), nz))

            def pull_check(g, z):
                x, ok = g["x"][g["ok"]], g["ok"]
                for i, (cx, cy, cz) in enumerate(C19):
                    so = i * ps + int(wrap1(np.array(g["y"] - cy), ny)) * sy + int(wrap1(np.array(z - cz), nz)) * sz
                    e = so + wrap1(x - cx, nx)
                    want = i * ps + wrap1(x - cx, nx) + ((g["y"] - cy) % ny) * sy + ((z - cz) % nz) * sz
                    assert (e == want).all() and (e >= 0).all() and (e < 19 * N).all()
                return ok

            q = k - 2
            if q >= z0:
                for g in segs:
                    if not g["interior"]:
                        continue
                    pull_check(g, q)
                    x = g["x"][g["ok"]]
                    ids = x + nx * (g["y"] + ny * q)
                    stored[ids] += 1
                    c = g["r"] * W + x + 1
                    for (cx, cy, cz) in C19[1:] + [(0, 0, 0)]:
                        pl = (q + cz + 4) & 3
                        pos = pl * slot + c + cx + cy * W
                        assert (pos >= 0).all() and (pos < 4 * slot).all()
                        want = ((x + cx) % nx) + nx * (((g["y"] + cy) % ny) + ny * ((q + cz) % nz))
                        assert (lds_node[pos] == want).all(), (b, k, cx, cy, cz)
                        assert (lds_k[pos] == q + cz).all()
            if k <= z1:
                for g in segs:
                    if not g["active"]:
                        continue
                    pull_check(g, zk)
                    x = g["x"][g["ok"]]
                    ids = x + nx * (g["y"] + ny * zk)
                    base = ((k + 4) & 3) * slot + g["r"] * W
                    for pos, sel in ((base + x + 1, slice(None)), (base + 0 * x, x == nx - 1),
                                     (base + nx + 1 + 0 * x, x == 0)):
                        assert (pos[sel] >= 0).all() and (pos[sel] < 4 * slot).all()
                        lds_node[pos[sel]] = ids[sel]
                        lds_k[pos[sel]] = k
    assert (stored == 1).all(), (stored.min(), stored.max())


# (shape, waves per workgroup, rows per tile (0: the largest), z-chunk): one
# to four 64-node segments per row (partial last segments), tiles and chunks
# that do not divide the lattice, a single plane, one-row tiles
CASES = [((24, 24, 24), 8, 0, 5), ((24, 24, 24), 16, 1, 1), ((70, 13, 11), 12, 3, 4),
         ((70, 13, 11), 16, 0, 16), ((130, 5, 3), 8, 0, 2), ((130, 5, 3), 12, 2, 1),
         ((64, 7, 1), 8, 4, 3), ((192, 14, 5), 6, 6, 4), ((130, 9, 3), 6, 0, 2), ((200, 3, 4), 16, 0, 3), ((256, 2, 3), 8, 0, 2), ((256, 9, 3), 12, 5, 2),
         ((192, 6, 49), 16, 3, 48)]


@pytest.mark.parametrize("kind", ["soa", "rows"])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_row_slab_indexing(case, kind):
    (nx, ny, nz), nw, by, zc = CASES[case]
    emulate(nx, ny, nz, nw, by or max_rows(nw, nx), zc, kind)


@pytest.mark.parametrize("kind", ["soa", "rows"])
@pytest.mark.parametrize("case", [0, 2, 4, 7, 11])
@pytest.mark.parametrize("rev", ["last", "offset"])
def test_row_slab_indexing_frames_on_last_waves(case, kind, rev):
    """The frame rows on the last waves (wave w computes frame segment
    w - foff in its own column) cover the same rows and write the same LDS
    positions."""
    (nx, ny, nz), nw, by, zc = CASES[case]
    emulate(nx, ny, nz, nw, by or max_rows(nw, nx), zc, kind, rev=rev)


@pytest.mark.parametrize("nw", WAVES)
def test_row_slab_c3_launch_shapes(nw):
    """Rows of 192 nodes (the 192^3 spinodal box) at each workgroup's largest
    tile and two z-chunks, on a lattice cut to three row tiles and two
    z-chunks plus one plane (every tile and chunk runs the same march)."""
    by = max_rows(nw, 192)
    for zc in (6, 12):
        emulate(192, 2 * by + 3, 2 * zc + 1, nw, by, zc, "rows")


@pytest.mark.parametrize("kind", ["soa", "rows"])
@pytest.mark.parametrize("es", [4, 2])
def test_row_slab_byte_offsets_stay_32_bit(kind, es):
    """The kernel forms every load and store offset in 32-bit byte units:
    slot i ps es, row y sy es, plane z sz es (each product in 32 bits), their
    sums, plus x es per lane.  The host admits the form only for
    19 N es < 2^32 (scr_lattice_ok, csrc/lb_fused.hip); on the largest such
    lattices (nx up to 512) every product and partial sum of the largest
    offset stays below 2^32, and one node more per plane row breaks the rule."""
    for nx in (512, 192, 130):
        n_max = ((1 << 32) - 1) // (19 * es)             # largest N the host admits
        ny = 64
        nz = n_max // (nx * ny)
        N = nx * ny * nz
        assert 19 * N * es < (1 << 32) and 19 * nx * ny * (nz + 1) * es >= (1 << 32)
        ps, sy, sz = layout(kind, nx, ny, nz)
        terms = [18 * ps * es, (ny - 1) * sy * es, (nz - 1) * sz * es, (nx - 1) * es, sz * es, sy * es]
        assert all(t < (1 << 32) for t in terms)
        rz = (ny - 1) * sy * es + (nz - 1) * sz * es       # the row + plane sum
        assert rz < (1 << 32) and 18 * ps * es + rz < (1 << 32)
        top = 18 * ps * es + rz + (nx - 1) * es            # the last population of the last node
        assert top == (19 * N - 1) * es if kind == "rows" else top == (18 * N + N - 1) * es
        assert top < (1 << 32)
