"""Input decks (parameter files, BMP and STL solids) and the VTK writer.

The BMP / STL decoders are checked against masks this test encodes itself
(every BMP bit depth and row order, binary and ASCII STL of shapes with an
exact inside test); the parameter reader against hand-written files and the
two sample decks kept in tests/golden/decks (fig31_force = the reference's
"external forces/1_fig31" body-force channel, fig52_angle90 = "Fig52 -
Contact Angles/angle90").  The reference voxelises STL with cvmlcpp, which
is not available here, so STL parity with the reference is unpinned; the
checks below pin the voxeliser to the exact geometry instead."""
import os
import struct
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from lattibol_amd import deck

DECKS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "decks")


# ---------------------------------------------------------------------------
# parameter files
# ---------------------------------------------------------------------------
def test_params_all_lines(tmp_path):
    p = tmp_path / "a.dat"
    p.write_text("2\n1\n0.05\n0.01\n0.00\n1.5\n0.1\n25\n1e-5\n-2e-6\n0\n-120\n-187.16\n")
    d = deck.read_params(str(p))
    assert d["device_count"] == 2 and d["max_dim"] == 1 and d["update_steps"] == 25
    assert d["ux"] == np.float32(0.05) and d["uy"] == np.float32(0.01) and d["rho"] == np.float32(1.5)
    assert d["fx"] == np.float32(1e-5) and d["fy"] == np.float32(-2e-6)
    assert d["G"] == np.float32(-120) and d["Gads"] == np.float32(-187.16)
    # tau_inv = 1/(3 nu + 0.5) with the float viscosity promoted to double
    assert d["tau_inv"] == np.float32(1.0 / (3.0 * np.float64(np.float32(0.1)) + 0.5))
    assert d["max_u"] == np.float32(np.float32(0.05) * np.float32(1.5))
    assert d["min_rho"] == np.float32(np.float32(1.5) * np.float32(0.98))


def test_params_short_file_and_max_dim(tmp_path):
    p = tmp_path / "b.dat"
    p.write_text("1\n256\n0.1\n0\n0\n1\n0.03\n1\n")     # 3D sample decks stop here
    d = deck.read_params(str(p))
    assert d["max_dim"] == 256 and d["fx"] == 0 and d["G"] == 0 and d["Gads"] == 0
    p.write_text("1\n0\n0.1\n")
    assert deck.read_params(str(p))["max_dim"] == 1         # <= 1 means a 2D deck
    p.write_text("1\n1\nabc\n")
    assert deck.read_params(str(p))["ux"] == 0.0            # an unreadable line reads as 0


def test_params_missing_file_raises(tmp_path):
    with pytest.raises(Exception):
        deck.read_params(str(tmp_path / "none.dat"))


# ---------------------------------------------------------------------------
# BMP
# ---------------------------------------------------------------------------
def encode_bmp(mask, bpp, top_down=False, black_index=1):
    """A BI_RGB BMP whose black pixels are the mask's ones."""
    h, w = mask.shape
    stride = ((w * bpp + 31) // 32) * 4
    rows = []
    ncol = (1 << bpp) if bpp <= 8 else 0
    pal = b""
    if bpp <= 8:
        cols = [(255, 255, 255, 0)] * ncol
        cols[black_index] = (0, 0, 0, 0)
        if ncol > 2:
            cols[(black_index + 1) % ncol] = (200, 10, 30, 0)     # a non-black, non-white colour
        pal = b"".join(bytes(c) for c in cols)
    for y in (range(h) if top_down else range(h - 1, -1, -1)):
        if bpp <= 8:
            bits = np.zeros(stride * 8, np.uint8)
            idx = np.where(mask[y] == 1, black_index, (black_index + 1) % ncol if ncol > 2 else 1 - black_index)
            for x in range(w):
                for b in range(bpp):
                    bits[x * bpp + b] = (idx[x] >> (bpp - 1 - b)) & 1
            rows.append(np.packbits(bits).tobytes())
        else:
            px = np.where(mask[y][:, None] == 1, 0, np.array([40, 90, 250]))
            row = b"".join(bytes([int(v) for v in p] + ([255] if bpp == 32 else [])) for p in px)
            rows.append(row + b"\0" * (stride - len(row)))
    data = b"".join(rows)
    off = 14 + 40 + len(pal)
    hdr = struct.pack("<2sIHHI", b"BM", off + len(data), 0, 0, off)
    info = struct.pack("<IiiHHIIiiII", 40, w, -h if top_down else h, 1, bpp, 0, len(data), 2835, 2835,
                       ncol if bpp <= 8 else 0, 0)
    return hdr + info + pal + data


@pytest.mark.parametrize("bpp", [1, 4, 8, 24, 32])
@pytest.mark.parametrize("top_down", [False, True])
@pytest.mark.parametrize("w", [13, 64, 194])
def test_bmp_decoder(tmp_path, bpp, top_down, w):
    rng = np.random.default_rng(bpp * 1000 + w)
    h = 7
    mask = (rng.random((h, w)) < 0.3).astype(np.uint8)
    p = tmp_path / "m.bmp"
    p.write_bytes(encode_bmp(mask, bpp, top_down, black_index=0 if bpp == 1 and w == 13 else 1))
    got = deck.read_bmp(str(p))
    np.testing.assert_array_equal(got, mask)


def test_bmp_rejects_compressed(tmp_path):
    b = bytearray(encode_bmp(np.zeros((4, 8), np.uint8), 8))
    b[30] = 1                                                # BI_RLE8
    p = tmp_path / "c.bmp"
    p.write_bytes(bytes(b))
    with pytest.raises(Exception):
        deck.read_bmp(str(p))


# ---------------------------------------------------------------------------
# STL
# ---------------------------------------------------------------------------
def box_tris(lo, hi):
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    v = np.array([[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0],
                  [x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]], np.float32)
    f = [(0, 2, 1), (0, 3, 2), (4, 5, 6), (4, 6, 7), (0, 1, 5), (0, 5, 4), (2, 3, 7), (2, 7, 6),
         (1, 2, 6), (1, 6, 5), (0, 4, 7), (0, 7, 3)]
    return np.array([[v[a], v[b], v[c]] for a, b, c in f], np.float32)


def corner_tris(hi):
    """Two tiny facets at (0,0,0) and hi that fix the bounding box; no voxel
    centre column passes through them."""
    e = 0.05
    return np.array([[[0, 0, 0], [e, 0, 0], [0, e, 0]],
                     [[hi[0], hi[1], hi[2]], [hi[0] - e, hi[1], hi[2]], [hi[0], hi[1] - e, hi[2]]]], np.float32)


def sphere_tris(c, r, n=48):
    th = np.linspace(0, np.pi, n + 1)
    ph = np.linspace(0, 2 * np.pi, 2 * n + 1)
    P = lambda t, p: np.array([c[0] + r * np.sin(t) * np.cos(p), c[1] + r * np.sin(t) * np.sin(p),
                               c[2] + r * np.cos(t)], np.float64)
    tris = []
    for i in range(n):
        for j in range(2 * n):
            a, b, d, e = P(th[i], ph[j]), P(th[i + 1], ph[j]), P(th[i + 1], ph[j + 1]), P(th[i], ph[j + 1])
            if i > 0:
                tris.append([a, b, e])
            if i < n - 1:
                tris.append([b, d, e])
    return np.array(tris, np.float32)


def write_stl(path, tris, ascii_=False):
    if ascii_:
        lines = ["solid t"]
        for t in tris:
            lines += ["facet normal 0 0 0", " outer loop"] + ["  vertex %.9g %.9g %.9g" % tuple(v) for v in t] + \
                     [" endloop", "endfacet"]
        path.write_text("\n".join(lines + ["endsolid t"]) + "\n")
    else:
        body = b"".join(struct.pack("<3f", 0, 0, 0) + t.astype("<f4").tobytes() + b"\0\0" for t in tris)
        path.write_bytes(b"MESH-" * 16 + struct.pack("<I", len(tris)) + body)


@pytest.mark.parametrize("ascii_", [False, True])
def test_stl_box_in_box(tmp_path, ascii_):
    lo, hi = np.array([3.0, 2.0, 1.0]), np.array([11.0, 7.0, 5.0])
    tris = np.concatenate([box_tris(lo, hi), corner_tris((20, 10, 8))])
    p = tmp_path / "b.stl"
    write_stl(p, tris, ascii_)
    v = deck.voxelize_stl(str(p), 40)          # voxel 0.5
    assert v.shape == (16, 20, 40)
    z, y, x = np.meshgrid(np.arange(16), np.arange(20), np.arange(40), indexing="ij")
    c = np.stack([(x + 0.5) * 0.5, (y + 0.5) * 0.5, (z + 0.5) * 0.5], -1)
    want = np.all((c > lo) & (c < hi), axis=-1)
    np.testing.assert_array_equal(v.astype(bool), want)


def test_stl_sphere(tmp_path):
    c, r = (16.0, 12.0, 10.0), 8.0
    tris = np.concatenate([sphere_tris(c, r), corner_tris((32, 24, 20))])
    p = tmp_path / "s.stl"
    write_stl(p, tris)
    v = deck.voxelize_stl(str(p), 32)          # voxel 1.0
    assert v.shape == (20, 24, 32)
    z, y, x = np.meshgrid(np.arange(20), np.arange(24), np.arange(32), indexing="ij")
    d = np.sqrt((x + 0.5 - c[0]) ** 2 + (y + 0.5 - c[1]) ** 2 + (z + 0.5 - c[2]) ** 2)
    inside, outside = d < r - 0.2, d > r + 0.02    # the faceted surface lies within 0.2 of the sphere
    assert v[inside].all() and not v[outside].any()


def test_reference_sphere_deck_voxelises_to_its_sphere():
    """The reference's 3D sample deck (samples/esfera.dat + esfera.stl): a
    sphere inside the 600 x 300 x 300 extent set by four corner triangles,
    voxelised to maxDim = 256.  The sphere's centre and radius come from the
    mesh's own vertices; the solid voxels must have the sphere's volume and
    centre in voxel units and lie inside / outside it away from the surface."""
    path = os.path.join(DECKS, "esfera.stl")
    raw = open(path, "rb").read()
    n = int(np.frombuffer(raw[80:84], "<u4")[0])
    rec = np.dtype([("n", "<f4", 3), ("v", "<f4", (3, 3)), ("a", "<u2")])
    verts = np.frombuffer(raw[84:84 + 50 * n], rec)["v"].reshape(-1, 3).astype(np.float64)
    lo, hi = verts.min(0), verts.max(0)
    on_box = np.any(np.isclose(verts, lo) | np.isclose(verts, hi), axis=1)
    sph = verts[~on_box]
    centre = (sph.min(0) + sph.max(0)) / 2
    radius = np.linalg.norm(sph - centre, axis=1).mean()
    p = deck.read_params(os.path.join(DECKS, "esfera.dat"))
    assert p["max_dim"] == 256
    v = deck.voxelize_stl(path, p["max_dim"])
    scale = (hi - lo).max() / 256                     # mesh units per voxel
    assert v.shape == tuple(int(round(e / scale)) for e in (hi - lo)[::-1])
    r_vox, c_vox = radius / scale, (centre - lo) / scale
    count = int(v.sum())
    assert abs(count / (4 / 3 * np.pi * r_vox ** 3) - 1) < 0.01, (count, r_vox)
    cz, cy, cx = np.argwhere(v).mean(0) + 0.5
    np.testing.assert_allclose([cx, cy, cz], c_vox, atol=0.05)
    z, y, x = np.meshgrid(*(np.arange(k) + 0.5 for k in v.shape), indexing="ij")
    d = np.sqrt((x - c_vox[0]) ** 2 + (y - c_vox[1]) ** 2 + (z - c_vox[2]) ** 2)
    assert v[d < r_vox - 0.5].all() and not v[d > r_vox + 0.5].any()


# ---------------------------------------------------------------------------
# VTK
# ---------------------------------------------------------------------------
def parse_vti(path, n_x):
    root = ET.parse(path).getroot()
    arrs = {a.get("Name"): np.array(a.text.split(), np.float32) for a in root.iter("DataArray")}
    return root, arrs


@pytest.mark.parametrize("dims", [(1, 5, 8), (3, 4, 6)])
def test_vtk_writer(tmp_path, dims):
    D, H, W = dims
    rng = np.random.default_rng(0)
    f = {k: rng.standard_normal((D, H, W)).astype(np.float32) for k in ("rho", "ux", "uy", "uz")}
    deck.write_vtk(str(tmp_path), 7, f, device_count=2)
    idx = ET.parse(tmp_path / "r0007.pvti").getroot()
    pieces = [p.get("Source") for p in idx.iter("Piece")]
    assert pieces == ["r0007p00.vti", "r0007p01.vti"]
    w = W // 2
    for dev, name in enumerate(pieces):
        root, arrs = parse_vti(tmp_path / name, w)
        piece = next(root.iter("Piece"))
        assert piece.get("Extent") == "%d %d 0 %d 0 %d" % (dev * w, (dev + 1) * w, H - 1, D - 1)
        # each row of w values is followed by one pad value
        rho = arrs["Pressure"].reshape(-1, w + 1)
        np.testing.assert_allclose(rho[:, :w].ravel(), f["rho"][:, :, dev * w:(dev + 1) * w].ravel(), rtol=1e-5)
        assert (rho[:, w] == 0).all()
        vel = arrs["Velocity"].reshape(-1, w + 1, 3)
        sub = {k: f[k][:, :, dev * w:(dev + 1) * w].ravel() for k in f}
        np.testing.assert_allclose(vel[:, :w, 0].ravel(), sub["ux"], rtol=1e-5)
        if D > 1:
            np.testing.assert_allclose(vel[:, :w, 1].ravel(), sub["uy"], rtol=1e-5)
            np.testing.assert_allclose(vel[:, :w, 2].ravel(), sub["uz"], rtol=1e-5)
        else:      # 2D: physical y points up while rows grow downward
            np.testing.assert_allclose(vel[:, :w, 1].ravel(), -sub["uy"], rtol=1e-5)
            assert (vel[:, :w, 2] == 0).all()


# ---------------------------------------------------------------------------
# sample decks
# ---------------------------------------------------------------------------
def test_sample_decks_load():
    d = deck.load(os.path.join(DECKS, "fig31_force.dat"))
    p, s = d["params"], d["solids"]
    assert s.shape == (256, 1024) and p["max_dim"] == 1 and p["fx"] == np.float32(2.035e-6)
    assert s[0].all() and s[-1].all() and not s[1:-1].any()        # a plane channel
    d = deck.load(os.path.join(DECKS, "fig52_angle90.dat"))
    p, s = d["params"], d["solids"]
    assert s.shape == (200, 192) and p["G"] == -120 and p["Gads"] == np.float32(-187.16) and p["rho"] == 200
    assert s[180:].all() and not s[:180].any()                     # fluid above a 20-row floor
