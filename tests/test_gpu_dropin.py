"""The drop-in boundary from a compiled C++ caller.

tests/dropin/ref_caller.cpp includes include/lattibol_amd.h, checks every
Data struct field offset against the reference layout at compile time, and
runs the reference's own step loop (simulation_d2q9.cu:462-554,
simulation_d3q19.cu:352-430) through the twelve launchers with the
reference's call sites (HIP dim3, the stream by reference).  Its populations,
macros, plot buffer and probe must equal the oracle's restatement of the same
loop bit for bit."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CALLER = os.path.join(ROOT, "tests", "dropin", "ref_caller")

DECKS = {
    "d2q9_zouhe_3dev": dict(shape=(16, 96), ndev=3, kw=dict(ux=0.05, rho=1.0, viscosity=0.05), steps=120),
    "d2q9_force": dict(shape=(16, 64), ndev=1, kw=dict(ux=0.0, rho=1.0, viscosity=0.1, Fx=2.0e-6), steps=100),
    "d2q9_shanchen_adhesion": dict(shape=(32, 64), ndev=1,
                                   kw=dict(ux=0.0, rho=200.0, viscosity=1 / 6, G=-120.0, Gads=-187.16), steps=150),
    "d3q19_zouhe_2dev": dict(shape=(6, 12, 64), ndev=2, kw=dict(ux=0.03, rho=1.0, viscosity=0.05), steps=60),
}


def _solids(shape):
    s = np.zeros(shape, np.uint8)
    if len(shape) == 2:
        H, W = shape
        s[0, :] = s[-1, :] = 1
        s[H // 3:H // 3 + 4, W // 4:W // 4 + 5] = 1
    else:
        D, H, W = shape
        s[:, 0, :] = s[:, -1, :] = 1
        s[2:5, 4:7, 6:10] = 1
    return s


def _write_input(path, ref, kw, steps, cursor):
    p = dict(ux=0.0, uy=0.0, uz=0.0, rho=1.0, Fx=0.0, Fy=0.0, G=0.0, Gads=0.0)
    p.update(kw)
    with open(path, "wb") as fh:
        np.array([ref.Q, ref.nd, ref.W, ref.H, ref.D, steps] + list(cursor), np.int32).tofile(fh)
        np.array([p["ux"], p["uy"], p["uz"], p["rho"], ref.tau_inv, p["Fx"], p["Fy"], p["G"], p["Gads"]],
                 np.float32).tofile(fh)
        for d in range(ref.nd):
            ref.f[d].tofile(fh)
            ref.solid[d].tofile(fh)
            ref.sb[d].tofile(fh)
            for a in (ref.rho, ref.ux, ref.uy):
                a[d].tofile(fh)
            if ref.Q == 9:
                for a in (ref.psi, ref.ueq_x, ref.ueq_y):
                    a[d].tofile(fh)


def test_caller_built():
    assert os.access(CALLER, os.X_OK), "build tests/dropin first (__graft_entry__.build())"


@pytest.mark.parametrize("deck", sorted(DECKS))
def test_compiled_reference_loop_matches_oracle(gpu, deck):
    d = DECKS[deck]
    s = _solids(d["shape"])
    cursor = (5, 6, 1 if len(d["shape"]) == 3 else 0)
    ref = O.RefOracle(s, deviceCount=d["ndev"], cursor=cursor, **d["kw"])
    with tempfile.TemporaryDirectory() as tmp:
        fin, fout = os.path.join(tmp, "in.bin"), os.path.join(tmp, "out.bin")
        _write_input(fin, ref, d["kw"], d["steps"], cursor)
        print("deck", deck, d["shape"], "subdomains", d["ndev"], flush=True)
        r = subprocess.run([CALLER, fin, fout], capture_output=True, text=True, timeout=100)
        assert r.returncode == 0, r.stdout + r.stderr
        out = np.fromfile(fout, np.uint8)
    ref.runLBM(d["steps"])
    N = ref.N
    nplot = (4 if ref.Q == 19 else 8) * N
    per = ref.Q * N + 3 * N + (N if ref.Q == 19 else 0) + nplot
    vals = out[:(per * ref.nd + 4) * 4].view(np.float32)
    err = int(out[(per * ref.nd + 4) * 4:].view(np.int32)[0])
    assert err == 0, err
    for dev in range(ref.nd):
        v = vals[dev * per:(dev + 1) * per]
        o = 0
        np.testing.assert_array_equal(v[o:o + ref.Q * N].reshape(ref.Q, N), ref.f[dev]); o += ref.Q * N
        np.testing.assert_array_equal(v[o:o + N], ref.rho[dev]); o += N
        np.testing.assert_array_equal(v[o:o + N], ref.ux[dev]); o += N
        np.testing.assert_array_equal(v[o:o + N], ref.uy[dev]); o += N
        if ref.Q == 19:
            np.testing.assert_array_equal(v[o:o + N], ref.uz[dev]); o += N
        np.testing.assert_array_equal(v[o:o + nplot], ref.plot[dev])
    np.testing.assert_array_equal(vals[per * ref.nd:], ref.value_cur)
