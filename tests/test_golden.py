"""Committed fixtures (tests/golden/*.npz, written by make_fixtures.py from
the oracle): the oracle must keep reproducing them bit for bit (CPU), and the
HIP path must produce the same bytes (GPU).  The reference ships no expected
outputs, so these pin regressions of both sides against one frozen state."""
import glob
import json
import os

import numpy as np
import pytest

import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FILES = sorted(glob.glob(os.path.join(GOLDEN, "*.npz")))
NAMES = [os.path.basename(f)[:-4] for f in FILES]


def load(name):
    d = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    return json.loads(str(d["meta"])), d


def test_fixtures_present():
    assert len(NAMES) >= 15, NAMES


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_fixture(name):
    meta, d = load(name)
    if meta["kind"] == "ref":
        ref = O.RefOracle(d["solid"], deviceCount=meta["deviceCount"], cursor=tuple(meta["cursor"]), **meta["kwargs"])
        ref.runLBM(meta["steps"])
        np.testing.assert_array_equal(ref.f, d["f"])
        np.testing.assert_array_equal(ref.plot, d["plot"])
        np.testing.assert_array_equal(ref.value_cur, d["value_cur"])
    else:
        orc = O.PullOracle(meta["cfg"], d["solid"])
        orc.init_equilibrium(*d["fields"])
        orc.run(meta["steps"])
        np.testing.assert_array_equal(orc.A, d["pops"])
        np.testing.assert_array_equal(orc.macros_arr, d["macros"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_gpu_reproduces_fixture(gpu, name):
    meta, d = load(name)
    print("fixture", name, flush=True)
    if meta["kind"] == "ref":
        sim = gpu.Simulation(d["solid"], deviceCount=meta["deviceCount"], cursor=tuple(meta["cursor"]),
                             **meta["kwargs"])
        sim.runLBM(meta["steps"])
        for dev in range(meta["deviceCount"]):
            got = sim.download(dev)
            np.testing.assert_array_equal(got["f"].reshape(d["f"].shape[1:]), d["f"][dev])
            np.testing.assert_array_equal(got["plot"], d["plot"][dev])
        np.testing.assert_array_equal(sim.valueCur(), d["value_cur"])
        sim.close()
    else:
        sol = gpu.Solver(meta["cfg"], d["solid"])
        sol.init_equilibrium(*d["fields"])
        sol.step(meta["steps"])
        fluid = d["solid"].reshape(d["pops"].shape[1:]) == 0
        np.testing.assert_array_equal(sol.populations()[:, fluid], d["pops"][:, fluid])
        m = sol.macros()
        for k, key in enumerate(("rho", "ux", "uy", "uz")):
            np.testing.assert_array_equal(m[key][fluid], d["macros"][k][fluid].astype(np.float32))
        sol.close()
