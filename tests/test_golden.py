"""The committed golden fixture (tests/golden/c1_laplace2d_ellipse.npz,
written by tests/golden/make_golden.py from the oracle) is still what the
oracle computes: same inputs, bit-equal tree and skeletons, density and
field to 1e-12."""
import importlib.util
from pathlib import Path

import numpy as np

GOLD = Path(__file__).resolve().parent / "golden"


def load_fixture():
    return np.load(GOLD / "c1_laplace2d_ellipse.npz", allow_pickle=False)


def test_oracle_reproduces_golden_fixture():
    spec = importlib.util.spec_from_file_location(
        "make_golden", GOLD / "make_golden.py")
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    new = mg.compute()
    old = load_fixture()
    assert str(new["input_sha256"]) == str(old["input_sha256"])
    for k in ("box", "original", "original_off", "skel", "skel_off",
              "top_size"):
        assert np.array_equal(new[k], old[k]), k
    for k in ("mu", "field"):
        assert np.linalg.norm(new[k] - old[k]) <= \
            1e-12 * np.linalg.norm(old[k]), k
    # the fixture itself solves the PDE: field against the exact solution
    assert np.max(np.abs(old["field"] - old["exact"])) < 1e-10
