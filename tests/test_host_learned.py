"""The C++ host's learned-polynomial helpers against the Python ones (CPU):
Ritz extremes by Sturm bisection vs numpy's tridiagonal eigenvalues, the
Chebyshev residual coefficients, and the binary model file written by
cgxx_amd.learned.export_bin (round trip, and rejection of a bad file)."""
import os
import subprocess

import numpy as np
import pytest

from cgxx_amd import learned, spectrum
from cgxx_amd.cghip import LIB_PATH
from oracle import oracle
from tests.conftest import REPO
from tests.test_poly import laplace1d

HOST = os.path.join(REPO, "cgxx_amd", "host")


@pytest.fixture(scope="module")
def tool(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("learnedhost") / "learned_host")
    subprocess.run(["g++", "-std=c++17", "-O2", "-I" + HOST, "-I" + os.path.join(REPO, "include"),
                    "-o", out, os.path.join(REPO, "tests", "learned_host.cpp"),
                    os.path.join(HOST, "LearnedPoly.cpp"), "-L" + os.path.dirname(LIB_PATH),
                    "-lcghip", "-Wl,-rpath," + os.path.dirname(LIB_PATH)], check=True)
    return out


def run(tool, *args):
    return subprocess.run([tool, *map(str, args)], check=True, capture_output=True,
                          text=True).stdout.split()


@pytest.mark.parametrize("m", [1, 2, 5, 10])
def test_ritz_extremes_match_numpy(tool, m):
    n = 60
    A = laplace1d(n)
    k = np.random.default_rng(m).uniform(-1, 1, n)
    _, _, _, h = oracle.pcg(n=n, fmt="CRS", precond="jacobi", order=0, k=k, tol=1e-300,
                            max_iter=m, crs=(A.ptr, A.index, A.value), C_jac=np.full(n, 0.5))
    a, b = h["alpha"][:m], h["beta"][:max(m - 1, 0)]
    lo, hi = (float(v) for v in run(tool, "ritz", m, *[repr(float(v)) for v in a],
                                    *[repr(float(v)) for v in b]))
    th = spectrum.ritz_values(a, b)
    assert lo == pytest.approx(th[0], rel=1e-12) and hi == pytest.approx(th[-1], rel=1e-12)


@pytest.mark.parametrize("a,b,deg", [(0.01, 2.0, 8), (1e-4, 2.0, 8), (0.3, 1.5, 3), (0.05, 2.0, 16)])
def test_cheb_coefficients_match_python(tool, a, b, deg):
    c = np.array([float(v) for v in run(tool, "cheb", a, b, deg)])
    ref = spectrum.cheb_residual_coeffs(a, b, deg)
    np.testing.assert_allclose(c, ref, rtol=1e-9, atol=1e-12 * np.abs(ref).max())


def test_model_file_round_trip(tool, tmp_path):
    out = run(tool, "model", os.path.join(REPO, "cgxx_amd", "data", "poly_mlp.bin"))
    model = learned.load_model()
    dims = [int(v) for v in model["dims"]]
    assert [int(v) for v in out[:4]] == dims
    assert int(out[4]) == len(model["degrees"]) and int(out[5]) == int(model["fixed_degree"])
    assert int(out[6]) == model["params"].size
    assert float(out[7]) == float(model["rho_star"])
    assert [int(v) for v in out[8::2]] == [int(v) for v in model["degrees"]]
    assert [float(v) for v in out[9::2]] == [float(v) for v in model["lower_end"]]
    bad = tmp_path / "bad.bin"
    bad.write_bytes(open(os.path.join(REPO, "cgxx_amd", "data", "poly_mlp.bin"), "rb").read()[:100])
    r = subprocess.run([tool, "model", str(bad)], capture_output=True, text=True)
    assert r.returncode == 1 and "truncated" in r.stdout


def _synthetic_model(rng):
    D = len(learned.DEGREES)
    out = learned.N_OUT
    F = learned.N_FEATURES
    return {"dims": np.array(learned.DIMS), "degrees": np.array(learned.DEGREES), "fixed_degree": 8,
            "rho_star": 16.0, "lower_end": rng.uniform(2.0, 16.0, D),
            "feat_mean": np.zeros(F), "feat_std": np.ones(F), "feat_lo": -np.ones(F), "feat_hi": np.ones(F),
            "out_mean": np.concatenate([[rng.uniform(-8, -4)], rng.uniform(3, 6, D), rng.uniform(-3, 0, D),
                                        [rng.uniform(5, 8), rng.uniform(-3, -1)]]),
            "out_std": rng.uniform(0.1, 1.0, out),
            "params": np.zeros(sum(a * b + b for a, b in zip(learned.DIMS[:-1], learned.DIMS[1:])),
                               np.float32)}


@pytest.mark.parametrize("seed", range(6))
def test_decision_matches_python(tool, seed, tmp_path):
    """The C++ driver's choice (LearnedPoly.cpp decide) equals learned.decide
    on the same standardised outputs: the cheapest of Jacobi and each degree
    by predicted iterations x predicted ms per iteration, and the lower end
    c_k theta_min."""
    rng = np.random.default_rng(seed)
    model = _synthetic_model(rng)
    path = str(tmp_path / "m.bin")
    learned.export_bin(model, path)
    for _ in range(20):
        yn = rng.standard_normal(learned.N_OUT).astype(np.float32)
        got = run(tool, "decide", path, *[repr(float(v)) for v in yn])
        want = learned.decide(yn.astype(np.float64) * model["out_std"] + model["out_mean"], model)
        deg = int(got[0])
        assert deg == (0 if want["choice"] == "jacobi" else want["degree"])
        assert float(got[3]) == pytest.approx(want["cost_ms"]["jacobi"], rel=1e-6)
        assert float(got[4]) == pytest.approx(want["theta_min"], rel=1e-6)
        costs = [float(v) for v in got[5:]]
        assert costs == pytest.approx([want["cost_ms"][k] for k in learned.DEGREES], rel=1e-6)
        if deg:
            assert float(got[1]) == pytest.approx(want["lower_end"], rel=1e-6)
            assert float(got[2]) == pytest.approx(want["ms_per_iter"], rel=1e-6)
