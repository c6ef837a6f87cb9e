"""The heavy-tailed measurement module at a reduced size: CRS and ELL
FP64 Jacobi-PCG solves agree (same per-row order, dots folded along
different trees: within one iteration, solutions within 1e-7 relative),
both reach the tolerance, and the mixed solver converges on its CRS layout."""
import json

import pytest

from cgxx_amd import heavy_tail

pytestmark = pytest.mark.gpu


def test_heavy_tail_crs_ell_identical(capsys):
    print("heavy_tail n=30000 cap=1000", flush=True)
    rc = heavy_tail.main(["--n", "30000", "--cap", "1000", "--steps", "1"])
    cap = capsys.readouterr()
    out = json.loads(cap.out.strip().splitlines()[-1])
    assert rc == 0 and out["crs_ell_agree"], cap.err + json.dumps(out)
    assert out["ell_width"] == out["system"]["max_row"] and out["ell_width"] > 500
    for fmt in ("CRS", "ELL"):
        f = out["formats"][fmt]
        assert f["residual"] <= 1e-9 and f["true_relative_residual"] <= 2e-9
    mj = out["mixed_jacobi"]
    assert mj["layout"] == "CRS" and mj["true_relative_residual"] <= 1e-9
