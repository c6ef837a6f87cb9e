"""What the trained network gains on the device, on held-out systems.

24 held-out members of the anisotropic diffusion family (seeds 100000-100023,
64^3-96^3, never seen in training), each solved from x = 0 to FP64 true
relative residual 1e-9 by
  learned  the network's choice from the matrix statistics alone (Jacobi,
           or the polynomial of the degree it predicts cheapest on
           [c_k theta_min_pred, b]), statistics and inference inside the
           timed solve
  fixed    the method without the network: 10 FP64 PCG steps for theta10
           (inside the timed solve; their 10 SpMVs counted), Chebyshev of one
           degree on [theta10 / rho*, b], degree and rho* tuned on the
           training systems
  ritz     as fixed, Chebyshev of degree 8 on the raw 10-step Ritz interval
Each method runs once untimed, then twice timed; the faster run counts.
The summed wall times and SpMVs are printed.  The network must win on both
sums against the fixed rule (whose 10 FP64 steps count as SpMVs), take at
most 0.92x the raw interval's time and fewer SpMVs.  It minimises
predicted time, not SpMVs: a higher degree needs more SpMVs than a lower
one (CG is optimal among Krylov polynomials) but its Clenshaw steps are
cheaper than CG iterations, so its SpMV saving over the raw interval
(about 8%) is smaller than its time saving (about 14%).
"""
import os

import pytest

from cgxx_amd import learned, problems

pytestmark = pytest.mark.gpu

SEEDS = range(100000, 100024)
SIDES = (64, 80, 96)


@pytest.mark.skipif(not os.path.exists(learned.MODEL_PATH), reason="no trained model file")
def test_learned_beats_fixed_rule_in_summed_time():
    from cgxx_amd.cghip import Context, Mlp

    model = learned.load_model()
    rho, fixed_degree = float(model["rho_star"]), int(model["fixed_degree"])
    ctx = Context(0)
    mlp = Mlp(tuple(int(v) for v in model["dims"]), 1)
    mlp.set_params(model["params"])

    def chebyshev(n, rho_, degree):
        return learned.solve_fixed_rule(ctx, n, rho_, degree)

    def best(run):
        run()
        a, b_ = run(), run()
        return a if a["wall_s"] <= b_["wall_s"] else b_

    tot = {"learned": 0.0, "fixed": 0.0, "ritz": 0.0}
    spmvs = {"learned": 0, "fixed": 0, "ritz": 0}
    choices = {}
    try:
        for seed in SEEDS:
            m = problems.anisotropic_family(seed, sides=SIDES)
            ctx.set_size(m.n)
            ctx.upload_crs(m.ptr, m.index, m.value)
            ctx.set_format("CRS")
            ctx.upload_jacobi(1.0 / m.diagonal())
            ctx.prepare_fp32()
            ctx.upload("K", m.ones_rhs())
            runs = {"learned": best(lambda: learned.solve_learned(ctx, mlp, model, m.n)),
                    "fixed": best(lambda: chebyshev(m.n, rho, fixed_degree)),
                    "ritz": best(lambda: chebyshev(m.n, 1.0, 8))}
            for name, r in runs.items():
                assert r["true_residual"] <= 1.05e-9, (seed, name, r)
                tot[name] += r["wall_s"]
                spmvs[name] += r["spmvs"]
            dec = runs["learned"]["decision"]
            key = "jacobi" if dec["choice"] == "jacobi" else f"k{dec['degree']}"
            choices[key] = choices.get(key, 0) + 1
    finally:
        mlp.close()
        ctx.close()
    print(f"\nsummed ms over {len(SEEDS)} held-out systems: "
          + ", ".join(f"{k} {1e3 * v:.1f}" for k, v in tot.items())
          + f"; SpMVs {spmvs}; learned/fixed {tot['learned'] / tot['fixed']:.4f}, "
            f"learned/ritz {tot['learned'] / tot['ritz']:.4f}; choices {choices}")
    assert tot["learned"] <= tot["fixed"]
    assert tot["learned"] <= 0.92 * tot["ritz"]
    assert spmvs["learned"] <= spmvs["fixed"]
    assert spmvs["learned"] < spmvs["ritz"]
