"""Register budget of the solve-loop kernels, from the compiler's resource
remarks that the library build records (cgxx_amd/lib/kernel_resources.json).

Measured on gfx950 (Clenshaw step, 256^3, grid of 2048 blocks): at 80 SGPRs
and 64 VGPRs a CU holds 8 blocks of 256 threads and the step takes 200 us;
at 82-91 SGPRs it takes 275-295 us although the compiler still reports 8
waves per SIMD, because fewer blocks are resident and the grid runs in two
rounds.  Every kernel of the FP64 and mixed-precision loops therefore stays
at <= 80 SGPRs, <= 64 VGPRs and no scratch, with one measured exception:
the Clenshaw steps' four-rows-per-trip instances (k_poly_step<..., 4>) hold
up to 96 VGPRs (5-6 waves per SIMD) for twice the loads in flight, 75.4 ->
74.3 us per 256^3 step; the library takes them only where every block
walks at least 8 chunks and the 64-VGPR two-row instances elsewhere.
"""
import json
import os
import re

import pytest

from cgxx_amd import build

HOT = [
    r"k_pcg_matvec_dot_crs\(", r"k_pcg_matvec_dot_ell<true>", r"k_matvec_crs<double, double>",
    r"k_matvec_ell<double, double, true>", r"k_pcg_update<", r"k_pcg_xpay<", r"k_mp_matvec_dot<",
    r"k_mp_update\(", r"k_mp_xpay\(", r"k_poly_step<", r"k_mp_xpay_poly\(",
]


@pytest.fixture(scope="module")
def resources():
    if not os.path.exists(build.RESOURCES):
        pytest.fail(f"{build.RESOURCES} missing: run python -m cgxx_amd.build")
    with open(build.RESOURCES) as f:
        return json.load(f)


def test_every_hot_kernel_is_recorded(resources):
    for pat in HOT:
        assert any(re.search(pat, k) for k in resources), pat


# template arguments FINAL, K, MODE, HALF, NT, ROWS, SYM
FOUR_ROWS = re.compile(r"k_poly_step<[^(]*, 4, (true|false)>\(")


@pytest.mark.parametrize("pat", HOT)
def test_hot_kernel_register_budget(resources, pat):
    for name, r in resources.items():
        if not re.search(pat, name):
            continue
        vgpr_max = 96 if FOUR_ROWS.search(name) else 64
        assert r["sgpr"] <= 80, (name, r)
        assert r["vgpr"] + r.get("agpr", 0) <= vgpr_max, (name, r)
        assert r["scratch"] == 0, (name, r)


def test_four_row_instances_exist_and_two_row_ones_keep_the_budget(resources):
    four = [k for k in resources if FOUR_ROWS.search(k)]
    two = [k for k in resources if re.search(r"k_poly_step<", k) and not FOUR_ROWS.search(k)]
    assert four and two
    for name in two:
        assert resources[name]["vgpr"] <= 64, name


def test_remark_parser():
    text = "\n".join([
        "a.hip:1:1: remark: Function Name: _Z3foov [-Rpass-analysis=kernel-resource-usage]",
        "a.hip:1:1: remark:     TotalSGPRs: 81 [-Rpass-analysis=kernel-resource-usage]",
        "a.hip:1:1: remark:     VGPRs: 40 [-Rpass-analysis=kernel-resource-usage]",
        "    1 | __global__ void foo() {}",
        "a.hip:1:1: remark:     ScratchSize [bytes/lane]: 0 [-Rpass-analysis=kernel-resource-usage]",
        "a.hip:1:1: remark:     Occupancy [waves/SIMD]: 8 [-Rpass-analysis=kernel-resource-usage]",
    ])
    got = build.parse_resource_remarks(text)
    assert list(got.values()) == [{"sgpr": 81, "vgpr": 40, "scratch": 0, "occupancy": 8}]
