"""HIP solver against the reference's golden fixtures (FP32 and FP64).

Tolerances: the HIP kernels use the rectilinear form of the support operators
(differences first), exact box volumes and fused multiply-adds, so FP32 results
differ from the reference's FP32 by rounding only; the bar is the north-star
FP32 relative L2 error of 1e-4 on every output, and 1e-9 for FP64.
"""
import numpy as np
import pytest

from conftest import golden, rel_l2, run_case
import cases

pytestmark = pytest.mark.gpu
TOL = {4: 1e-4, 8: 1e-9}


@pytest.mark.parametrize('realsize', [4, 8])
@pytest.mark.parametrize('case', sorted(cases.CASES))
def test_hip_matches_reference(case, realsize):
    print('case', case, 'realsize', realsize, flush=True)
    ref = golden(case, realsize)
    out = run_case(case, 'hip', realsize)
    assert sorted(out) == sorted(ref)
    errs = {}
    for k in ref:
        assert out[k].shape == ref[k].shape, k
        errs[k] = rel_l2(out[k], ref[k])
    print(case, realsize, {k: '%.2e' % v for k, v in errs.items()}, flush=True)
    bad = {k: v for k, v in errs.items() if not v <= TOL[realsize]}
    assert not bad, bad


@pytest.mark.parametrize('case', ['explosion_pml', 'fault_sw'])
def test_stage_path_matches_fused_path(case):
    """Stage kernels only (no fused update) give the same results."""
    a = run_case(case, 'hip', 4, fused=False)
    b = run_case(case, 'hip', 4, fused=True)
    for k in a:
        assert rel_l2(a[k], b[k]) < 1e-6, k
