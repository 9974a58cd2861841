"""The CPU oracle against the reference's golden fixtures (no GPU).

The oracle restates the reference SORD solver statement by statement; the
fixtures are the output files of the amdflang build of the reference on the
same sord.in.  FP32 and FP64 agree to the last bit apart from sums whose
order the compilers may choose (statistics of the fault energies).
"""
import numpy as np
import pytest

from conftest import golden, rel_l2, run_case
import cases

ORDER_FREE = {'stats-efric', 'stats-estrain', 'stats-moment', 'stats-mw'}


@pytest.mark.parametrize('realsize', [4, 8])
@pytest.mark.parametrize('case', sorted(cases.CASES))
def test_oracle_matches_reference(case, realsize):
    ref = golden(case, realsize)
    out = run_case(case, 'oracle', realsize)
    assert sorted(out) == sorted(ref)
    for k in ref:
        assert out[k].shape == ref[k].shape, k
        if k in ORDER_FREE:
            assert rel_l2(out[k], ref[k]) < 1e-5, k
        else:
            assert np.array_equal(out[k], ref[k]), (k, rel_l2(out[k], ref[k]))


def test_oracle_catches_planted_error():
    """A wrong PML coefficient in the oracle input must break parity: the
    comparison is sensitive to the damped planes."""
    ref = golden('explosion_pml', 4)
    out = run_case('explosion_pml', 'oracle', 4, prm_update={'vpml': 3000.0})
    assert rel_l2(out['vx'], ref['vx']) > 1e-4
