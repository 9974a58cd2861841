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


def test_kostrov_constant_known_value():
    """Dahlen's C for the reference's parameters (vrup = 0.9 vs,
    vp = 6000, vs = 3464) is the 0.82 the reference's test names as its
    alternative constant (cst/tests/sord_kostrov.py:58), within 0.1 %."""
    import kostrov
    c = kostrov.cee(0.9 * 3464.0 / 6000.0, 0.9)
    assert abs(c / 0.82 - 1) < 1e-3, c


def test_reference_kostrov_run_matches_closed_form():
    """The reference's own run of the Kostrov case (the fixture) against the
    closed form at the last step (cst/tests/sord_kostrov.py:54-76): slip rate
    at 2 km on four azimuths, reduced time (nt - 1.5) dt - r / vrup.  The
    reference states 1.5 % for its quarter model; on this full-plane mesh its
    run sits 1.4-2.0 % above the closed form (numerical dispersion of the
    100 m mesh), so the bar is 3 %; a wrong stress drop, rupture speed or
    time origin moves the points by far more."""
    import kostrov
    prm = cases.CASES['kostrov']()[0]
    nt, dt = prm['shape'][3], prm['delta'][3]
    vr = prm['vrup']
    for realsize in (4, 8):
        ref = golden('kostrov', realsize)
        for p, (x, y) in zip('abcd', cases.KOSTROV_POINTS):
            r = prm['delta'][0] * np.hypot(x - cases.KOSTROV_CENTRE[0], y - cases.KOSTROV_CENTRE[1])
            v = kostrov.slip_rate(2670.0, 6000.0, 3464.0, vr, 10e6, r, (nt - 1.5) * dt - r / vr)
            err = (ref['p20' + p][-1] - v) / v
            assert abs(err) < 0.03, (realsize, p, err)


@pytest.mark.parametrize('realsize', [4, 8])
def test_oracle_cube_ops_match_closed_form(realsize):
    """Cube ops ('=@', '*@', '+@', tests/cases.py cube_ops): the cells whose
    centres lie in each box, from the mesh's closed-form cell centres, get
    the op in the order listed (the material written at initialisation)."""
    dt = np.float32 if realsize == 4 else np.float64
    out = run_case('cube_ops', 'oracle', realsize)
    nx, ny, nz = cases.cube_ops()[0]['shape'][:3]
    c = [(np.arange(n - 1) + 0.5) * 100.0 for n in (nx, ny, nz)]
    X, Y, Z = np.meshgrid(*c, indexing='ij')

    def inbox(b):
        return ((X >= b[0][0]) & (X <= b[1][0]) & (Y >= b[0][1]) & (Y <= b[1][1]) &
                (Z >= b[0][2]) & (Z <= b[1][2]))
    rho = np.full(X.shape, dt(2670.0))
    rho[inbox(cases.CUBE_BOXES[0])] = dt(2500.0)
    rho[inbox(cases.CUBE_BOXES[1])] *= dt(0.98)
    vs = np.full(X.shape, dt(3464.0))
    vs[inbox(cases.CUBE_BOXES[2])] += dt(-150.0)
    gam = np.zeros(X.shape, dt)
    m = inbox(cases.CUBE_BOXES[3])
    m[:1] = m[12:] = False  # the op's index range: cells 1..11 along x
    gam[m] = dt(0.02)
    for name, want in (('rho', rho), ('vs', vs), ('gam', gam)):
        got = out['out-' + name].reshape(nz - 1, ny - 1, nx - 1).T
        assert np.array_equal(got, want), name
    assert (rho != 2670.0).sum() > 500 and m.sum() > 500


@pytest.mark.parametrize('realsize', [4, 8])
def test_oracle_point_series_match_pulses(realsize):
    """'.<' (a point time series read from a file, one value per step): the
    series of each step's pulse value (tests/cases.py point_series) gives the
    run of the same sources as '.' ops with those pulses (point_pulses): the
    record of step it + 1 is applied at that step, with the '.' op's
    weights.  A series shifted by one step misses by 0.1 to 0.7 (relative
    L2).  FP32: the file's values are rounded FP64 pulses, the '.' op's are
    FP32 pulses."""
    a = run_case('point_series', 'oracle', realsize)
    b = run_case('point_pulses', 'oracle', realsize)
    assert sorted(a) == sorted(b)
    bar = {4: 1e-5, 8: 1e-12}[realsize]
    for k in a:
        assert rel_l2(a[k], b[k]) <= bar, (k, rel_l2(a[k], b[k]))
