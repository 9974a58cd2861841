"""HIP solver against the reference's golden fixtures (FP32 and FP64).

Tolerances: the HIP kernels use the rectilinear form of the support operators
(differences first), exact box volumes and fused multiply-adds, so FP32 results
differ from the reference's FP32 by rounding only; the bar is the target
FP32 relative L2 error of 1e-4 on every output, and 1e-9 for FP64.
"""
import numpy as np
import pytest

from conftest import golden, rel_l2, run_case
import cases

pytestmark = pytest.mark.gpu
TOL = {4: 1e-4, 8: 1e-9}
# statistics sums whose order of summation is free (fault energies)
ORDER_FREE = {'stats-efric': 1e-6, 'stats-estrain': 1e-6, 'stats-moment': 1e-6, 'stats-mw': 1e-6}


@pytest.mark.parametrize('realsize', [4, 8])
@pytest.mark.parametrize('case', sorted(cases.CASES))
def test_hip_matches_reference(case, realsize):
    """Every output file against the reference run in the same precision.

    Rupture runs in FP32 are checked against the FP64 reference instead: the
    reference's own FP32 run departs from its FP64 run by more than 1e-4 on
    slip-rate outputs (threshold crossings of the friction law amplify
    rounding), so the bar there is to be no farther from the FP64 oracle than
    the reference's FP32 run is (plus 1e-4).
    """
    print('case', case, 'realsize', realsize, flush=True)
    ref = golden(case, realsize)
    out = run_case(case, 'hip', realsize)
    assert sorted(out) == sorted(ref)
    rupture = cases.CASES[case]()[0].get('faultnormal', 'none') != 'none'
    ref64 = golden(case, 8)
    errs, bars = {}, {}
    for k in ref:
        assert out[k].shape == ref[k].shape, k
        if realsize == 4 and rupture:
            errs[k] = rel_l2(out[k], ref64[k])
            bars[k] = max(TOL[4], 1.5 * rel_l2(ref[k], ref64[k]))
        else:
            errs[k] = rel_l2(out[k], ref[k])
            bars[k] = ORDER_FREE.get(k, TOL[realsize]) if realsize == 8 else TOL[realsize]
    print(case, realsize, {k: '%.2e/%.0e' % (errs[k], bars[k]) for k in errs}, flush=True)
    bad = {k: (errs[k], bars[k]) for k in errs if not errs[k] <= bars[k]}
    assert not bad, bad


@pytest.mark.parametrize('case', ['explosion_pml', 'layered_free', 'fault_sw', 'finite_potency', 'finite_moment',
                                  'finite_force'])
def test_fused_kernel_matches_stage_kernels(case):
    """The fused interior kernel (differences of true gradients) and the stage
    kernels (volume-integrated operators) are the same algebra: in FP64 they
    agree to rounding; in FP32 to the case's own FP32 rounding noise, measured
    as the distance of the FP32 stage run from the FP64 stage run (the rupture
    case amplifies rounding through its friction law).  The finite sources
    enter the fused kernels' stress (potency, moment) or force from a box the
    source kernel fills each step."""
    names = []
    a64 = run_case(case, 'hip', 8, fused=False)
    b64 = run_case(case, 'hip', 8, fused=True, probe=lambda s: names.append(s.bench(1)[3]))
    assert names == ['k_fused'], names  # the fused kernels ran the steps
    for k in a64:
        assert rel_l2(b64[k], a64[k]) < 1e-9, (8, k, rel_l2(b64[k], a64[k]))
    a = run_case(case, 'hip', 4, fused=False)
    b = run_case(case, 'hip', 4, fused=True)
    for k in a:
        if k.startswith('stats-'):
            continue
        bar = max(3e-5, 2 * rel_l2(a[k], a64[k]))
        assert rel_l2(b[k], a[k]) < bar, (4, k, rel_l2(b[k], a[k]), bar)


@pytest.mark.parametrize('case', ['explosion_pml', 'fault_sw'])
def test_results_do_not_depend_on_tiles(case):
    """The tile shape of the fused kernels changes which block computes a
    node, not the algebra of that node.  It can change rounding: the cell
    computation is inlined several times (tile prologue, the two planes of the
    unrolled loop, each wave count's kernel) and the compiler may contract
    other multiply-adds in each copy, so FP64 outputs agree to 1e-12 (1e-9
    with a fault, whose friction law amplifies rounding).  The base run puts
    the whole boundary of these small meshes on the general kernel; with the
    tiles set, the face slabs run on their own kernels (the one-axis
    instantiations with the PML memory loaded a plane ahead), so the two are
    checked against each other."""
    bar = 1e-9 if case == 'fault_sw' else 1e-12
    base = run_case(case, 'hip', 8)
    for tiles in ((8, 3, 5), (8, 40, 2), (16, 64, 16), (4, 64, 16)):
        out = run_case(case, 'hip', 8, tiles=tiles)
        for k in base:
            assert rel_l2(out[k], base[k]) < bar, (tiles, k, rel_l2(out[k], base[k]))


@pytest.mark.parametrize('case', ['explosion_pml', 'layered_free'])
def test_side_streams_do_not_change_results(case):
    """The boundary kernels on side streams beside the interior kernel
    (opt-in), or in order on one stream (the default): the same kernels
    compute every node, so within one process the outputs are identical
    (FP64)."""
    a = run_case(case, 'hip', 8, side_streams=True)
    b = run_case(case, 'hip', 8)
    for k in a:
        assert np.array_equal(a[k], b[k]), (k, rel_l2(a[k], b[k]))


PML_CASES = ['explosion_pml', 'layered_free', 'sources_scattered', 'stages', 'strided_input']


@pytest.mark.parametrize('realsize', [4, 8])
@pytest.mark.parametrize('case', PML_CASES)
def test_face_slab_kernels_match_reference(case, realsize):
    """With the tile shape set, each PML face slab runs on its own boundary
    kernel (the one-damped-axis instantiations, which load the PML memory a
    plane ahead) instead of the general kernel that small meshes use by
    default: every output against the reference at the bars of
    test_hip_matches_reference."""
    print('case', case, 'realsize', realsize, flush=True)
    ref = golden(case, realsize)
    names = []
    out = run_case(case, 'hip', realsize, tiles=(8, 16, 4), probe=lambda s: names.append(s.boundary_masks()))
    assert any(m in (1, 2, 4) for m in names[0]), names  # one-axis kernels ran
    for k in ref:
        bar = ORDER_FREE.get(k, TOL[realsize]) if realsize == 8 else TOL[realsize]
        err = rel_l2(out[k], ref[k])
        assert err <= bar, (k, err, bar)
