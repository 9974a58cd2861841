"""Caller-owned device buffers (sord_set_tensor): the kernels work in the
caller's tensors, with the results of the library's own buffers."""
import numpy as np
import pytest

from conftest import rel_l2, run_case
import cases

pytestmark = pytest.mark.gpu

STATE = ['vx', 'vy', 'vz', 'ux', 'uy', 'uz']
MATERIAL = ['lam', 'mu', 'yy', 'mr', 'gam']


def _adopt(names, keep):
    """probe for run_case: adopt a fresh tensor (filled with NaN, which the
    adoption's copy overwrites) for each name"""
    import torch

    def probe(s):
        n = s.tensor_numel()
        dt = torch.float32 if s.realsize == 4 else torch.float64
        for name in names:
            t = torch.full((n,), float('nan'), dtype=dt, device='cuda')  # overwritten by the adoption
            s.set_tensor(name, t)
            keep[name] = t
    return probe


@pytest.mark.parametrize('case', ['explosion_pml', 'layered_free', 'bc_mix'])
def test_adopted_buffers_give_the_same_run(case):
    """Every state buffer (both of each double-buffered vector) and every
    material array in caller tensors: identical outputs (FP64); after one
    step the current state lies in one of the caller's two buffers (the
    fused kernels write the next one, the stage kernels update in place)."""
    base = run_case(case, 'hip', 8)
    keep = {}
    names = STATE + [n + '.next' for n in STATE] + MATERIAL

    def probe(s):
        _adopt(names, keep)(s)
        s.step(1)
        for n in STATE:
            assert s.field(n)[0] in (keep[n].data_ptr(), keep[n + '.next'].data_ptr()), n

    out = run_case(case, 'hip', 8, probe=probe)
    for k in base:
        assert np.array_equal(out[k], base[k]), (k, rel_l2(out[k], base[k]))


def test_edited_material_enters_the_next_step():
    """bc_mix staged with one material and run on the material arrays of
    another (copied into adopted tensors before the first step) gives the
    run staged with the other material (FP64: same arithmetic on the same
    values)."""
    import torch
    import os
    import shutil
    import tempfile
    from coseis_amd import sord, solver
    other = {'rho': [2500.0], 'vp': [5000.0], 'vs': [2900.0], 'gam': [0.03]}
    ref = run_case('bc_mix', 'hip', 8, prm_update=other)
    # the material arrays of a handle staged with the other material
    prm, files = cases.CASES['bc_mix']()
    prm.update(other)
    d = tempfile.mkdtemp(prefix='sord-mat-')
    cwd = os.getcwd()
    try:
        os.chdir(d)
        cases.write_inputs(files, 8)
        text = sord.stage(prm, realsize=8)[0]
        src = solver.Solver(text, realsize=8)
    finally:
        os.chdir(cwd)
        shutil.rmtree(d)
    keep = {}

    def probe(s):
        _adopt(MATERIAL, keep)(s)
        n = s.tensor_numel()
        for name in MATERIAL:
            ptr, shape, strides, off = src.field(name)
            u = torch.empty(n, dtype=torch.float64, device='cuda')
            solver.copy_probe(ptr, u.data_ptr(), n * 8)
            keep[name].copy_(u)
        torch.cuda.synchronize()

    out = run_case('bc_mix', 'hip', 8, probe=probe)
    src.close()
    material_outputs = {'rho', 'vp', 'vs', 'lam', 'mu', 'nu', 'gam'}
    for k in ref:
        if k.split('.')[0] in material_outputs:
            continue
        assert rel_l2(out[k], ref[k]) < 1e-13, (k, rel_l2(out[k], ref[k]))


def test_adoption_refusals():
    """Too small, overlapping, unknown or wrong-precision buffers are refused
    with SORD_EINPUT and leave the handle working."""
    import torch
    from coseis_amd import solver

    def probe(s):
        n = s.tensor_numel()
        small = torch.zeros(n - 1, dtype=torch.float64, device='cuda')
        with pytest.raises(solver.SolverError) as e:
            s.set_tensor('vx', small)
        assert e.value.code == 1
        big = torch.zeros(n + 8, dtype=torch.float64, device='cuda')
        s.set_tensor('vx', big)
        with pytest.raises(solver.SolverError):  # same memory for a second array
            s.set_tensor('vy', big[4:])
        with pytest.raises(solver.SolverError):
            s.set_tensor('rho', torch.zeros(n, dtype=torch.float64, device='cuda'))
        with pytest.raises(ValueError):
            s.set_tensor('vz', torch.zeros(n, dtype=torch.float32, device='cuda'))

    base = run_case('explosion_pml', 'hip', 8)
    out = run_case('explosion_pml', 'hip', 8, probe=probe)
    for k in base:
        assert np.array_equal(out[k], base[k]), k
