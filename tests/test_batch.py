"""Batched ensembles: every scenario of a batch gives the output files of its
own run (the batch kernels are other instantiations of the same kernel
bodies, so FP64 agrees to rounding: 1e-12, 1e-9 with a fault, whose friction
law amplifies rounding), and each scenario's outputs match the reference's
own run of that case (tests/golden, FP64 bars of check_outputs)."""
import glob
import os
import shutil
import tempfile

import numpy as np
import pytest

from conftest import check_outputs, golden, rel_l2
import cases
from coseis_amd import sord

pytestmark = pytest.mark.gpu
NAMES = ['explosion_pml', 'layered_free', 'fault_sw', 'bc_mix', 'stages', 'finite_moment', 'finite_force']


def _stage(base, case, tag, realsize):
    prm, files = cases.CASES[case]()
    d = os.path.join(base, '%s-%s' % (case, tag))
    os.makedirs(d)
    cwd = os.getcwd()
    try:
        os.chdir(d)
        cases.write_inputs(files, realsize)
        text = sord.stage(prm, realsize=realsize)[0]
    finally:
        os.chdir(cwd)
    return d, text


def _outputs(d, realsize):
    out = {}
    for f in sorted(glob.glob(os.path.join(d, '*.bin'))):
        b = os.path.basename(f)
        if not b.startswith('prof-'):
            out[b] = np.fromfile(f, '<f%d' % realsize)
    return out


@pytest.mark.parametrize('graph', [False, True])
def test_batch_matches_separate_runs(graph):
    """graph: the merged launches replayed from HIP graphs (each ring slot's
    graph replayed many times, with the scenarios' own work between)."""
    from coseis_amd import solver
    realsize = 8
    base = tempfile.mkdtemp(prefix='sord-batch-')
    try:
        texts, dirs, single = [], [], []
        for c in NAMES:
            d, text = _stage(base, c, 'single', realsize)
            cwd = os.getcwd()
            os.chdir(d)
            try:
                with solver.Solver(text, realsize=realsize) as s:
                    s.run()
            finally:
                os.chdir(cwd)
            single.append(d)
            d, text = _stage(base, c, 'batch', realsize)
            texts.append(text)
            dirs.append(d)
        with solver.Batch(texts, dirs, realsize=realsize) as b:
            assert b.set_graph(graph) == graph
            b.run()
            assert b.set_graph(None) == graph  # the captures succeeded
        for c, ds, db in zip(NAMES, single, dirs):
            a, o = _outputs(ds, realsize), _outputs(db, realsize)
            assert sorted(a) == sorted(o), c
            bar = 1e-9 if c == 'fault_sw' else 1e-12
            for k in a:
                assert a[k].shape == o[k].shape, (c, k)
                assert rel_l2(o[k], a[k]) < bar, (c, k, rel_l2(o[k], a[k]))
            # and each scenario of the batch against the reference itself
            ref = golden(c, realsize)
            check_outputs(c, {k[:-4]: v for k, v in o.items() if k[:-4] in ref or k.startswith('stats-')},
                          ref, ref, realsize)
    finally:
        shutil.rmtree(base)
