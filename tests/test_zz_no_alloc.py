"""sord_step_static (no device allocation, scratch in a static device array)
against sord_step on the same runs: the same kernels in the same order, so
the output files are identical.  Two runs in a row on the static path, the
second on a smaller mesh (the static scratch keeps the first run's contents),
and one run alternating the two entry points on one handle (the statistics
slots move between the handle's buffers and the static scratch)."""
import os
import shutil
import tempfile

import numpy as np
import pytest

from conftest import collect
import cases
from coseis_amd import sord

pytestmark = pytest.mark.gpu


def run(case, realsize, mode, chunk=7):
    """Stage a case and advance it in chunks of `chunk` steps with
    sord_step ('step'), sord_step_static ('static') or the two in turn
    ('mixed'); returns the output files."""
    from coseis_amd import solver
    prm, files = cases.CASES[case]()
    d = tempfile.mkdtemp(prefix='sord-static-')
    cwd = os.getcwd()
    try:
        os.chdir(d)
        cases.write_inputs(files, realsize)
        text = sord.stage(prm, realsize=realsize)[0]
        with solver.Solver(text, realsize=realsize) as s:
            it, nt = s.status()
            q = 0
            while it < nt:
                static = mode == 'static' or (mode == 'mixed' and q % 2 == 1)
                (s.step_static if static else s.step)(chunk)
                it, nt = s.status()
                q += 1
        return collect(d, files, realsize)
    finally:
        os.chdir(cwd)
        shutil.rmtree(d)


@pytest.mark.parametrize('realsize', [4, 8])
def test_static_step_matches_step(realsize):
    # a rupture on 36 x 24 x 30 nodes, then an explosion on 21^3 nodes
    for case in ('fault_prop', 'explosion_pml'):
        print('static step', case, realsize, flush=True)
        a = run(case, realsize, 'step')
        b = run(case, realsize, 'static')
        assert sorted(a) == sorted(b)
        for k in a:
            assert np.array_equal(a[k], b[k]), (case, k)


def test_static_and_plain_steps_alternate():
    print('static step alternating', flush=True)
    a = run('layered_free', 4, 'step')
    b = run('layered_free', 4, 'mixed')
    assert sorted(a) == sorted(b)
    for k in a:
        assert np.array_equal(a[k], b[k]), k
