import glob
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, TESTS)

import cases  # noqa: E402
from coseis_amd import sord  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (HIP device)')


def oracle_exe(realsize):
    """The C oracle (test infrastructure), built on demand with make."""
    exe = os.path.join(ROOT, 'oracle', '_build', 'sord_oracle_r%d' % realsize)
    if not os.path.exists(exe):
        subprocess.run(['make', '-s', '-C', os.path.join(ROOT, 'oracle')], check=True)
    return exe


def golden(case, realsize=4):
    with np.load(os.path.join(TESTS, 'golden', '%s_r%d.npz' % (case, realsize))) as z:
        return {k: z[k] for k in z.files}


def collect(d, files, realsize):
    out = {}
    for f in sorted(glob.glob(os.path.join(d, '*.bin'))):
        b = os.path.basename(f)
        if b.startswith('prof-') or b in files:
            continue
        out[b[:-4]] = np.fromfile(f, '<f%d' % realsize)
    return out


def run_case(case, runner, realsize=4, prm_update=None, fused=True, tiles=None, side_streams=False, probe=None):
    """Stage a case in a scratch directory and run it; returns {name: array}.
    probe(solver), if given, is called on the HIP solver before the run."""
    prm, files = cases.CASES[case]()
    if prm_update:
        prm.update(prm_update)
    d = tempfile.mkdtemp(prefix='sord-%s-' % case)
    cwd = os.getcwd()
    try:
        os.chdir(d)
        cases.write_inputs(files, realsize)
        text = sord.stage(prm, realsize=realsize)[0]
        if runner == 'oracle':
            env = dict(os.environ, OMP_NUM_THREADS=os.environ.get('OMP_NUM_THREADS', '4'))
            subprocess.run([oracle_exe(realsize)], check=True, stdout=subprocess.DEVNULL, env=env)
        elif runner == 'hip':
            from coseis_amd import solver
            with solver.Solver(text, realsize=realsize, fused=fused, tiles=tiles, side_streams=side_streams) as s:
                if probe:
                    probe(s)
                s.run()
        else:
            raise ValueError(runner)
        return collect(d, files, realsize)
    finally:
        os.chdir(cwd)
        shutil.rmtree(d)


def rel_l2(a, b):
    a = np.asarray(a, 'f8')
    b = np.asarray(b, 'f8')
    n = np.linalg.norm(b)
    if n == 0:
        return float(np.linalg.norm(a - b))
    return float(np.linalg.norm(a - b) / n)
