import glob
import os
import re
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
sys.path.insert(0, os.path.join(ROOT, 'oracle'))  # closed forms of the checks (kostrov)

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


def run_case(case, runner, realsize=4, prm_update=None, fused=True, tiles=None, side_streams=False, probe=None,
             device_init=True):
    """Stage a case in a scratch directory and run it; returns {name: array}.
    probe(solver), if given, is called on the HIP solver before the run."""
    prm, files = (cases.CASES.get(case) or cases.ORACLE_CASES[case])()
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
            with solver.Solver(text, realsize=realsize, fused=fused, tiles=tiles, side_streams=side_streams,
                               device_init=device_init) as s:
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


def vector_groups(keys):
    """Output names grouped into vectors: the x, y, z components of one
    receiver, plane or slip field ('p1-vx', 'p1-vy', 'p1-vz'; 'sux', 'suy',
    'suz') form one vector: a component that vanishes by symmetry has no
    relative error of its own."""
    groups, seen = [], set()
    for k in sorted(keys):
        if k in seen:
            continue
        m = re.match(r'^(.*(?:v|u|a|su|sv|sa|t|ts))([xyz])(.*)$', k)
        if m and not k.startswith('stats-'):
            trio = [m.group(1) + c + m.group(3) for c in 'xyz']
            if all(t in keys for t in trio):
                groups.append(trio)
                seen.update(trio)
                continue
        groups.append([k])
        seen.add(k)
    return groups


def group_rel_l2(out, ref, grp):
    return rel_l2(np.concatenate([np.ravel(out[g]) for g in grp]), np.concatenate([np.ravel(ref[g]) for g in grp]))


def rupture_times_agree(name, out, ref, dt, realsize):
    """Rupture time (trup) or arrest time (tarr) where the reference's fault
    node ruptured (trup < 1e8) or arrested (0 < tarr < 1e8): the same set of
    nodes, and times within one time step in FP32 or 1e-9 s in FP64.  Every
    rupture fixture ruptures somewhere; where the reference's run was too
    short for any node to arrest, none may arrest.  Returns the largest
    difference."""
    hit = (ref < 1e8) & (ref > 0 if name == 'tarr' else True)
    mine = (out < 1e8) & (out > 0 if name == 'tarr' else True)
    assert np.array_equal(mine, hit), (name, int(mine.sum()), int(hit.sum()), int((mine != hit).sum()))
    if name == 'tarr' and not hit.any():
        return 0.0
    assert hit.any(), name
    bar = dt if realsize == 4 else 1e-9
    err = float(np.abs(out[hit].astype('f8') - ref[hit]).max())
    assert err <= bar, (name, err, bar)
    return err


# Bars of the parity tests against the reference fixtures (relative L2; see
# tests/test_hip_parity.py for their reasons)
TOL = {4: 1e-4, 8: 1e-9}
# statistics sums whose order of summation is free (fault energies)
ORDER_FREE = {'stats-efric': 1e-6, 'stats-estrain': 1e-6, 'stats-moment': 1e-6, 'stats-mw': 1e-6}
TIMES = ('trup', 'tarr')


def check_outputs(case, out, ref, ref64, realsize, dt=None):
    """Every output of a run against the reference at the bars of the module."""
    assert sorted(out) == sorted(ref)
    if dt is None:
        dt = cases.CASES[case]()[0]['delta'][3]
    errs, bad, noisy = {}, {}, []
    if ref64 is not None and any(k in ref64 and ref64[k].shape != ref[k].shape for k in ref if k.startswith('stats-')):
        ref64 = None  # the FP64 fixture covers another step window: no conditioning estimate
    for grp in vector_groups(set(ref)):
        for k in grp:
            assert out[k].shape == ref[k].shape, k
        name = '|'.join(grp)
        if grp[0] in TIMES:
            errs[name] = rupture_times_agree(grp[0], out[grp[0]], ref[grp[0]], dt, realsize)
            continue
        if realsize == 8:
            bar = ORDER_FREE.get(grp[0], TOL[8])
            err = group_rel_l2(out, ref, grp)
        else:
            # the reference's own FP32-vs-FP64 distance
            same = ref64 is not None and all(g in ref64 and ref64[g].shape == ref[g].shape for g in grp)
            cond = group_rel_l2(ref, ref64, grp) if same else 0.0
            if cond <= TOL[4]:
                bar, err = TOL[4], group_rel_l2(out, ref, grp)
            else:
                noisy.append(name)
                bar, err = 1.5 * cond + TOL[4], group_rel_l2(out, ref64, grp)
        errs[name] = err
        if not err <= bar:
            bad[name] = (err, bar)
    print(case, realsize, {k: '%.2e' % v for k, v in errs.items()}, 'noisy:', noisy, flush=True)
    assert not bad, bad
    return errs


