"""
Write the golden fixtures of tests/cases.py from the reference solver.

Runs the amdflang build of the reference SORD (oracle/_ref/r{4,8}/sord.x, made
by oracle/build_ref.sh from /root/reference/sord) on every case in a scratch
directory and stores every output file it writes as one array of an .npz
archive per case and precision: tests/golden/<case>_r4.npz, <case>_r8.npz.
Profile files (timings) are not stored.  The cases are those of
tests/cases.py: CASES (small meshes) and WORKLOADS (the benchmark workloads
at their own sizes, whose definitions depend on the precision; the two
16 M-node meshes take 10 to 50 minutes each on 8 cores).

    python tests/golden/make_golden.py [--r4 | --r8] [case ...]
    python tests/golden/make_golden.py --material

--material writes tests/golden/material_json.json: the material.json text the
reference writes at initialisation (its list-directed layout) for every case
of CASES, in FP32 and FP64.
"""
import glob
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))
from coseis_amd import sord  # noqa: E402
import cases  # noqa: E402


def case_inputs(case, realsize):
    """(prm, files) of a case of tests/cases.py."""
    if case in cases.CASES:
        return cases.CASES[case]()
    return cases.WORKLOADS[case](realsize)


def run_reference(case, realsize, exe):
    prm, files = case_inputs(case, realsize)
    d = tempfile.mkdtemp(prefix='golden-')
    cwd = os.getcwd()
    try:
        os.chdir(d)
        cases.write_inputs(files, realsize)
        sord.stage(prm, realsize=realsize)
        subprocess.run([exe], check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
        out = {}
        dt = '<f%d' % realsize
        for f in sorted(glob.glob('*.bin')):
            if f.startswith('prof-') or f in files:
                continue
            out[f[:-4]] = np.fromfile(f, dt)
        return out
    finally:
        os.chdir(cwd)
        shutil.rmtree(d)


def material_texts():
    out = {}
    for case in cases.CASES:
        out[case] = {}
        for rs in (4, 8):
            exe = os.path.join(ROOT, 'oracle', '_ref', 'r%d' % rs, 'sord.x')
            prm, files = cases.CASES[case]()
            prm['shape'] = prm['shape'][:3] + [1]
            d = tempfile.mkdtemp(prefix='golden-')
            cwd = os.getcwd()
            try:
                os.chdir(d)
                cases.write_inputs(files, rs)
                sord.stage(prm, realsize=rs)
                subprocess.run([exe], check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
                out[case][str(rs)] = open('material.json').read()
            finally:
                os.chdir(cwd)
                shutil.rmtree(d)
    import json
    with open(os.path.join(HERE, 'material_json.json'), 'w') as fh:
        json.dump(out, fh, indent=1, sort_keys=True)


def main():
    args = sys.argv[1:]
    if '--material' in args:
        return material_texts()
    sizes = [4, 8]
    if '--r4' in args or '--r8' in args:
        sizes = [rs for rs in sizes if '--r%d' % rs in args]
    names = [a for a in args if not a.startswith('--')] or list(cases.CASES)
    for rs in sizes:
        exe = os.path.join(ROOT, 'oracle', '_ref', 'r%d' % rs, 'sord.x')
        if not os.path.exists(exe):
            sys.exit('reference not built: run oracle/build_ref.sh')
        for case in names:
            out = run_reference(case, rs, exe)
            path = os.path.join(HERE, '%s_r%d.npz' % (case, rs))
            np.savez_compressed(path, **out)
            print(path, sorted(out))


if __name__ == '__main__':
    main()
