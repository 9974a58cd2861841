"""The cst.sord staging logic (host side, no GPU)."""
import json
import os

import numpy as np
import pytest

from coseis_amd import sord


def line1(prm):
    return sord.stage(prm, write=False)[0].split('\n')[0].split()


def test_sord_in_order_and_pml_box():
    prm = {'shape': [21, 22, 23, 5], 'bc1': ['pml', 'free', 'pml'], 'bc2': ['pml', 'pml', 'free'], 'npml': 4}
    text, meta, p, fio, job = sord.stage(prm, write=False)
    toks = text.split('\n')[0].split()
    # affine(9) bc1(3) bc2(3) debug delta(4) faultnormal faultopening gam_max gam_min
    # gridnoise hourglass(2) hypocenter(3) i1pml(3) i2pml(3) ...
    assert toks[9:12] == ['10', '0', '10'] and toks[12:15] == ['10', '10', '0']
    i1 = toks[30:33]
    i2 = toks[33:36]
    # npml damped planes on PML faces and none elsewhere (defect D13 fixed)
    assert i1 == ['3', '-1', '3']
    assert i2 == ['18', '19', '24']
    assert text.split('\n')[0].endswith(' ')
    assert p['nfieldio'] == len(fio) == 4  # rho vp vs gam


def test_slices_whole_axis_and_node_shape():
    prm = {'shape': [10, 11, 12, 7], 'vx': [([[], [], -1, -1], '=>', 'vx.bin')],
           'ux': [([[2, 8, 2], 3, [], [0, None, 2]], '=>', 'ux.bin')]}
    text, meta, p, fio, job = sord.stage(prm, write=False)
    rec = {r[9]: r for r in fio}
    vx = rec['vx.bin'][2]
    assert vx == [[0, 10, 1], [0, 11, 1], [11, 12, 1], [6, 7, 1]]  # [nx, ny, nz] (D5)
    assert meta['shapes']['vx.bin'] == [10, 11]
    assert meta['shapes']['ux.bin'] == [3, 12, 4]
    assert meta['indices']['ux.bin'] == [[2, 8, 2], 3, [0, 12], [0, 7, 2]]


def test_subcell_op_records():
    prm = {'shape': [10, 10, 10, 5], 'mxx': [([4.5, 4.0, 3.25, []], '.', 1e15, 'brune', 0.1)]}
    text, meta, p, fio, job = sord.stage(prm, write=False)
    r = [x for x in fio if x[0] == 'mxx'][0]
    assert r[4] == [4.5, 4.0, 3.25]
    assert r[2][:3] == [[4, 5, 1], [4, 5, 1], [3, 4, 1]]
    assert r[8] == '.' and r[9] == 'brune'


def test_fault_hypocenter_bc():
    prm = {'shape': [10, 10, 10, 5], 'faultnormal': '+z', 'hypocenter': [5.0, 5.0, 0.0]}
    text, meta, p, fio, job = sord.stage(prm, write=False)
    assert p['bc1'][2] == -2  # fault at the boundary: -cell symmetry (cst/sord.py:313-317)


def test_errors():
    with pytest.raises(sord.SordError):
        sord.stage({'shape': [10, 10, 10, 5], 'vx': [([], '!!', 1.0)]}, write=False)
    with pytest.raises(sord.SordError):
        sord.stage({'shape': [10, 10, 10, 5], 'vm2': [([], '=', 1.0)]}, write=False)
    with pytest.raises(sord.SordError):
        sord.stage({'shape': [10, 10, 10, 5], 'bc1': ['pml'] * 3, 'bc2': ['pml'] * 3, 'npml': 5}, write=False)
    with pytest.raises(sord.SordError):
        sord.stage({'shape': [10, 10, 10, 5], 'vx': [([], '=', 1.0, 'nosuch', 0.1)]}, write=False)


def test_stage_writes_files(tmp_path):
    os.chdir(tmp_path)
    sord.stage({'shape': [8, 8, 8, 3], 'vx': [([[], [], [], -1], '=>', 'vx.bin')]})
    for f in ['sord.in', 'parameters.json', 'job.json', 'sord.json', 'meta.json']:
        assert os.path.exists(f)
    meta = json.load(open('meta.json'))
    assert meta['dtype'] == '<f4' and meta['shapes']['vx.bin'] == [8, 8, 8]
