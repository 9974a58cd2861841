"""
Support Operator Rupture Dynamics: the ``cst.sord`` user interface.

Same parameter dictionary, field names, boundary-condition ids and time
functions as cst/sord.py:16-186, same ``sord.in`` text format
(cst/sord.py:538-544, read back by sord/input_output.f90:51-92) and the same
``meta.json`` / flat-binary output contract (cst/sord.py:546-553,
cst/bin2npy.py:10-27).  ``run()`` hands the ``sord.in`` text to the HIP
solver library through its C ABI (``include/sord_amd.h``) instead of
launching ``sord.x`` as a child process (cst/job.py:192-194).

The staging logic restores what cst/sord.py intends where the shipped code
raises or mis-indexes:

* D2  ``typed_dict.__setitem__`` raises on matching types (cst/sord.py:219-223):
      plain dictionaries are used here.
* D3  ``slice(*[])`` for the ``[]`` shorthand (cst/sord.py:228-235): ``[]`` is
      the whole axis, as README.md:166-170 documents.
* D4  tuple item assignment for sub-cell ``.`` ops (cst/sord.py:407-412):
      slices are lists here.
* D5  node shape ``[nz, ny, nz]`` (cst/sord.py:355): ``[nx, ny, nz]``.
* D13 PML index box (cst/sord.py:319-330 with sord/input_output.f90:62).  The
      shipped values make the Fortran treat ``npml + 1`` planes per PML face
      as damped while the damping tables hold ``npml`` entries
      (sord/boundary_pml.f90:136-150), so ``dn1(npml+1)`` etc. are read out of
      bounds (sord/acceleration.f90:300-306, sord/stress.f90:49-55), and every
      non-PML face gets a one-plane PML with the strongest coefficient.  Here
      ``i1pml = npml - 1`` and ``i2pml = n - npml + 1`` on PML faces and
      ``-1`` / ``n + 1`` elsewhere, which gives exactly ``npml`` damped node and
      cell planes per PML face and none on other faces (the original SORD
      layout), and keeps every table index in bounds.
"""
import copy
import json
import os
import sys

parameters = {
    'nproc3': [1, 1, 1],
    'nthread': -1,
    'mpin': 1,
    'mpout': 1,
    'itstats': 10,
    'itio': 50,
    'itbuff': 10,
    'debug': 0,
    'shape': [41, 41, 41, 41],
    'delta': [100.0, 100.0, 100.0, 0.0075],
    'affine': [[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]],
    'rexpand': 1.06,
    'n1expand': [0, 0, 0],
    'n2expand': [0, 0, 0],
    'gridnoise': 0.0,
    'tm0': 0.0,
    'rho': [2670.0],
    'vp': [6000.0],
    'vs': [3464.0],
    'gam': [0.0],
    'rho_min': -1.0,
    'rho_max': -1.0,
    'vp_min': -1.0,
    'vp_max': -1.0,
    'vs_min': -1.0,
    'vs_max': -1.0,
    'gam_min': -1.0,
    'gam_max': 0.8,
    'hourglass': [1.0, 1.0],
    'vdamp': -1.0,
    'bc1': ['free', 'free', 'free'],
    'bc2': ['free', 'free', 'free'],
    'npml': 10,
    'ppml': 2,
    'vpml': -1.0,
    'nsource': 0,
    'source': 'potency',
    'hypocenter': [0.0, 0.0, 0.0],
    'slipvector': [1.0, 0.0, 0.0],
    'faultnormal': 'none',
    'faultopening': 0,
    'vrup': -1.0,
    'rcrit': 1000.0,
    'trelax': 0.075,
    'svtol': 0.001,
}

# n node, c cell, f fault, ~ time varying, < input  (cst/sord.py:64-163)
fieldnames = {
    'x': ['n<', 'x', 'Node coordinate'],
    'y': ['n<', 'y', 'Node coordinate'],
    'z': ['n<', 'z', 'Node coordinate'],
    'fx': ['n~<', 'f_x', 'Force'],
    'fy': ['n~<', 'f_y', 'Force'],
    'fz': ['n~<', 'f_z', 'Force'],
    'ax': ['n~<', 'a_x', 'Acceleration'],
    'ay': ['n~<', 'a_y', 'Acceleration'],
    'az': ['n~<', 'a_z', 'Acceleration'],
    'vx': ['n~<', 'v_x', 'Velocity'],
    'vy': ['n~<', 'v_y', 'Velocity'],
    'vz': ['n~<', 'v_z', 'Velocity'],
    'ux': ['n~<', 'u_x', 'Displacement'],
    'uy': ['n~<', 'u_y', 'Displacement'],
    'uz': ['n~<', 'u_z', 'Displacement'],
    'am2': ['n~', '|a|', 'Acceleration magnitude'],
    'vm2': ['n~', '|v|', 'Velocity magnitude'],
    'um2': ['n~', '|u|', 'Displacement magnitude'],
    'rho': ['c<', '\\rho', 'Density'],
    'vp': ['c<', 'V_p', 'P-wave velocity'],
    'vs': ['c<', 'V_s', 'S-wave velocity'],
    'gam': ['c<', '\\gamma', 'Viscosity'],
    'xc': ['c', 'x', 'Cell coordinate'],
    'yc': ['c', 'y', 'Cell coordinate'],
    'zc': ['c', 'z', 'Cell coordinate'],
    'vc': ['c', 'V^C', 'Cell volume'],
    'nu': ['c', '\\nu', 'Poisson ratio'],
    'mu': ['c', '\\mu', 'Elastic modulus'],
    'lam': ['c', '\\lambda', 'Elastic modulus'],
    'pxx': ['c~<', 'p_{xx}', 'Seismic potency'],
    'pyy': ['c~<', 'p_{yy}', 'Seismic potency'],
    'pzz': ['c~<', 'p_{zz}', 'Seismic potency'],
    'pyz': ['c~<', 'p_{yz}', 'Seismic potency'],
    'pzx': ['c~<', 'p_{zx}', 'Seismic potency'],
    'pxy': ['c~<', 'p_{xy}', 'Seismic potency'],
    'mxx': ['c~<', 'm_{xx}', 'Seismic moment'],
    'myy': ['c~<', 'm_{yy}', 'Seismic moment'],
    'mzz': ['c~<', 'm_{zz}', 'Seismic moment'],
    'myz': ['c~<', 'm_{yz}', 'Seismic moment'],
    'mzx': ['c~<', 'm_{zx}', 'Seismic moment'],
    'mxy': ['c~<', 'm_{xy}', 'Seismic moment'],
    'wxx': ['c~<', '\\sigma_{xx}', 'Stress'],
    'wyy': ['c~<', '\\sigma_{yy}', 'Stress'],
    'wzz': ['c~<', '\\sigma_{zz}', 'Stress'],
    'wyz': ['c~<', '\\sigma_{yz}', 'Stress'],
    'wzx': ['c~<', '\\sigma_{zx}', 'Stress'],
    'wxy': ['c~<', '\\sigma_{xy}', 'Stress'],
    'wm2': ['c~', '||\\sigma||_F', 'Stress Frobenius norm'],
    'mus': ['f<', '\\mu_s', 'Static friction coefficient'],
    'mud': ['f<', '\\mu_d', 'Dynamic friction coefficient'],
    'dc': ['f<', 'D_c', 'Slip weakening distance'],
    'co': ['f<', 'co', 'Cohesion'],
    'tn': ['f<', '\\tau_n', 'Pre-traction normal component'],
    'ts': ['f<', '\\tau_s', 'Pre-traction strike component'],
    'td': ['f<', '\\tau_d', 'Pre-traction dip component'],
    'sxx': ['f<', '\\sigma_{xx}', 'Pre-stress'],
    'syy': ['f<', '\\sigma_{yy}', 'Pre-stress'],
    'szz': ['f<', '\\sigma_{zz}', 'Pre-stress'],
    'syz': ['f<', '\\sigma_{yz}', 'Pre-stress'],
    'szx': ['f<', '\\sigma_{zx}', 'Pre-stress'],
    'sxy': ['f<', '\\sigma_{xy}', 'Pre-stress'],
    'nsx': ['f', 'n_x', 'Fault surface normal'],
    'nsy': ['f', 'n_y', 'Fault surface normal'],
    'nsz': ['f', 'n_z', 'Fault surface normal'],
    '_f0': ['f<', 'f_0', 'Steady state friction at V_0'],
    '_fw': ['f<', 'f_w', 'Fully weakened fiction'],
    '_v0': ['f<', 'V_0', 'Reference slip velocity'],
    '_vw': ['f<', 'V_w', 'Weakening slip velocity'],
    '_ll': ['f<', 'L', 'State evolution distance'],
    '_af': ['f<', 'a', 'Direct effect parameter'],
    '_bf': ['f<', 'b', 'Evolution effect parameter'],
    'tx': ['f~', '\\tau_x', 'Traction'],
    'ty': ['f~', '\\tau_y', 'Traction'],
    'tz': ['f~', '\\tau_z', 'Traction'],
    'tsx': ['f~', '\\tau^s_x', 'Shear traction'],
    'tsy': ['f~', '\\tau^s_y', 'Shear traction'],
    'tsz': ['f~', '\\tau^s_z', 'Shear traction'],
    'tnm': ['f~', '\\tau^n', 'Normal traction'],
    'tsm': ['f~', '|\\tau^s|', 'Shear traction magnitude'],
    'fr': ['f~', '\\tau_c', 'Friction'],
    'sax': ['f~', '\\ddot s_x', 'Slip acceleration'],
    'say': ['f~', '\\ddot s_y', 'Slip acceleration'],
    'saz': ['f~', '\\ddot s_z', 'Slip acceleration'],
    'sam': ['f~', '|\\ddot s|', 'Slip acceleration magnitude'],
    'svx': ['f~', '\\dot s_x', 'Slip velocity'],
    'svy': ['f~', '\\dot s_y', 'Slip velocity'],
    'svz': ['f~', '\\dot s_z', 'Slip velocity'],
    'svm': ['f~', '|\\dot s|', 'Slip velocity magnitude'],
    'psv': ['f~', '|\\dot s|_{peak}', 'Peak slip velocity'],
    'sux': ['f~', 's_x', 'Slip'],
    'suy': ['f~', 's_y', 'Slip'],
    'suz': ['f~', 's_z', 'Slip'],
    'sum': ['f~', '|s|', 'Slip magnitude'],
    'sl': ['f~', '\\ell', 'Slip path length'],
    'trup': ['f~', 't_{rupture}', 'Rupture time'],
    'tarr': ['f~', 't_{arrest}', 'Arrest time'],
    '_psi': ['f~', '\\psi', 'State variable'],
}

boundary_conditions = {
    'free': 0,
    'rigid': 3,
    '+node': 1,
    '+cell': 2,
    '-node': -1,
    '-cell': -2,
    'pml': 10,
}

time_functions = [
    'const',
    'delta',
    'step', 'integral_delta',
    'brune',
    'integral_brune',
    'hann',
    'integral_hann',
    'gaussian', 'integral_ricker1',
    'ricker1', 'integral_ricker2',
    'ricker2',
]

_nonfile_ops = ['.', '=', '+', '*', '=~', '+~', '*~']
_cube_ops = ['=@', '+@', '*@']
_file_ops = ['.<', '=<', '+<', '*<', '.>', '=>']


class SordError(Exception):
    """A staging or solver error (bad parameters, bad i/o op, solver abort)."""


def expand_slices(shape, slices):
    """Normalise user indices to ``[start, stop, step]`` lists.

    Follows cst/sord.py:226-239 with the ``[]`` shorthand meaning the whole
    axis (README.md:166-170) and lists instead of tuples.
    """
    n = len(shape)
    if len(slices) == 0:
        slices = n * [[]]
    elif len(slices) != n:
        raise SordError('error in indices: %r' % (slices,))
    ss = []
    for n, s in zip(shape, slices):
        if isinstance(s, (list, tuple)):
            s = list(slice(*s).indices(n)) if len(s) else [0, n, 1]
        else:
            s = [s % n, s % n + 1, 1]
        ss.append(s)
    return ss


def prepare(prm, fio):
    """Derive the solver input records; cst/sord.py:297-468."""
    prm = copy.deepcopy(prm)
    if prm['source'] not in ('potency', 'moment', 'force', 'none'):
        raise SordError('bad source type: %r' % prm['source'])
    if prm['faultnormal'] not in ('none', '-x', '-y', '-z', '+x', '+y', '+z'):
        raise SordError('bad faultnormal: %r' % prm['faultnormal'])

    # intervals
    nx, ny, nz, nt = shape = prm['shape']
    prm['itio'] = max(1, min(prm['itio'], nt))

    # rupture
    ifn = None
    irup = None
    if prm['faultnormal'] != 'none':
        ifn = {'x': 0, 'y': 1, 'z': 2}[prm['faultnormal'][1:]]
        for i in 0, 1, 2:
            prm['hypocenter'][i] %= shape[i]
            if prm['hypocenter'][i] > shape[i] - 1:
                raise SordError('hypocenter outside the mesh')
        irup = int(prm['hypocenter'][ifn])
        if irup == 0:
            prm['bc1'][ifn] = '-cell'
        if irup == shape[ifn] - 2:
            prm['bc2'][ifn] = '-cell'

    # pml (D13: npml damped planes per PML face, none on other faces)
    npml = prm['npml']
    i1 = [-1, -1, -1]
    i2 = [nx + 1, ny + 1, nz + 1]
    if npml > 0:
        for i in 0, 1, 2:
            if prm['bc1'][i] == 'pml':
                i1[i] = npml - 1
            if prm['bc2'][i] == 'pml':
                i2[i] = shape[i] - npml + 1
            if i1[i] + 1 >= i2[i] - 1:
                raise SordError('Error: model too small for PML')
    prm.update({'i1pml': i1, 'i2pml': i2})

    # convert boundary conditions to numeric ids
    for k in 'bc1', 'bc2':
        for i in 0, 1, 2:
            if isinstance(prm[k][i], str):
                prm[k][i] = boundary_conditions[prm[k][i]]

    # field i/o
    fio_ = []
    filenames = []
    shapes = {}
    deltas = {}
    indices = {}
    for field in sorted(fio):
        if field not in fieldnames:
            raise SordError('unknown field: %r' % field)
        ios = fio[field]
        tags = fieldnames[field][0]
        reg = tags[0]
        input_ = '<' in tags
        static = '~' not in tags
        it1, it2 = nt + 1, -1
        if reg == 'c':
            fshape = [nx - 1, ny - 1, nz - 1]
        else:
            fshape = [nx, ny, nz]
        if reg == 'f':
            if ifn is None:
                raise SordError('Error: field only for ruptures: %r' % field)
            del fshape[ifn]
        if not static:
            fshape += [nt]
        if isinstance(ios, (float, int)):
            ios = [ios]
        elif len(ios) > 1 and isinstance(ios[1], str):
            ios = [ios]
        ios_ = []
        for io in ios:
            fname, val, tau = 'const', 0.0, 0.0
            x1, x2 = [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]
            if isinstance(io, (float, int)):
                slices, op, rhs = [], '=', [io]
            else:
                slices, op, rhs = io[0], io[1], list(io[2:])
            if op == '#':
                pass
            elif op in _nonfile_ops:
                if len(rhs) == 1:
                    val, = rhs
                else:
                    val, fname, tau = rhs
            elif op in _cube_ops:
                if len(rhs) == 3:
                    x1, x2, val = rhs
                else:
                    x1, x2, val, fname, tau = rhs
                x1 = list(x1)
                x2 = list(x2)
            elif op in _file_ops:
                fname, = rhs
                if fname in filenames:
                    raise SordError('Error: duplicate filename: ' + fname)
                filenames.append(fname)
                fname = os.path.expanduser(fname)
            else:
                raise SordError('Error: bad i/o operation: %r' % (io,))

            # error check
            if len(fname) > 32:
                raise SordError('File/function name too long: ' + fname)
            if op not in _file_ops and op != '#' and fname not in time_functions:
                raise SordError('invalid time func: %r' % fname)
            if '>' not in op and not input_:
                raise SordError('Error: field is ouput only: %r' % field)

            # slices
            slices = expand_slices(fshape, slices)
            if reg == 'f':
                slices.insert(ifn, [irup, irup + 1, 1])
            if static:
                slices.append([-1, 0, 1])
            if '.' in op:
                full = [nx - 1, ny - 1, nz - 1] if reg == 'c' else [nx, ny, nz]
                for i in 0, 1, 2:
                    x1[i] = float(slices[i][0])
                    if x1[i] > full[i] - 1:
                        raise SordError('sub-cell location outside the mesh')
                    slices[i][0] = int(slices[i][0])
                    slices[i][1] = int(slices[i][1])
            it1 = min(it1, slices[-1][0])
            it2 = max(it2, slices[-1][1])

            # buffer size
            nn = [(s[1] - s[0] - 1) // s[2] + 1 for s in slices]
            nb = (min(prm['itio'], nt) - 1) // slices[3][2] + 1
            nb = max(1, min(nb, nn[3]))
            n = nn[0] * nn[1] * nn[2]
            if n > (nx + ny + nz) ** 2:
                nb = 1
            elif n > 1:
                nb = min(nb, prm['itbuff'])

            ios_ += [[field, reg, slices, nb, x1, x2, val, tau, op, fname]]

            # metadata
            if op[-1] in '<>':
                ii = []
                nn = []
                dd = []
                for i, s in enumerate(slices):
                    start, stop, step = s
                    d = prm['delta'][i] * step
                    n = (stop - start - 1) // step + 1
                    if n == 1:
                        ii += [start]
                    else:
                        dd += [d]
                        nn += [n]
                        if step == 1:
                            ii += [[start, stop]]
                        else:
                            ii += [[start, stop, step]]
                if nn != []:
                    shapes[fname] = nn
                if dd != []:
                    deltas[fname] = dd
                if '.' in op:
                    indices[fname] = x1
                else:
                    indices[fname] = ii

        fio_ += [(it1, it2, ios_)]

    # sort by time window (cst/sord.py:460)
    fio_ = [g for f in sorted(fio_) for g in f[2]]
    del prm['itbuff']
    prm.update({'nfieldio': len(fio_)})
    meta = {
        'indices': indices,
        'deltas': deltas,
        'shapes': shapes,
    }
    return prm, fio_, meta


def split_args(args=None, **kwargs):
    """Split user arguments into parameters, field i/o and job keys."""
    args = dict(args or {})
    args.update(kwargs)
    prm = {}
    fio = {}
    job = {}
    for k, v in parameters.items():
        if k in fieldnames:
            fio[k] = copy.deepcopy(v)
        else:
            prm[k] = copy.deepcopy(v)
    for k, v in args.items():
        if k in fieldnames:
            fio[k] = copy.deepcopy(v)
        elif k in parameters:
            prm[k] = copy.deepcopy(v)
        else:
            job[k] = v
    return prm, fio, job


def sord_in_text(prm, fio):
    """The ``sord.in`` text: cst/sord.py:538-544.

    Each line ends in a blank before the newline: flang's list-directed
    reader (sord/input_output.f90:51,82) rejects a newline directly after a
    value inside one internal record, and a blank is a plain separator.
    """
    out = [prm[i] for i in sorted(prm)]
    out = json.dumps(out) + ' \n'
    for i in fio:
        out += json.dumps(i) + ' \n'
    for i in '",[]':
        out = out.replace(i, '')
    return out


def stage(args=None, realsize=4, write=True, **kwargs):
    """Prepare a run.  Returns ``(text, meta, prm, fio, job)``.

    With ``write`` the files cst/sord.py:535-553 writes are created in the
    current directory: ``sord.in``, ``parameters.json``, ``job.json``,
    ``sord.json``, ``meta.json``.
    """
    prm, fio, job = split_args(args, **kwargs)
    prm, fio_, meta = prepare(prm, fio)
    nproc = prm['nproc3'][0] * prm['nproc3'][1] * prm['nproc3'][2]
    if nproc > 1:
        raise SordError('one GPU per run: nproc3 must be [1, 1, 1]')
    text = sord_in_text(prm, fio_)
    d = {'little': '<', 'big': '>'}[sys.byteorder]
    meta['dtype'] = d + 'f%s' % realsize
    job['name'] = 'sord'
    if write:
        if prm['debug'] > 2 and not os.path.exists('debug'):
            os.mkdir('debug')
        with open('sord.in', 'w') as fh:
            fh.write(text)
        sprm = dict(prm)
        sprm['~fieldio'] = fio_
        allargs = dict(args or {})
        allargs.update(kwargs)
        for name, obj in (('parameters.json', allargs), ('job.json', job),
                          ('sord.json', sprm), ('meta.json', meta)):
            with open(name, 'w') as fh:
                json.dump(obj, fh, indent=4, sort_keys=True)
    return text, meta, prm, fio_, job


def run(args=None, realsize=4, device=0, **kwargs):
    """Stage and run SORD on the GPU; returns ``meta`` (cst/sord.py:558-561).

    Output files are written to the current directory, exactly as the
    reference solver writes them.
    """
    from . import solver
    text, meta, prm, fio_, job = stage(args, realsize=realsize, **kwargs)
    with solver.Solver(text, realsize=realsize, device=device) as s:
        s.run()
    return meta
