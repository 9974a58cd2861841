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

# field-I/O operation kinds: value ops (optionally times a time function),
# cube ops (a value inside a coordinate box), file ops (read '<', write '>')
VALUE_OPS = ('.', '=', '+', '*', '=~', '+~', '*~')
CUBE_OPS = ('=@', '+@', '*@')
FILE_OPS = ('.<', '=<', '+<', '*<', '.>', '=>')
AXES = {'x': 0, 'y': 1, 'z': 2}


class SordError(Exception):
    """A staging or solver error (bad parameters, bad i/o op, solver abort)."""


def _axis_range(n, s):
    """One axis of a user index as [start, stop, step] over n points: a list
    is Python slice notation ([] is the whole axis, README.md:166-170), a
    number is a single (possibly negative) index."""
    if not isinstance(s, (list, tuple)):
        return [s % n, s % n + 1, 1]
    if len(s) == 0:
        return [0, n, 1]
    return list(slice(*s).indices(n))


def expand_slices(shape, slices):
    """User indices of a field (one entry per axis of `shape`, or none for the
    whole field) as [start, stop, step] lists (cst/sord.py:226-239)."""
    if len(slices) == 0:
        slices = [[]] * len(shape)
    if len(slices) != len(shape):
        raise SordError('error in indices: %r' % (slices,))
    return [_axis_range(n, s) for n, s in zip(shape, slices)]


def _count(r):
    """Number of indices in a [start, stop, step] range."""
    start, stop, step = r
    return (stop - start - 1) // step + 1


def _fault_plane(prm, nn):
    """Fault-normal axis and the index of the fault's first node plane (None,
    None without a fault).  The hypocentre wraps like a negative index; a
    fault on the first or last cell plane of its axis makes that face a
    cell-centred mirror (cst/sord.py:306-317)."""
    if prm['faultnormal'] == 'none':
        return None, None
    ax = AXES[prm['faultnormal'][1]]
    hyp = prm['hypocenter']
    for i in range(3):
        hyp[i] %= nn[i]
        if hyp[i] > nn[i] - 1:
            raise SordError('hypocenter outside the mesh')
    irup = int(hyp[ax])
    if irup == 0:
        prm['bc1'][ax] = '-cell'
    if irup == nn[ax] - 2:
        prm['bc2'][ax] = '-cell'
    return ax, irup


def _pml_box(prm, nn):
    """Last damped node on the low side and first on the high side, per
    axis, as the solver reads them (it adds 1): exactly npml damped node and
    cell planes per PML face, none on other faces (-1 and n + 1).  The
    shipped stager's box (cst/sord.py:319-330) damps npml + 1 planes against
    npml table entries (defect D13)."""
    npml = prm['npml']
    lo = [-1, -1, -1]
    hi = [n + 1 for n in nn]
    if npml <= 0:
        return lo, hi
    for i in range(3):
        if prm['bc1'][i] == 'pml':
            lo[i] = npml - 1
        if prm['bc2'][i] == 'pml':
            hi[i] = nn[i] - npml + 1
        if lo[i] + 1 >= hi[i] - 1:
            raise SordError('Error: model too small for PML')
    return lo, hi


def _op_list(value):
    """The operations given for one field: a number, one operation tuple, or
    a list of numbers and tuples."""
    if isinstance(value, (float, int)):
        return [value]
    if len(value) > 1 and isinstance(value[1], str):
        return [value]
    return list(value)


class _Op(object):
    """One field-I/O operation decoded from its tuple."""

    def __init__(self, item, files):
        self.fname, self.val, self.tau = 'const', 0.0, 0.0
        self.x1, self.x2 = [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]
        if isinstance(item, (float, int)):
            self.index, self.op, args = [], '=', [item]
        else:
            self.index, self.op, args = item[0], item[1], list(item[2:])
        op = self.op
        if op == '#':
            pass
        elif op in VALUE_OPS:
            if len(args) == 1:
                self.val = args[0]
            else:
                self.val, self.fname, self.tau = args
        elif op in CUBE_OPS:
            if len(args) == 3:
                x1, x2, self.val = args
            else:
                x1, x2, self.val, self.fname, self.tau = args
            self.x1, self.x2 = list(x1), list(x2)
        elif op in FILE_OPS:
            self.fname, = args
            if self.fname in files:
                raise SordError('Error: duplicate filename: ' + self.fname)
            files.add(self.fname)
            self.fname = os.path.expanduser(self.fname)
        else:
            raise SordError('Error: bad i/o operation: %r' % (item,))
        if len(self.fname) > 32:
            raise SordError('File/function name too long: ' + self.fname)
        if op not in FILE_OPS and op != '#' and self.fname not in time_functions:
            raise SordError('invalid time func: %r' % (self.fname,))


class _Stager(object):
    """The solver's field-I/O records and the output metadata of one run."""

    def __init__(self, prm, fault_axis, irup):
        self.prm = prm
        self.nn = prm['shape'][:3]
        self.nt = prm['shape'][3]
        self.fault_axis, self.irup = fault_axis, irup
        self.files = set()
        self.meta = {'indices': {}, 'deltas': {}, 'shapes': {}}

    def extent(self, reg, static):
        """Index extent of a field: nodes or cells, without the fault-normal
        axis on the fault plane, plus time for time-varying fields."""
        ext = [n - 1 for n in self.nn] if reg == 'c' else list(self.nn)
        if reg == 'f':
            del ext[self.fault_axis]
        return ext if static else ext + [self.nt]

    def buffer_depth(self, ranges):
        """Time records buffered per file write (performance only): up to
        itio steps, 1 for a slice larger than (nx + ny + nz)^2 points, at
        most itbuff for other multi-point slices."""
        nx, ny, nz = self.nn
        nb = max(1, min((min(self.prm['itio'], self.nt) - 1) // ranges[3][2] + 1, _count(ranges[3])))
        points = _count(ranges[0]) * _count(ranges[1]) * _count(ranges[2])
        if points > (nx + ny + nz) ** 2:
            return 1
        return min(nb, self.prm['itbuff']) if points > 1 else nb

    def describe(self, o, ranges):
        """meta.json entries of a file op: the file's shape (axes of more than
        one point), their spacing and the indices it covers."""
        index, shape, delta = [], [], []
        for i, (start, stop, step) in enumerate(ranges):
            n = _count([start, stop, step])
            if n == 1:
                index.append(start)
                continue
            shape.append(n)
            delta.append(self.prm['delta'][i] * step)
            index.append([start, stop] if step == 1 else [start, stop, step])
        if shape:
            self.meta['shapes'][o.fname] = shape
        if delta:
            self.meta['deltas'][o.fname] = delta
        self.meta['indices'][o.fname] = o.x1 if '.' in o.op else index

    def field(self, name, value):
        """Records of one field and their time window (first, last step)."""
        if name not in fieldnames:
            raise SordError('unknown field: %r' % name)
        tags = fieldnames[name][0]
        reg, settable, static = tags[0], '<' in tags, '~' not in tags
        if reg == 'f' and self.fault_axis is None:
            raise SordError('Error: field only for ruptures: %r' % name)
        ext = self.extent(reg, static)
        recs, first, last = [], self.nt + 1, -1
        for item in _op_list(value):
            o = _Op(item, self.files)
            if '>' not in o.op and not settable:
                raise SordError('Error: field is ouput only: %r' % name)
            ranges = expand_slices(ext, o.index)
            if reg == 'f':
                ranges.insert(self.fault_axis, [self.irup, self.irup + 1, 1])
            if static:
                ranges.append([-1, 0, 1])  # applied at initialisation (it = 0)
            if '.' in o.op:
                # a sub-cell point: x1 keeps the fractional node (or cell)
                # position, the range the cell that holds it
                top = [n - 1 for n in self.nn] if reg == 'c' else list(self.nn)
                for i in range(3):
                    o.x1[i] = float(ranges[i][0])
                    if o.x1[i] > top[i] - 1:
                        raise SordError('sub-cell location outside the mesh')
                    ranges[i][0], ranges[i][1] = int(ranges[i][0]), int(ranges[i][1])
            first, last = min(first, ranges[3][0]), max(last, ranges[3][1])
            recs.append([name, reg, ranges, self.buffer_depth(ranges), o.x1, o.x2, o.val, o.tau, o.op, o.fname])
            if o.op[-1] in '<>':
                self.describe(o, ranges)
        return first, last, recs


def prepare(prm, fio):
    """Solver parameters (line 1 of sord.in), field-I/O records (one line
    each) and output metadata from the user's parameters and field I/O;
    the semantics of cst/sord.py:297-468 with defects D3-D5 and D13 repaired.

    Records are grouped per field and the groups ordered by time window
    (first step, last step, then the records themselves), which is the order
    the solver scans them in (sord/input_output.f90:103-151)."""
    prm = copy.deepcopy(prm)
    if prm['source'] not in ('potency', 'moment', 'force', 'none'):
        raise SordError('bad source type: %r' % prm['source'])
    if prm['faultnormal'] not in ('none', '-x', '-y', '-z', '+x', '+y', '+z'):
        raise SordError('bad faultnormal: %r' % prm['faultnormal'])
    nn, nt = prm['shape'][:3], prm['shape'][3]
    prm['itio'] = max(1, min(prm['itio'], nt))
    ax, irup = _fault_plane(prm, nn)
    prm['i1pml'], prm['i2pml'] = _pml_box(prm, nn)
    for key in ('bc1', 'bc2'):
        prm[key] = [boundary_conditions[b] if isinstance(b, str) else b for b in prm[key]]
    st = _Stager(prm, ax, irup)
    groups = sorted(st.field(name, fio[name]) for name in sorted(fio))
    records = [r for g in groups for r in g[2]]
    del prm['itbuff']
    prm['nfieldio'] = len(records)
    return prm, records, st.meta


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
