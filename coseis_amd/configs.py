"""
Benchmark workloads as cst.sord parameter dictionaries.

Every function returns (prm, files): the parameters and a dict of input files
{name: array} to write before the run.  Indices are 0-based node / cell
indices of the cst.sord interface.
"""
import numpy as np


def explosion_61(nt=200):
    """Explosion in a homogeneous full space: 61^3 nodes, dx = 100 m,
    dt = 0.0075 s, 10-node PML on all faces, isotropic moment M0 = 1e16 N m
    spread over the 8 cells around the centre node, Brune pulse tau = 0.1 s."""
    prm = {
        'shape': [61, 61, 61, nt], 'delta': [100.0, 100.0, 100.0, 0.0075],
        'bc1': ['pml'] * 3, 'bc2': ['pml'] * 3, 'npml': 10,
        'rho': [2670.0], 'vp': [6000.0], 'vs': [3464.0], 'gam': [0.0],
        'vx': [([[], [], [], -1], '=>', 'vx.bin')],
        'vy': [([[], [], [], -1], '=>', 'vy.bin')],
        'vz': [([[], [], [], -1], '=>', 'vz.bin')],
    }
    for f in 'mxx', 'myy', 'mzz':
        prm[f] = [([30.0, 30.0, 30.0, []], '.', 1e16, 'brune', 0.1)]
    for f in 'vx', 'vy', 'vz':
        prm[f].append(([40.0, 30.0, 30.0, []], '.>', 'p1-%s.bin' % f))
    return prm, {}


def layered_321(nt=1500, receivers=True):
    """A 1 km layer (rho 2600, vp 4000, vs 2000) over a half-space (rho 2700,
    vp 6000, vs 3464): 321 x 321 x 161 nodes, dx = 50 m, dt = 0.004 s, free
    surface on top, 10-node PML on the other faces, viscous hourglass on
    (hourglass = [1, 2]); strike-slip double couple M0 = 1e18 N m at 2 km
    depth, Brune tau = 0.1 s; velocity on a 16 x 16 surface receiver grid."""
    layer = ([[], [], [0, 20]], '=')
    prm = {
        'shape': [321, 321, 161, nt], 'delta': [50.0, 50.0, 50.0, 0.004],
        'bc1': ['pml', 'pml', 'free'], 'bc2': ['pml', 'pml', 'pml'], 'npml': 10,
        'rho': [2700.0, layer + (2600.0,)],
        'vp': [6000.0, layer + (4000.0,)],
        'vs': [3464.0, layer + (2000.0,)],
        'gam': [0.0],
        'hourglass': [1.0, 2.0],
        'mxy': [([160.0, 160.0, 40.0, []], '.', 1e18, 'brune', 0.1)],
    }
    if receivers:
        for f in 'vx', 'vy', 'vz':
            prm[f] = [([[8, 321, 20], [8, 321, 20], 0, []], '=>', 'rcv-%s.bin' % f)]
    return prm, {}


def rupture_400(nt=1500):
    """Spontaneous rupture on a vertical planar fault (double nodes, normal
    +y): 400 x 200 x 200 nodes, dx = 100 m, dt = 0.008 s, homogeneous
    (2670, 6000, 3464), 10-node PML; 30 x 15 km fault with mu_s = 0.677,
    mu_d = 0.525, dc = 0.4 m, tn = -120 MPa, ts = 70 MPa, a 3 x 3 km
    nucleation square at 81.6 MPa, mu_s = 1e4 outside the fault.  Fault
    parameters are read from files (the reference executes non-disk ops on
    fault fields outside the fault arrays).  Outputs: slip, slip rate,
    traction and rupture time at the last step."""
    nx, ny, nz = 400, 200, 200
    prm = {
        'shape': [nx, ny, nz, nt], 'delta': [100.0, 100.0, 100.0, 0.008],
        'bc1': ['pml'] * 3, 'bc2': ['pml'] * 3, 'npml': 10,
        'rho': [2670.0], 'vp': [6000.0], 'vs': [3464.0], 'gam': [0.0],
        'faultnormal': '+y', 'hypocenter': [199.5, 99.0, 99.5], 'vrup': -1.0,
    }
    shp = (nx, nz)
    mus = np.full(shp, 1e4, 'f4')
    mus[50:350, 25:175] = 0.677
    ts = np.full(shp, 70e6, 'f4')
    ts[185:215, 85:115] = 81.6e6
    files = {
        'in-mus.bin': mus, 'in-mud.bin': np.full(shp, 0.525, 'f4'),
        'in-dc.bin': np.full(shp, 0.4, 'f4'), 'in-tn.bin': np.full(shp, -120e6, 'f4'),
        'in-ts.bin': ts,
    }
    for f in ['mus', 'mud', 'dc', 'tn', 'ts']:
        prm[f] = [([[], []], '=<', 'in-%s.bin' % f)]
    for f in ['sux', 'suy', 'suz', 'svm', 'tsm', 'trup']:
        prm[f] = [([[], [], -1], '=>', f + '.bin')]
    return prm, files


def ensemble(n=64, seed=0, nt=800):
    """Batched ensemble: n scenarios with seeded uneven meshes.  Edges per
    axis drawn log-uniformly from 48 to 160 nodes, dx = 100 m, dt = 0.008 s,
    10-node PML on all faces.  Each model is a stack of 1 to 4 layers with vs
    uniform in [1500, 3500] m/s (increasing with depth), vp = 1.732 vs and
    Gardner density rho = 310 vp^0.25; one double couple (random fault normal
    and slip direction) with M0 log-uniform in [1e15, 1e18] N m and a Brune
    pulse, tau uniform in [0.1, 0.3] s, at a random node at least 4 nodes
    inside the PML.  Output: vz of the top interior plane at the last step.
    Returns a list of (prm, files)."""
    rng = np.random.default_rng(seed)
    out = []
    for q in range(n):
        shape = [int(round(float(np.exp(rng.uniform(np.log(48), np.log(160)))))) for _ in range(3)]
        nz = shape[2]
        nl = int(rng.integers(1, 5))
        vs = np.sort(rng.uniform(1500.0, 3500.0, nl))
        cuts = np.sort(rng.choice(np.arange(1, nz - 2), size=nl - 1, replace=False)) if nl > 1 else []
        vp = 1.732 * vs
        rho = 310.0 * vp ** 0.25
        # layer 0 on top (l = 0), the deepest layer fills the rest
        prm = {
            'shape': shape + [nt], 'delta': [100.0, 100.0, 100.0, 0.008],
            'bc1': ['pml'] * 3, 'bc2': ['pml'] * 3, 'npml': 10,
            'rho': [float(rho[-1])], 'vp': [float(vp[-1])], 'vs': [float(vs[-1])], 'gam': [0.0],
        }
        top = 0
        for i, c in enumerate(cuts):
            zone = ([[], [], [int(top), int(c)]], '=')
            prm['rho'].append(zone + (float(rho[i]),))
            prm['vp'].append(zone + (float(vp[i]),))
            prm['vs'].append(zone + (float(vs[i]),))
            top = c
        m0 = 10.0 ** rng.uniform(15.0, 18.0)
        nrm = rng.normal(size=3)
        nrm /= np.linalg.norm(nrm)
        slip = np.cross(nrm, rng.normal(size=3))
        slip /= np.linalg.norm(slip)
        x = [float(rng.integers(14, e - 14)) for e in shape]
        tau = float(rng.uniform(0.1, 0.3))
        comps = {'mxx': (0, 0), 'myy': (1, 1), 'mzz': (2, 2), 'myz': (1, 2), 'mzx': (2, 0), 'mxy': (0, 1)}
        for f, (a, b) in comps.items():
            prm[f] = [(x + [[]], '.', float(m0 * (slip[a] * nrm[b] + slip[b] * nrm[a])), 'brune', tau)]
        prm['vz'] = [([[], [], 10, -1], '=>', 'vz-top.bin')]
        out.append((prm, {}))
    return out


def gaussian_field(shape, corr, rng):
    """Zero-mean, unit-std random field with Gaussian correlation
    exp(-r^2 / corr^2) (corr in cells): white noise filtered in the Fourier
    domain by the square root of that correlation's spectrum."""
    w = rng.standard_normal(shape)
    k = [np.fft.fftfreq(n) * 2 * np.pi for n in shape[:-1]] + [np.fft.rfftfreq(shape[-1]) * 2 * np.pi]
    k2 = k[0][:, None, None] ** 2 + k[1][None, :, None] ** 2 + k[2][None, None, :] ** 2
    f = np.fft.irfftn(np.fft.rfftn(w) * np.exp(-k2 * corr * corr / 8.0), s=shape, axes=(0, 1, 2))
    f -= f.mean()
    return f / f.std()


def adjoint_200(n=200, nt=1500, seed=0, nrec=64, nsrc=8):
    """Adjoint-gradient workload: n^3 nodes, dx = 50 m, dt = 0.0035 s, free
    surface on every face.  Material (per cell): the homogeneous background
    (rho 2700, vp 6000, vs 3464) times (1 + 0.05 g), g a unit-std Gaussian
    random field with a 1 km correlation length (20 cells); the observed data
    come from a second field of the next seed.  Sources: nsrc double couples
    (random fault normal and slip, M0 = 1e16 N m, Brune tau = 0.1 s) at the
    centres of random cells at least 10 cells inside (one cell each); receivers: nrec random nodes of the top
    face (l = 0) at least 10 nodes inside.  Returns a dict: prm (without
    material), model and observed (rho, vp, vs) cell arrays, receivers (nrec,
    3) 0-based node indices."""
    rng = np.random.default_rng(seed)
    cells = (n - 1,) * 3
    bg = (2700.0, 6000.0, 3464.0)
    models = []
    for s in (seed, seed + 1):
        g = 1.0 + 0.05 * gaussian_field(cells, 20.0, np.random.default_rng(1000 + s))
        models.append(tuple(np.float32(b) * g.astype('f4') for b in bg))
    prm = {
        'shape': [n, n, n, nt], 'delta': [50.0, 50.0, 50.0, 0.0035],
        'bc1': ['free'] * 3, 'bc2': ['free'] * 3, 'gam': [0.0],
    }
    comps = {'mxx': (0, 0), 'myy': (1, 1), 'mzz': (2, 2), 'myz': (1, 2), 'mzx': (2, 0), 'mxy': (0, 1)}
    for f in comps:
        prm[f] = []
    for _ in range(nsrc):
        nrm = rng.normal(size=3)
        nrm /= np.linalg.norm(nrm)
        slip = np.cross(nrm, rng.normal(size=3))
        slip /= np.linalg.norm(slip)
        x = [float(rng.integers(10, n - 11)) + 0.5 for _ in range(3)]
        for f, (a, b) in comps.items():
            prm[f].append((x + [[]], '.', float(1e16 * (slip[a] * nrm[b] + slip[b] * nrm[a])), 'brune', 0.1))
    rec = np.zeros((nrec, 3), np.int32)
    rec[:, :2] = rng.integers(10, n - 10, size=(nrec, 2))
    return {'prm': prm, 'model': models[0], 'observed': models[1], 'receivers': rec}


def write_inputs(files, realsize=4):
    """Input files: reals in the run's precision, integer arrays as int32."""
    for name, a in files.items():
        a = np.asarray(a)
        a.astype('<i4' if np.issubdtype(a.dtype, np.integer) else '<f%d' % realsize).T.tofile(name)
