"""Workload generators (coseis_amd.configs) and the autotuner's buckets, on
the CPU: the properties their definitions fix."""
import numpy as np

from coseis_amd import autotune, configs


def test_gaussian_field_moments_and_correlation():
    """Zero mean, unit std, and the correlation exp(-r^2 / corr^2) at lags of
    0.5, 1 and 1.5 correlation lengths along each axis (a wrong filter
    exponent moves the lag-corr value from 0.37 to 0.14 or 0.61)."""
    corr = 6.0
    f = configs.gaussian_field((96, 80, 72), corr, np.random.default_rng(3))
    assert abs(f.mean()) < 1e-12 and abs(f.std() - 1) < 1e-12
    for axis in range(3):
        for lag in (3, 6, 9):
            a = np.moveaxis(f, axis, 0)
            c = float((a[:-lag] * a[lag:]).mean())
            assert abs(c - np.exp(-(lag / corr) ** 2)) < 0.08, (axis, lag, c)


def test_adjoint_workload():
    n = 40
    w = configs.adjoint_200(n=n, nt=10, nsrc=3, nrec=16)
    prm = w['prm']
    assert prm['shape'] == [n, n, n, 10]
    assert prm['bc1'] == ['free'] * 3 and prm['bc2'] == ['free'] * 3
    for model in (w['model'], w['observed']):
        for a, bg in zip(model, (2700.0, 6000.0, 3464.0)):
            assert a.shape == (n - 1,) * 3
            rel = a / bg - 1
            assert abs(rel.std() - 0.05) < 1e-3 and abs(rel.mean()) < 1e-3
    assert not np.allclose(w['model'][0], w['observed'][0])
    # vp / vs is the background ratio everywhere (one field scales all three)
    assert np.allclose(w['model'][1] / w['model'][2], 6000.0 / 3464.0, rtol=1e-5)
    rec = w['receivers']
    assert rec.shape == (16, 3) and (rec[:, 2] == 0).all()
    assert (rec[:, :2] >= 10).all() and (rec[:, :2] < n - 10).all()
    for f in ('mxx', 'myy', 'mzz', 'myz', 'mzx', 'mxy'):
        assert len(prm[f]) == 3
        for x, op, val, fn, tau in prm[f]:
            assert op == '.' and fn == 'brune' and tau == 0.1
            assert all(v % 1 == 0.5 for v in x[:3])  # cell centres
    # each double couple: symmetric moment with zero trace and M0 = 1e16
    for i in range(3):
        m = np.zeros((3, 3))
        for f, (a, b) in {'mxx': (0, 0), 'myy': (1, 1), 'mzz': (2, 2), 'myz': (1, 2), 'mzx': (2, 0),
                          'mxy': (0, 1)}.items():
            m[a, b] = m[b, a] = prm[f][i][2]
        assert abs(np.trace(m)) < 1e4
        assert abs(np.sqrt((m * m).sum() / 2) - 1e16) < 1e10


def test_ensemble_workload():
    ens = configs.ensemble(n=24, seed=5, nt=20)
    assert len(ens) == 24
    shapes = [p['shape'] for p, _ in ens]
    assert all(48 <= e <= 160 for s in shapes for e in s[:3])
    assert len({tuple(s) for s in shapes}) == 24
    for prm, files in ens:
        assert prm['bc1'] == ['pml'] * 3 and prm['bc2'] == ['pml'] * 3 and prm['npml'] == 10
        vs = np.array([prm['vs'][0]] + [z[-1] for z in prm['vs'][1:]])
        vp = np.array([prm['vp'][0]] + [z[-1] for z in prm['vp'][1:]])
        rho = np.array([prm['rho'][0]] + [z[-1] for z in prm['rho'][1:]])
        assert ((vs >= 1500) & (vs <= 3500)).all()
        assert np.allclose(vp, 1.732 * vs) and np.allclose(rho, 310 * vp ** 0.25)
        m = np.zeros((3, 3))
        for f, (a, b) in {'mxx': (0, 0), 'myy': (1, 1), 'mzz': (2, 2), 'myz': (1, 2), 'mzx': (2, 0),
                          'mxy': (0, 1)}.items():
            m[a, b] = m[b, a] = prm[f][0][2]
        m0 = np.sqrt((m * m).sum() / 2)
        assert 1e15 * 0.999 <= m0 <= 1e18 * 1.001


def test_autotune_buckets():
    assert autotune.bucket([48, 60, 71, 10]) == 0
    assert autotune.bucket([48, 72, 50, 10]) == 1
    assert autotune.bucket([107, 60, 50, 10]) == 1
    assert autotune.bucket([160, 60, 50, 10]) == 2
