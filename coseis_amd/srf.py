"""
Standard Rupture Format (SRF) kinematic ruptures as SORD finite-source input.

`read` parses an SRF text (Graves' format, versions 1 and 2: an optional PLANE
header block, then a POINTS block of subfaults, each with up to three slip
components and their slip-rate histories) into metadata and numpy arrays in
SI units, as cst/srf.py:13-117 does.  `write_sord` writes the files the
solver's finite source reads (source-*.bin, sord/kinematic_source.f90:10-63),
the semantics of cst/srf.py:197-300: every (subfault, slip component) with a
history is one point source whose potency tensor is the subfault area times
the symmetric dyad of the fault normal and that component's slip direction,
with the cumulative slip as its time history.

Where the shipped writer departs from its own intent, the intent is kept:
its W22 file is opened under the W23 name and then overwritten (cst/srf.py:
259-261), and the history loops of the second and third components run over
the last subfault's sample count instead of the subfaults (:283-297).
Integer files (nt) are written as 4-byte integers, the solver's integer size,
at either real precision.
"""
import math

import numpy as np

KM, CM, CM2 = 1000.0, 0.01, 0.0001  # SRF units in metres, square metres


def _fields(fh):
    """Whitespace-separated tokens of the remaining lines, one line at a time."""
    for line in fh:
        t = line.split()
        if t:
            yield t


def read(fh):
    """Parse an SRF file object; returns (meta, data).

    meta: version, plane (list of segment dicts: shape, length, area, delta,
    topcenter, strike, dip, hypocenter), nsource, nsource_nonzero,
    hypocenter (lon, lat, depth of the earliest subfault), area, potency,
    displacement.  data: per subfault lon, lat, dep (m), stk, dip, rake
    (degrees), area (m^2), t0, dt (s), slip1..3 (m), nt1..3, and the
    concatenated slip-rate samples sv1..3 (m/s).  With one plane the
    per-subfault arrays are shaped (along strike, down dip)."""
    lines = _fields(fh)
    meta = {'version': next(lines)[0]}
    head = next(lines)
    meta['plane'] = []
    if head[0] == 'PLANE':
        for _ in range(int(head[1])):
            t = next(lines) + next(lines)
            if len(t) != 11:
                raise ValueError('SRF: bad PLANE segment record')
            nx, nz = int(t[2]), int(t[3])
            length = [float(t[4]) * KM, float(t[5]) * KM]
            meta['plane'].append({
                'shape': [nx, nz], 'length': length, 'area': length[0] * length[1],
                'delta': [length[0] / nx, length[1] / nz],
                'topcenter': [float(t[0]), float(t[1]), float(t[8])],
                'strike': float(t[6]), 'dip': float(t[7]),
                'hypocenter': [float(t[9]) * KM, float(t[10]) * KM],
            })
        head = next(lines)
    if head[0] != 'POINTS':
        raise ValueError('SRF: POINTS block expected')
    n = meta['nsource'] = int(head[1])
    floats = ('lon', 'lat', 'dep', 'stk', 'dip', 'rake', 'area', 't0', 'dt', 'slip1', 'slip2', 'slip3')
    data = {k: np.empty(n, 'f') for k in floats}
    data.update({k: np.empty(n, 'i') for k in ('nt1', 'nt2', 'nt3')})
    rates = ([], [], [])
    for i in range(n):
        t = next(lines) + next(lines)
        if len(t) != 15:
            raise ValueError('SRF: bad record of subfault %d' % i)
        lon, lat, dep, stk, dip, area, t0, dt, rake = (float(x) for x in t[:9])
        data['lon'][i], data['lat'][i], data['dep'][i] = lon, lat, dep * KM
        data['stk'][i], data['dip'][i], data['rake'][i] = stk, dip, rake
        data['area'][i], data['t0'][i], data['dt'][i] = area * CM2, t0, dt
        counts = []
        for c in range(3):
            data['slip%d' % (c + 1)][i] = float(t[9 + 2 * c]) * CM
            counts.append(int(t[10 + 2 * c]))
            data['nt%d' % (c + 1)][i] = counts[-1]
        samples = []
        while len(samples) < sum(counts):
            samples += next(lines)
        if len(samples) != sum(counts):
            raise ValueError('SRF: slip-rate samples of subfault %d' % i)
        start = 0
        for c in range(3):
            rates[c].extend(float(x) * CM for x in samples[start:start + counts[c]])
            start += counts[c]
    for c in range(3):
        data['sv%d' % (c + 1)] = np.array(rates[c], 'f')
    if len(meta['plane']) == 1:
        shape = meta['plane'][0]['shape']
        for k in floats + ('nt1', 'nt2', 'nt3'):
            data[k] = data[k].reshape(shape[::-1]).T
    meta['nsource_nonzero'] = int(sum((data['nt%d' % c] > 0).sum() for c in (1, 2, 3)))
    first = int(np.argmin(data['t0']))
    meta['hypocenter'] = [float(data[k].flat[first]) for k in ('lon', 'lat', 'dep')]
    meta['area'] = float(data['area'].sum(dtype='d'))
    comps = [(data['area'] * data['slip%d' % c]).sum(dtype='d') for c in (1, 2, 3)]
    meta['potency'] = float(math.sqrt(sum(p * p for p in comps)))
    meta['displacement'] = meta['potency'] / meta['area'] if meta['area'] else 0.0
    return meta, data


def fault_frame(strike, dip, rake):
    """Unit vectors (slip direction 1, slip direction 2, fault normal) in
    (east, north, up) for strike, dip, rake in degrees, the Aki-Richards
    convention (dip to the right of the strike direction): the rows of the
    Z-X-Z Euler rotation by (90 - strike, dip, rake), cst/coord.py:164-202.
    Arrays broadcast; returns three arrays of shape (3,) + strike.shape."""
    a, b, c = np.broadcast_arrays(*(np.radians(np.asarray(x, 'd')) for x in (90.0 - np.asarray(strike, 'd'), dip,
                                                                            rake)))
    ca, sa, cb, sb, cc, sc = np.cos(a), np.sin(a), np.cos(b), np.sin(b), np.cos(c), np.sin(c)
    # R = Rz(rake) Rx(dip) Rz(90 - strike), Rz(t) = [[c, s, 0], [-s, c, 0], [0, 0, 1]],
    # Rx(t) = [[1, 0, 0], [0, c, s], [0, -s, c]]
    r1 = np.array([cc * ca - sc * cb * sa, cc * sa + sc * cb * ca, sc * sb])
    r2 = np.array([-sc * ca - cc * cb * sa, -sc * sa + cc * cb * ca, cc * sb])
    r3 = np.array([sb * sa, -sb * ca, cb])
    return r1, r2, r3


def potency_dyad(normal, slip):
    """Symmetric dyad of a fault normal and a slip direction as the solver's
    six potency components (W11, W22, W33, W23, W31, W12)."""
    n, s = normal, slip
    return (n[0] * s[0], n[1] * s[1], n[2] * s[2], 0.5 * (n[1] * s[2] + n[2] * s[1]),
            0.5 * (n[2] * s[0] + n[0] * s[2]), 0.5 * (n[0] * s[1] + n[1] * s[0]))


def local_rotation(lon, lat, projection, eps=100.0):
    """Clockwise angle (degrees) from geographic north to the projected grid's
    y axis at (lon, lat): the mean of the projected east and north directions'
    deviations, from points eps metres away (cst/coord.py:205-230)."""
    rearth = 6370000.0
    dlon = eps * 180.0 / (math.pi * rearth) * np.cos(np.radians(lat))
    dlat = eps * 180.0 / (math.pi * rearth)
    xe, ye = projection(np.array([lon - dlon, lon + dlon]), np.array([lat, lat]))
    xn, yn = projection(np.array([lon, lon]), np.array([lat - dlat, lat + dlat]))
    east = math.degrees(1) * np.arctan2(xe[1] - xe[0], ye[1] - ye[0]) - 90.0
    north = math.degrees(1) * np.arctan2(xn[1] - xn[0], yn[1] - yn[0])
    return 0.5 * (east + north)


def write_sord(path, srf, delta=(1.0, 1.0, 1.0), proj=None, dbytes=4):
    """Write the solver's finite-source files path + 'xi1.bin' ... (with path
    'source-' they are the files sord/kinematic_source.f90 reads): one point
    source per (subfault, slip component) with samples, grouped by component
    (all first components, then the second, then the normal one).

    xi1..3: position in grid units (projected x, y or lon, lat, and depth over
    delta); t0, dt, nt: start, sampling interval, sample count; w11..w12:
    area times the normal-slip dyad; history: cumulative slip (m) at each
    sample.  proj(lon, lat) -> (x, y) projects to the mesh, whose rotation
    against north is added to the strike.  Returns the number of point
    sources (the solver's nsource)."""
    meta, data = srf
    real = '<f%d' % dbytes
    flat = {k: np.asarray(v).ravel(order='F') for k, v in data.items() if not k.startswith('sv')}
    x, y, z = flat['lon'].astype('d'), flat['lat'].astype('d'), flat['dep'].astype('d')
    strike = flat['stk'].astype('d')
    if proj is not None:
        strike = strike + local_rotation(x, y, proj)
        x, y = proj(x, y)
    xi = (np.asarray(x, 'd') / delta[0], np.asarray(y, 'd') / delta[1], z / delta[2])
    frame = fault_frame(strike, flat['dip'], flat['rake'])
    normal = frame[2]
    area = flat['area'].astype('d')
    # arrays in the SRF's record order (a single plane's arrays are shaped
    # (along strike, down dip) with the strike index fastest in the records),
    # which is the order of the concatenated slip-rate samples
    out = {k: [] for k in ('xi1', 'xi2', 'xi3', 't0', 'dt', 'nt', 'w11', 'w22', 'w33', 'w23', 'w31', 'w12')}
    history = []
    for c in range(3):
        rate = np.asarray(data['sv%d' % (c + 1)], 'd')
        ntc = flat['nt%d' % (c + 1)].astype(int)
        starts = np.concatenate([[0], np.cumsum(ntc)])
        for i in np.flatnonzero(ntc > 0):
            for a in range(3):
                out['xi%d' % (a + 1)].append(xi[a][i])
            out['t0'].append(flat['t0'][i])
            out['dt'].append(flat['dt'][i])
            out['nt'].append(ntc[i])
            w = potency_dyad(normal[:, i], frame[c][:, i])
            for k, v in zip(('w11', 'w22', 'w33', 'w23', 'w31', 'w12'), w):
                out[k].append(area[i] * v)
            history.append(np.cumsum(rate[starts[i]:starts[i + 1]]) * float(flat['dt'][i]))
    for k, v in out.items():
        np.asarray(v, '<i4' if k == 'nt' else real).tofile(path + k + '.bin')
    np.concatenate(history or [np.zeros(0)]).astype(real).tofile(path + 'history.bin')
    return len(out['nt'])
