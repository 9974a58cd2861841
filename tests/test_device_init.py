"""Device-side initialisation against the host's (sord_set_device_init).

sord_init computes the material model, the node masses and viscosity and
the rectilinear cell volumes on the device by default (coseis_amd/csrc/
sord_init.hip), from the same input arrays as the host restatement
(sord_init.inc), which is pinned against the reference through the golden
fixtures.  On every case, at both precisions: the material arrays after
sord_init (lambda, mu, hourglass stiffness, 1/node mass, viscosity, cell
volume) and material.json are identical, and so is every output file of the
run, bit for bit."""
import numpy as np
import pytest

from conftest import run_case
import cases

pytestmark = pytest.mark.gpu
FIELDS = ('lam', 'mu', 'yy', 'mr', 'gam', 'vc')


def device_array(s, name):
    """The whole padded device array, copied to the host with the HIP
    runtime the library runs on (hipMemcpy, device to host)."""
    import ctypes
    hip = ctypes.CDLL('libamdhip64.so.7')
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    p = s.field(name)[0]  # (sord_field waits for the library's stream)
    out = np.empty(s.tensor_numel(), '<f%d' % s.realsize)
    assert hip.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0
    return out


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view('<u4' if a.dtype.itemsize == 4 else '<u8')


@pytest.mark.parametrize('realsize', [4, 8])
@pytest.mark.parametrize('case', sorted(cases.CASES) + sorted(cases.ORACLE_CASES))
def test_device_init_matches_host_init(case, realsize):
    print('case', case, 'realsize', realsize, flush=True)
    got = {}

    def grab(tag):
        def probe(s):
            got[tag] = {n: device_array(s, n) for n in FIELDS}
            got[tag]['material.json'] = open('material.json').read()
        return probe

    out_d = run_case(case, 'hip', realsize, probe=grab('device'))
    out_h = run_case(case, 'hip', realsize, probe=grab('host'), device_init=False)
    for n in FIELDS:
        same = np.array_equal(bits(got['device'][n]), bits(got['host'][n]))
        assert same, (n, int((bits(got['device'][n]) != bits(got['host'][n])).sum()))
    assert got['device']['material.json'] == got['host']['material.json']
    assert sorted(out_d) == sorted(out_h)
    for k in out_d:
        assert np.array_equal(bits(out_d[k]), bits(out_h[k])), k
