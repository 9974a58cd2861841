"""
ctypes binding of the HIP solver library (C ABI: include/sord_amd.h).

The library is loaded from this package directory; there is no fallback: a
missing or unloadable library raises ``ImportError``-class errors at first use.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, 'libsord_amd.so')

# error codes of include/sord_amd.h
ERRORS = {
    1: 'EINPUT', 2: 'ECFL', 3: 'EMODULI', 4: 'EVOLUME', 5: 'ENAN', 6: 'ENANIN',
    7: 'EBADOP', 8: 'EFILE', 9: 'EHIP', 10: 'EUNSUPPORTED', 11: 'ESTATE',
}
EXPORTS = [
    'sord_create', 'sord_init', 'sord_step', 'sord_step_static', 'sord_run', 'sord_status', 'sord_field', 'sord_set_tensor',
    'sord_tensor_numel',
    'sord_bench', 'sord_bench_boundary', 'sord_set_fused', 'sord_set_device_init', 'sord_set_streams', 'sord_set_tiles', 'sord_plan', 'sord_copy_probe',
    'sord_batch_create', 'sord_batch_set_graph', 'sord_batch_set_streams', 'sord_batch_set_tiles', 'sord_batch_init', 'sord_batch_step', 'sord_batch_run',
    'sord_batch_bench', 'sord_batch_size', 'sord_batch_destroy',
    'sord_adjoint_setup', 'sord_adjoint_forward', 'sord_adjoint_backward', 'sord_adjoint_timing',
    'sord_adjoint_set_resident', 'sord_adjoint_resident',
    'sord_last_error', 'sord_destroy',
]

_lib = None


class SolverError(RuntimeError):
    """Error returned by the solver library; ``code`` is the SORD_E* value."""

    def __init__(self, code, msg):
        self.code = code
        self.name = ERRORS.get(code, str(code))
        RuntimeError.__init__(self, 'SORD_%s: %s' % (self.name, msg))


def load():
    """Load libsord_amd.so (once) and declare the C signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise ImportError('HIP solver library not built: %s (python -m coseis_amd.build)' % LIBPATH)
    lib = ctypes.CDLL(LIBPATH)
    vp = ctypes.c_void_p
    lib.sord_create.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(vp)]
    lib.sord_init.argtypes = [vp]
    lib.sord_step.argtypes = [vp, ctypes.c_int]
    lib.sord_step_static.argtypes = [vp, ctypes.c_int]
    lib.sord_run.argtypes = [vp]
    lib.sord_status.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    i64p = ctypes.POINTER(ctypes.c_int64)
    lib.sord_field.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(vp), i64p, i64p, i64p]
    lib.sord_set_tensor.argtypes = [vp, ctypes.c_char_p, vp, ctypes.c_int64]
    lib.sord_tensor_numel.argtypes = [vp, i64p]
    lib.sord_bench.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                               ctypes.POINTER(ctypes.c_double), i64p, ctypes.c_char_p, ctypes.c_int]
    lib.sord_set_fused.argtypes = [vp, ctypes.c_int]
    lib.sord_set_device_init.argtypes = [vp, ctypes.c_int]
    dp = ctypes.POINTER(ctypes.c_double)
    lib.sord_bench_boundary.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                        dp, dp, dp, ctypes.c_int]
    lib.sord_set_tiles.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.sord_set_streams.argtypes = [vp, ctypes.c_int]
    lib.sord_plan.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), ctypes.c_int]
    lib.sord_copy_probe.argtypes = [vp, vp, ctypes.c_int64, ctypes.c_int]
    lib.sord_last_error.restype = ctypes.c_char_p
    lib.sord_last_error.argtypes = []
    lib.sord_destroy.argtypes = [vp]
    ccp = ctypes.POINTER(ctypes.c_char_p)
    lib.sord_batch_create.argtypes = [ctypes.c_int, ccp, ccp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(vp)]
    lib.sord_batch_set_tiles.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.sord_batch_set_graph.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    lib.sord_batch_set_streams.argtypes = [vp, ctypes.c_int]
    lib.sord_batch_init.argtypes = [vp]
    lib.sord_batch_step.argtypes = [vp, ctypes.c_int]
    lib.sord_batch_run.argtypes = [vp]
    lib.sord_batch_bench.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_double), i64p]
    lib.sord_batch_size.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
    lib.sord_batch_destroy.argtypes = [vp]
    lib.sord_adjoint_setup.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int32), ctypes.c_int]
    lib.sord_adjoint_forward.argtypes = [vp, vp]
    lib.sord_adjoint_backward.argtypes = [vp, vp, vp, vp, vp]
    lib.sord_adjoint_timing.argtypes = [vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    lib.sord_adjoint_set_resident.argtypes = [vp, ctypes.c_int64]
    lib.sord_adjoint_resident.argtypes = [vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]
    for name in EXPORTS:
        getattr(lib, name).restype = ctypes.c_int
    lib.sord_last_error.restype = ctypes.c_char_p
    lib.sord_destroy.restype = None
    lib.sord_batch_destroy.restype = None
    _lib = lib
    return lib


def _check(rc):
    if rc != 0:
        msg = _lib.sord_last_error().decode(errors='replace')
        raise SolverError(rc, msg)


def copy_probe(src_ptr, dst_ptr, nbytes, device=0):
    """Copy nbytes between two device buffers with the library's calibration
    kernel (one 4-byte word per lane per access) and wait for it."""
    _check(load().sord_copy_probe(src_ptr, dst_ptr, nbytes, device))


class Solver(object):
    """One solver handle: ``sord.in`` text in, reference output files out.

    The current directory at ``init()`` and during stepping receives the
    output files, as the reference ``sord.x`` writes them into its cwd.
    """

    def __init__(self, text, realsize=4, device=0, fused=True, tiles=None, side_streams=False, device_init=True):
        lib = load()
        self._h = ctypes.c_void_p()
        self._adopted = {}  # caller tensors the handle uses (set_tensor)
        _check(lib.sord_create(text.encode(), realsize, device, ctypes.byref(self._h)))
        self.realsize = realsize
        if not device_init:
            _check(lib.sord_set_device_init(self._h, 0))
        if not fused:
            _check(lib.sord_set_fused(self._h, 0))
        if side_streams:
            _check(lib.sord_set_streams(self._h, 1))
        if tiles:
            self.set_tiles(*tiles)
        _check(lib.sord_init(self._h))

    def set_streams(self, side=True):
        """Boundary kernels on side streams (opt-in) or in order on the
        handle's one stream (the default)."""
        _check(_lib.sord_set_streams(self._h, 1 if side else 0))

    def set_tiles(self, waves=8, planes=16, planes_boundary=16):
        """Tile shape of the fused kernels (performance only; see sord_set_tiles)."""
        _check(_lib.sord_set_tiles(self._h, waves, planes, planes_boundary))

    def _caller_sync(self):
        # The library's stream is not ordered against torch's: work the caller
        # queued on adopted tensors (fills, edits) must finish first.
        if self._adopted:
            import torch
            torch.cuda.current_stream().synchronize()

    def boundary_masks(self):
        """Damped-axis masks of the planned boundary launches (sord_plan;
        None when the stage kernels run the steps)."""
        n = ctypes.c_int()
        m = (ctypes.c_int * 64)()
        _check(_lib.sord_plan(self._h, ctypes.byref(n), m, 64))
        return None if n.value < 0 else [m[i] for i in range(min(n.value, 64))]

    def step(self, n=1):
        self._caller_sync()
        _check(_lib.sord_step(self._h, n))

    def step_static(self, n=1):
        """As step(), with no device allocation (sord_step_static)."""
        self._caller_sync()
        _check(_lib.sord_step_static(self._h, n))

    def run(self):
        self._caller_sync()
        _check(_lib.sord_run(self._h))

    def status(self):
        it, nt = ctypes.c_int(), ctypes.c_int()
        _check(_lib.sord_status(self._h, ctypes.byref(it), ctypes.byref(nt)))
        return it.value, nt.value

    def bench(self, nsteps):
        """Advance nsteps timed: (ms all steps, 0 (the launches are not
        instrumented; see bench_boundary), nodes per dominant-kernel launch,
        kernel name)."""
        tot, ker = ctypes.c_double(), ctypes.c_double()
        nodes = ctypes.c_int64()
        name = ctypes.create_string_buffer(64)
        _check(_lib.sord_bench(self._h, nsteps, ctypes.byref(tot), ctypes.byref(ker), ctypes.byref(nodes),
                               name, 64))
        return tot.value, ker.value, nodes.value, name.value.decode()

    def bench_boundary(self, nsteps):
        """Advance nsteps steps timing the interior launch and each boundary
        launch (one stream): [(mask, mean ms, nodes, damped-axis memberships)]
        per launch of a step, the interior launch first with mask -1
        (sord_bench_boundary)."""
        n = ctypes.c_int()
        cap = 64
        m = (ctypes.c_int * cap)()
        ms, nodes, damped = (ctypes.c_double * cap)(), (ctypes.c_double * cap)(), (ctypes.c_double * cap)()
        _check(_lib.sord_bench_boundary(self._h, nsteps, ctypes.byref(n), m, ms, nodes, damped, cap))
        return [(m[i], ms[i], nodes[i], damped[i]) for i in range(min(n.value, cap))]

    # ---- adjoint gradient (sord_adjoint_*) ----
    def adjoint_setup(self, receivers, every, resident=0):
        """Receivers: (n, 3) 0-based node indices; checkpoint every `every`
        steps; keep the last `resident` forward states in device memory (-1:
        as many as fit; sord_adjoint_set_resident)."""
        import numpy as np
        _check(_lib.sord_adjoint_set_resident(self._h, int(resident)))
        rec = np.ascontiguousarray(receivers, dtype=np.int32)
        self._nrec = len(rec)
        _check(_lib.sord_adjoint_setup(self._h, self._nrec, rec.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                       every))

    def adjoint_forward(self):
        """Run all steps with checkpoints; returns the traces (nt, nrec, 3):
        velocity at the receivers at the start of each step."""
        import numpy as np
        it, nt = self.status()
        tr = np.zeros((nt, self._nrec, 3), dtype='f%d' % self.realsize)
        _check(_lib.sord_adjoint_forward(self._h, tr.ctypes.data))
        return tr

    def adjoint_backward(self, residual, cells):
        """Gradient of J given dJ/dtraces (nt, nrec, 3); returns dJ/drho,
        dJ/dvp, dJ/dvs of shape `cells` (nx-1, ny-1, nz-1)."""
        import numpy as np
        res = np.ascontiguousarray(residual, dtype='f%d' % self.realsize)
        n = cells[0] * cells[1] * cells[2]
        g = [np.zeros(n, dtype='f%d' % self.realsize) for _ in range(3)]
        _check(_lib.sord_adjoint_backward(self._h, res.ctypes.data, g[0].ctypes.data, g[1].ctypes.data,
                                          g[2].ctypes.data))
        return [x.reshape(cells[::-1]).T for x in g]

    def adjoint_resident(self):
        """(first resident state, bytes) of the adjoint's resident tail."""
        a, b = ctypes.c_int64(), ctypes.c_int64()
        _check(_lib.sord_adjoint_resident(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def adjoint_timing(self):
        """Wall time (ms) of the last adjoint forward and backward."""
        f, b = ctypes.c_double(), ctypes.c_double()
        _check(_lib.sord_adjoint_timing(self._h, ctypes.byref(f), ctypes.byref(b)))
        return f.value, b.value

    def field(self, name):
        """(device pointer, padded shape, strides, offset) of a device array."""
        p = ctypes.c_void_p()
        shp = (ctypes.c_int64 * 3)()
        st = (ctypes.c_int64 * 3)()
        off = ctypes.c_int64()
        _check(_lib.sord_field(self._h, name.encode(), ctypes.byref(p), shp, st, ctypes.byref(off)))
        return p.value, tuple(shp), tuple(st), off.value

    def tensor_numel(self):
        """Elements of one padded solver array (what set_tensor needs)."""
        n = ctypes.c_int64()
        _check(_lib.sord_tensor_numel(self._h, ctypes.byref(n)))
        return n.value

    def set_tensor(self, name, tensor):
        """Adopt a caller-owned device tensor (flat, contiguous, of the
        solver's precision, at least tensor_numel() elements) for the named
        array (sord_set_tensor): its current contents are copied in and the
        kernels use it from then on.  The tensor must outlive the handle's use
        of it; the handle keeps a reference to it."""
        import torch
        want = torch.float32 if self.realsize == 4 else torch.float64
        if tensor.dtype != want or not tensor.is_cuda or not tensor.is_contiguous():
            raise ValueError('set_tensor: a contiguous %s device tensor' % want)
        # the adoption copies the handle's array into the tensor on the
        # library's stream: the caller's pending work on it goes first
        torch.cuda.current_stream().synchronize()
        _check(_lib.sord_set_tensor(self._h, name.encode(), ctypes.c_void_p(tensor.data_ptr()), tensor.numel()))
        self._adopted[name] = tensor

    def close(self):
        if self._h:
            _lib.sord_destroy(self._h)
            self._h = None
        self._adopted = {}

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch(object):
    """Many scenarios stepped together on one device (sord_batch_*): texts are
    sord.in texts, dirs the run directory of each (its output files go
    there).  tiles: optional list of (planes, planes_boundary) per scenario."""

    def __init__(self, texts, dirs=None, realsize=4, device=0, tiles=None):
        lib = load()
        n = len(texts)
        t = (ctypes.c_char_p * n)(*[x.encode() for x in texts])
        d = (ctypes.c_char_p * n)(*[(x or '').encode() for x in (dirs or [''] * n)])
        self._h = ctypes.c_void_p()
        _check(lib.sord_batch_create(n, t, d, realsize, device, ctypes.byref(self._h)))
        self.n = n
        if tiles:
            for q, (p, pb) in enumerate(tiles):
                _check(lib.sord_batch_set_tiles(self._h, q, p, pb))
        _check(lib.sord_batch_init(self._h))

    def set_graph(self, on=True):
        """Replay each step's merged launches from HIP graphs (on, the
        default) or launch them directly (off); None only queries; returns
        whether graphs are in use."""
        u = ctypes.c_int()
        _check(_lib.sord_batch_set_graph(self._h, -1 if on is None else int(bool(on)), ctypes.byref(u)))
        return bool(u.value)

    def set_streams(self, side=True):
        """Merged boundary launches on side streams (opt-in) or in order on
        the batch's one stream (the default)."""
        _check(_lib.sord_batch_set_streams(self._h, 1 if side else 0))

    def set_tiles(self, index, planes, planes_boundary):
        """Planes per work item of one scenario; call replan() to apply."""
        _check(_lib.sord_batch_set_tiles(self._h, index, planes, planes_boundary))

    def replan(self):
        _check(_lib.sord_batch_set_tiles(self._h, -1, 1, 1))

    def step(self, n=1):
        _check(_lib.sord_batch_step(self._h, n))

    def run(self):
        _check(_lib.sord_batch_run(self._h))

    def bench(self, nsteps):
        """Advance nsteps timed: (ms, node-updates)."""
        ms, nu = ctypes.c_double(), ctypes.c_int64()
        _check(_lib.sord_batch_bench(self._h, nsteps, ctypes.byref(ms), ctypes.byref(nu)))
        return ms.value, nu.value

    def close(self):
        if self._h:
            _lib.sord_batch_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
