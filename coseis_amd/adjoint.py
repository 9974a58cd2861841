"""
Gradient of a waveform misfit with respect to per-cell rho, vp, vs.

`SordGradient` is a torch.autograd.Function: its forward stages the material
as cell input files, runs the solver with checkpoints every `every` steps
(sqrt(nt) by default) and returns the velocity traces at the receivers; its
backward runs the hand-written adjoint step kernels (sord_adjoint.hip) over
the recomputed segments (or, with Problem(resident=-1), over the forward's
last states kept in device memory) and returns dJ/drho, dJ/dvp, dJ/dvs for
the incoming dJ/dtraces.  The mesh must be rectilinear with free-surface faces, no fault;
sources are moment tensors (field-I/O ops on mxx..mxy).

    prob = Problem(prm, receivers)          # prm: cst.sord parameters without rho/vp/vs
    tr = SordGradient.apply(rho, vp, vs, prob)
    J = 0.5 * ((tr - obs) ** 2).sum(); J.backward()
"""
import math
import os
import shutil
import tempfile
import weakref

import numpy as np
import torch

from . import solver, sord


class Problem(object):
    """Mesh, sources and receivers of a gradient computation."""

    def __init__(self, prm, receivers, every=None, realsize=4, device=0, resident=0):
        self.prm = dict(prm)
        for f in 'rho', 'vp', 'vs':
            self.prm.pop(f, None)
        self.receivers = np.asarray(receivers, dtype=np.int32).reshape(-1, 3)
        nt = self.prm['shape'][3]
        self.every = every or max(1, int(round(math.sqrt(nt))))
        self.realsize = realsize
        self.device = device
        # forward states kept in device memory instead of recomputed (-1: as
        # many as fit in the free memory)
        self.resident = resident
        self.cells = tuple(n - 1 for n in self.prm['shape'][:3])
        self.last_timing = None
        self.last_resident = None  # (first resident state, bytes) of the last backward

    def open(self, rho, vp, vs):
        """Stage the run in a scratch directory; returns (solver, directory)."""
        d = tempfile.mkdtemp(prefix='sord-adj-')
        cwd = os.getcwd()
        dt = '<f%d' % self.realsize
        prm = dict(self.prm)
        try:
            os.chdir(d)
            for name, a in (('rho', rho), ('vp', vp), ('vs', vs)):
                a = np.asarray(a, dtype=dt)
                if a.shape != self.cells:
                    raise ValueError('%s: cell array of shape %s expected' % (name, self.cells))
                a.T.tofile(name + '.bin')
                prm[name] = [([[], [], []], '=<', name + '.bin')]
            text = sord.stage(prm, realsize=self.realsize)[0]
            s = solver.Solver(text, realsize=self.realsize, device=self.device)
            # the scratch directory goes with the solver (a graph whose backward
            # never runs frees both when it is dropped)
            weakref.finalize(s, shutil.rmtree, d, True)
            s.adjoint_setup(self.receivers, self.every, self.resident)
        except Exception:
            os.chdir(cwd)
            shutil.rmtree(d)
            raise
        os.chdir(cwd)
        return s, d


class SordGradient(torch.autograd.Function):
    """traces = SordGradient.apply(rho, vp, vs, problem); see the module."""

    @staticmethod
    def forward(ctx, rho, vp, vs, problem):
        s, d = problem.open(rho.detach().cpu().numpy(), vp.detach().cpu().numpy(), vs.detach().cpu().numpy())
        cwd = os.getcwd()
        try:
            os.chdir(d)
            tr = s.adjoint_forward()
        finally:
            os.chdir(cwd)
        ctx.solver, ctx.dir, ctx.problem = s, d, problem
        return torch.from_numpy(tr).to(rho.device)

    @staticmethod
    def backward(ctx, grad):
        s, d, problem = ctx.solver, ctx.dir, ctx.problem
        cwd = os.getcwd()
        try:
            os.chdir(d)
            g = s.adjoint_backward(grad.detach().cpu().numpy(), problem.cells)
            problem.last_timing = s.adjoint_timing()
            problem.last_resident = s.adjoint_resident()
        finally:
            os.chdir(cwd)
            s.close()
            shutil.rmtree(d)
        dev = grad.device
        return (torch.from_numpy(np.ascontiguousarray(g[0])).to(dev),
                torch.from_numpy(np.ascontiguousarray(g[1])).to(dev),
                torch.from_numpy(np.ascontiguousarray(g[2])).to(dev), None)


def traces(rho, vp, vs, problem):
    """Forward only: the receiver traces (numpy) for a material."""
    s, d = problem.open(rho, vp, vs)
    cwd = os.getcwd()
    try:
        os.chdir(d)
        return s.adjoint_forward()
    finally:
        os.chdir(cwd)
        s.close()
        shutil.rmtree(d)
