"""
Benchmark of the SORD solver step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--workload layered|explosion|rupture]

Workload: the layered half-space model, 321 x 321 x 161 nodes, dx = 50 m,
dt = 0.004 s, free surface on top, 10-node PML elsewhere, viscous hourglass,
strike-slip double couple at 2 km, 16 x 16 surface receivers, FP32
(coseis_amd.configs.layered_321; --workload explosion / rupture time the
61^3 explosion and the 400x200x200 rupture configurations instead, with the
same contract).  The run is staged with W + K time steps;
W steps run untimed, then exactly K steps are timed (HIP events on the solver
stream, bracketed by device synchronisation), with their field i/o and
statistics as the reference performs them.  value = node-updates per second
over the K steps (nodes nx*ny*nz per step, halo excluded).

roofline: the fused interior kernel (k_fused) is the dominant kernel; its
algorithmic bytes per node-update are 68 B in FP32 (read u, v: 24 B, write u,
v: 24 B, cell lambda, mu, yy: 12 B, node 1/m, gamma: 8 B); one launch advances
the interior box, so achieved = 68 B x nodes / mean launch duration (a pair of
HIP events per launch that the launch itself stamps with the kernel's start
and end, hipExtLaunchKernelGGL, on the solver stream, on which every kernel of
the step runs in order), over --boundary-steps steps run after the timed ones
(the timed steps carry no per-launch events: a launch that records its start
and end adds gaps to the step).  traffic: HBM bytes per launch from rocprofv3
FETCH_SIZE x 2 + WRITE_SIZE x 1 (factors from a 1 GiB copy with the same
4-byte access width), read from the committed summary in profiles/.
step_roofline: the whole step, 68 B per node plus 48 B per node and damped
axis (PML memory g and p), over the measured step time.  roofline.boundary:
each boundary launch of a step on the same bytes (its nodes and their damped
axes) over its mean duration, from --boundary-steps further steps with every
boundary launch stamping a pair of HIP events the same way.

batched_vs_plain: the batched ensemble (64 seeded scenarios, 48..160 nodes
per edge, PML, FP32; batched_leg) against the same scenarios run one after
another.  grad_over_forward: the adjoint gradient on 200^3 nodes (free
surfaces, 8 double couples, 64 surface receivers; the forward's last states
kept in device memory, sqrt(N) checkpoints before them; adjoint_leg)
against one plain forward run of the same model.

cpu_baseline: the reference solver itself (the amdflang -O3 -fopenmp build of
the reference Fortran, oracle/_ref/r4/sord.x) on this machine's host cores, on
a bounded sample of the same workload (the same mesh and sources, a few time
steps, timed from the per-step profile the solver writes); the C/OpenMP oracle
stands in when the reference build is absent ('port').
"""
import argparse
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

ALG_BYTES = {4: 68, 8: 136}
WORKLOADS = {
    'layered': ('layered_321', 'layered half-space 321x321x161, dx 50 m, dt 0.004 s, free surface + '
                               '10-node PML, viscous hourglass, double couple, 16x16 receivers'),
    'explosion': ('explosion_61', 'explosion in a full space 61^3, dx 100 m, dt 0.0075 s, 10-node PML, '
                                  'isotropic moment 1e16 N m'),
    'rupture': ('rupture_400', 'spontaneous rupture 400x200x200, dx 100 m, dt 0.008 s, 10-node PML, '
                               'slip-weakening vertical fault 30x15 km with nucleation'),
}
PML_BYTES = {4: 48, 8: 96}  # per node and damped axis: read + write of g and p (3 components each)


def slab_memberships(prm):
    """Sum over nodes of the number of axes along which the node is damped
    (npml planes on each PML face)."""
    n = prm['shape'][:3]
    npml = prm.get('npml', 10)
    tot = 0
    for a in range(3):
        planes = npml * ((prm['bc1'][a] == 'pml') + (prm['bc2'][a] == 'pml'))
        tot += planes * n[(a + 1) % 3] * n[(a + 2) % 3]
    return tot
HBM_PEAK_GBS = 8000.0
PROFILE = os.path.join(ROOT, 'profiles', 'k_fused_layered321_traffic.txt')


def parse():
    p = argparse.ArgumentParser()
    p.add_argument('--gpus', type=int, default=1)
    p.add_argument('--steps', type=int, default=1500)
    p.add_argument('--warmup', type=int, default=20)
    p.add_argument('--realsize', type=int, default=4)
    p.add_argument('--workload', default='layered', choices=sorted(WORKLOADS),
                   help='plain-forward workload (layered: the metric; explosion, rupture: the other configurations)')
    p.add_argument('--boundary-steps', type=int, default=100,
                   help='steps after the timed ones with the interior and every boundary launch timed (0: none)')
    p.add_argument('--no-cpu', action='store_true')
    p.add_argument('--cpu-steps', type=int, default=6)
    p.add_argument('--stage-only', action='store_true', help='stage kernels only (no fused kernel)')
    p.add_argument('--tiles', type=int, nargs=3, default=None, metavar=('W', 'NZ', 'NZB'),
                   help='tile shape: interior waves, planes interior/boundary')
    p.add_argument('--no-batch', action='store_true', help='skip the batched-ensemble comparison')
    p.add_argument('--batch-scenarios', type=int, default=64)
    p.add_argument('--batch-steps', type=int, default=100)
    p.add_argument('--batch-direct', action='store_true',
                   help='launch the batch steps directly instead of replaying HIP graphs')
    p.add_argument('--no-adjoint', action='store_true', help='skip the adjoint-gradient leg')
    p.add_argument('--adjoint-n', type=int, default=200, help='adjoint mesh edge (nodes)')
    p.add_argument('--adjoint-steps', type=int, default=1500)
    p.add_argument('--adjoint-resident', type=int, default=-1,
                   help='forward states kept in device memory (-1: as many as fit, 0: checkpoints only)')
    p.add_argument('--copy-probe', type=int, default=0, metavar='MIB',
                   help='only copy MIB MiB device to device (calibration of the HBM traffic counters)')
    return p.parse_args()


def copy_probe(mib, reps=5):
    """Plain device-to-device copies of known size: under rocprofv3 --pmc their
    FETCH_SIZE / WRITE_SIZE per dispatch calibrate the counters (bytes counted
    per byte moved).  k_copy_words moves one 4-byte word per lane per access,
    as the solver kernels do; the runtime's own copy (wide accesses) is run
    beside it for comparison."""
    import torch
    from coseis_amd import solver
    n = mib * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device='cuda')
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(reps):
        solver.copy_probe(x.data_ptr(), y.data_ptr(), n * 4, 0)
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({'copy_probe_bytes': n * 4, 'dispatches': reps}))


def traffic_from_profile():
    """HBM bytes per k_fused launch from the committed rocprofv3 counter summary."""
    if not os.path.exists(PROFILE):
        return None
    m = re.search(r'^traffic_bytes_per_launch:\s*([0-9.eE+]+)', open(PROFILE).read(), re.M)
    return float(m.group(1)) if m else None


def cpu_model():
    try:
        for line in open('/proc/cpuinfo'):
            if line.startswith('model name'):
                return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(prm, files, steps):
    """Time the reference solver (the amdflang build in oracle/_ref, when it was
    built) or else the oracle on `steps` steps of the same mesh, loop only."""
    import numpy as np
    from coseis_amd import configs, sord
    ref = os.path.join(ROOT, 'oracle', '_ref', 'r4', 'sord.x')
    kind = 'reference'
    exe = ref
    if not os.path.exists(ref):
        kind = 'port'
        exe = os.path.join(ROOT, 'oracle', '_build', 'sord_oracle_r4')
        if not os.path.exists(exe):
            subprocess.run(['make', '-s', '-C', os.path.join(ROOT, 'oracle')], check=True)
    cores = min(16, os.cpu_count() or 1)
    prm = dict(prm)
    prm['shape'] = prm['shape'][:3] + [steps]
    d = tempfile.mkdtemp(prefix='sord-cpu-')
    cwd = os.getcwd()
    try:
        os.chdir(d)
        configs.write_inputs(files)
        sord.stage(prm)
        env = dict(os.environ, OMP_NUM_THREADS=str(cores))
        t0 = time.perf_counter()
        r = subprocess.run([exe], check=True, capture_output=True, text=True, env=env)
        wall = time.perf_counter() - t0
        # per-step loop time in clock ticks (statistics.f90 prof-step.bin; the
        # oracle writes it too) at the rate the solver prints first
        # (sord/sord.f90:26, 'Clock rate')
        m = re.search(r'^\s*(\d+)\s+Clock rate', r.stdout, re.M)
        rate = float(m.group(1)) if m else 1000.0
        ms = np.fromfile('prof-step.bin', '<i4').astype(float) * (1000.0 / rate)
    finally:
        os.chdir(cwd)
        shutil.rmtree(d)
    nodes = prm['shape'][0] * prm['shape'][1] * prm['shape'][2]
    loop = ms.sum() / 1e3
    return {
        'value': round(nodes * steps / loop / 1e9, 5),
        'unit': 'Gnode-updates/s',
        'cores': cores,
        'kind': kind,
        'cpu': cpu_model(),
        'sample': '%d time steps of the same %dx%dx%d mesh and sources, FP32, OMP_NUM_THREADS=%d; '
                  'loop %.2f s of %.1f s wall' % (steps, prm['shape'][0], prm['shape'][1], prm['shape'][2],
                                                  cores, loop, wall),
    }


def batched_leg(nscen, steps, warmup=5, tune_steps=10, graph=True):
    """Batched ensemble against sequential plain runs of the same scenarios.

    The ensemble (coseis_amd.configs.ensemble): nscen seeded scenarios with
    edges log-uniform in 48..160 nodes, dx = 100 m, dt = 0.008 s, 1-4 layers,
    a double couple each, 10-node PML, FP32.  Batched: one solver.Batch; after
    warmup steps the autotuner picks the planes per grid-size bucket (its
    trials advance the batch), then `steps` steps are timed (HIP events),
    each step's merged launches replayed from HIP graphs (graph=False:
    launched directly).  Plain: each scenario in its own solver.Solver, warmup
    steps, then the same `steps` timed; the times add up.  Both exclude
    initialisation."""
    from coseis_amd import autotune, configs, solver, sord
    ens = configs.ensemble(n=nscen, nt=warmup + tune_steps * 14 + steps + 8)
    shapes = [p['shape'] for p, _ in ens]
    base = tempfile.mkdtemp(prefix='sord-ens-')
    cwd = os.getcwd()
    try:
        texts, dirs = [], []
        for q, (prm, files) in enumerate(ens):
            d = os.path.join(base, 'b%02d' % q)
            os.makedirs(d)
            os.chdir(d)
            configs.write_inputs(files)
            texts.append(sord.stage(prm)[0])
            dirs.append(d)
        os.chdir(cwd)
        t0 = time.perf_counter()
        with solver.Batch(texts, dirs) as b:
            init_s = time.perf_counter() - t0
            b.step(warmup)
            choice, _, trials = autotune.tune_batch(b, shapes, steps=tune_steps)
            graphs = b.set_graph(graph)
            ms_b, nu_b = b.bench(steps)
        ms_p, nu_p = 0.0, 0
        for q, text in enumerate(texts):
            d = os.path.join(base, 'p%02d' % q)
            os.makedirs(d)
            os.chdir(d)
            with solver.Solver(text) as s:
                s.step(warmup)
                ms, _, _, _ = s.bench(steps)
            os.chdir(cwd)
            ms_p += ms
            nu_p += shapes[q][0] * shapes[q][1] * shapes[q][2] * steps
    finally:
        os.chdir(cwd)
        shutil.rmtree(base)
    rb, rp = nu_b / ms_b / 1e6, nu_p / ms_p / 1e6
    return round(rb / rp, 3), {
        'scenarios': nscen, 'steps': steps, 'nodes': sum(s[0] * s[1] * s[2] for s in shapes),
        'batched_gnode_s': round(rb, 3), 'batched_ms_per_step': round(ms_b / steps, 4),
        'plain_sequential_gnode_s': round(rp, 3), 'plain_ms_per_step_sum': round(ms_p / steps, 4),
        'tiles_by_bucket': {str(k): list(v) for k, v in sorted(choice.items())},
        'autotune_trials': len(trials), 'init_s': round(init_s, 2), 'graphs': bool(graphs),
    }


def adjoint_leg(n, nt, resident=-1):
    """Gradient of the waveform misfit against one plain forward run.

    Workload (coseis_amd.configs.adjoint_200): n^3 nodes, dx = 50 m, nt
    steps, free surfaces, 5% Gaussian heterogeneity (1 km correlation), 8
    double couples, 64 surface receivers, FP32.  Observed traces come from
    the second seeded model.  Plain: the forward run of the model, timed with
    HIP events (sord_bench).  Gradient: the checkpointed forward (every
    round(sqrt(nt)) steps, traces gathered each step) plus the backward sweep
    (recomputation of each segment, adjoint steps), both HIP-event timed
    inside the library.  Initialisation is excluded from both.  resident=-1:
    the forward's last states stay in device memory as far as 95 % of the
    free memory allows, and only the segments before them are checkpointed
    and recomputed (sord_adjoint_set_resident)."""
    import numpy as np
    import torch
    from coseis_amd import adjoint, configs
    w = configs.adjoint_200(n=n, nt=nt)
    prob = adjoint.Problem(w['prm'], w['receivers'], realsize=4, resident=resident)
    obs = adjoint.traces(*w['observed'], prob)
    s, d = prob.open(*w['model'])
    cwd = os.getcwd()
    try:
        os.chdir(d)
        ms_plain = s.bench(nt)[0]
    finally:
        os.chdir(cwd)
        s.close()
        shutil.rmtree(d)
    m = [torch.tensor(a, requires_grad=True) for a in w['model']]
    t0 = time.perf_counter()
    tr = adjoint.SordGradient.apply(m[0], m[1], m[2], prob)
    J = 0.5 * ((tr - torch.from_numpy(obs)) ** 2).sum()
    J.backward()
    wall = time.perf_counter() - t0
    ms_f, ms_b = prob.last_timing
    first, res_bytes = prob.last_resident
    g = [x.grad.numpy() for x in m]
    return round((ms_f + ms_b) / ms_plain, 3), {
        'nodes': n ** 3, 'steps': nt, 'checkpoint_every': prob.every, 'receivers': len(w['receivers']),
        'plain_forward_ms': round(ms_plain, 2), 'checkpointed_forward_ms': round(ms_f, 2),
        'backward_ms': round(ms_b, 2), 'gradient_wall_s': round(wall, 3),
        'misfit': float(J.detach()), 'grad_norm': [float(np.linalg.norm(x)) for x in g],
        'plain_gnode_s': round(n ** 3 * nt / ms_plain / 1e6, 3),
        'resident_from_step': first, 'resident_gb': round(res_bytes / 1e9, 1),
    }


def boundary_rooflines(launches, realsize, steps):
    """Per boundary launch of a step (k_fusedp, one damped-axis mask each):
    algorithmic bytes = 68 B per node (FP32) + 48 B per node and damped axis,
    over the launch's mean duration (HIP events that the launch stamps with
    the kernel's start and end, on the solver stream, over `steps` steps run
    after the timed ones)."""
    out = []
    for mask, ms, nodes, damped in launches:
        alg = ALG_BYTES[realsize] * nodes + PML_BYTES[realsize] * damped
        gbs = alg / (ms / 1e3) / 1e9 if ms > 0 else None
        out.append({'kernel': 'k_fusedp mask %d' % mask, 'nodes': int(nodes), 'damped_memberships': int(damped),
                    'alg_bytes': int(alg), 'launch_ms': round(ms, 5),
                    'achieved': round(gbs, 1) if gbs else None,
                    'frac': round(gbs / HBM_PEAK_GBS, 4) if gbs else None, 'steps': steps})
    return out


def progress(msg):
    print('bench: ' + msg, file=sys.stderr, flush=True)


def main():
    a = parse()
    if a.copy_probe:
        copy_probe(a.copy_probe)
        return
    if a.gpus != 1:
        print('bench.py: this build runs on one GPU; a run on %d GPUs is not provided '
              '(the multi-GPU domain decomposition is design-only)' % a.gpus, file=sys.stderr)
        sys.exit(2)
    import torch
    from coseis_amd import configs, solver, sord
    nt = a.warmup + a.steps + a.boundary_steps
    prm, files = getattr(configs, WORKLOADS[a.workload][0])(nt=nt)
    nx, ny, nz = prm['shape'][:3]
    d = tempfile.mkdtemp(prefix='sord-bench-')
    cwd = os.getcwd()
    try:
        os.chdir(d)
        configs.write_inputs(files, a.realsize)
        text = sord.stage(prm, realsize=a.realsize)[0]
        torch.cuda.init()
        s = solver.Solver(text, realsize=a.realsize, device=0, fused=not a.stage_only, tiles=a.tiles)
        if a.warmup:
            s.step(a.warmup)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ms_total, _, knodes, kname = s.bench(a.steps)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        # per-launch durations from further steps whose launches stamp their
        # own start and end (kept out of the timed steps: the stamps add gaps)
        bnd = s.bench_boundary(a.boundary_steps) if a.boundary_steps and not a.stage_only else []
        s.close()
    finally:
        os.chdir(cwd)
        shutil.rmtree(d)
    ms_kernel = next((ms for mask, ms, _, _ in bnd if mask == -1), 0.0)
    bnd = [b for b in bnd if b[0] != -1]
    nodes = nx * ny * nz
    secs = ms_total / 1e3
    value = nodes * a.steps / secs / 1e9
    alg = ALG_BYTES[a.realsize] * knodes
    step_bytes = ALG_BYTES[a.realsize] * nodes + PML_BYTES[a.realsize] * slab_memberships(prm)
    achieved = alg / (ms_kernel / 1e3) / 1e9 if ms_kernel > 0 else None
    line = {
        'metric': 'Gnode-updates/s', 'value': round(value, 3), 'unit': 'Gnode-updates/s',
        'n_gpus': 1, 'steps': a.steps, 'warmup': a.warmup,
        'ms_per_step': round(ms_total / a.steps, 4),
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'f32' if a.realsize == 4 else 'f64', 'data': 'synthetic',
        'config': {'workload': WORKLOADS[a.workload][1],
                   'nodes': nodes, 'realsize': a.realsize,
                   'path': 'stage kernels' if a.stage_only else
                           'fused interior kernel, then the fused boundary kernels, in order on one stream'},
        'wall_s': round(wall, 4),
        'roofline': {
            'bound': 'hbm', 'kernel': kname,
            'achieved': round(achieved, 1) if achieved else None, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
            'frac': round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
            # the committed counter summary measures the layered FP32 fused path
            'traffic': traffic_from_profile() if (a.workload, a.realsize, a.stage_only) == ('layered', 4, False)
                       else None,
            'alg_bytes_per_launch': alg, 'nodes_per_launch': knodes,
            'launch_ms': round(ms_kernel, 5),
        },
        'step_roofline': {
            'alg_bytes_per_step': step_bytes,
            'achieved': round(step_bytes * a.steps / secs / 1e9, 1), 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
            'frac': round(step_bytes * a.steps / secs / 1e9 / HBM_PEAK_GBS, 4),
        },
        'batched_vs_plain': None,
        'grad_over_forward': None,
    }
    if bnd:
        line['roofline']['boundary'] = boundary_rooflines(bnd, a.realsize, a.boundary_steps)
    progress('plain forward: %.3f Gnode/s' % value)
    if not a.no_batch:
        line['batched_vs_plain'], line['batched'] = batched_leg(a.batch_scenarios, a.batch_steps, graph=not a.batch_direct)
        progress('batched leg: %s x plain' % line['batched_vs_plain'])
    if not a.no_adjoint:
        line['grad_over_forward'], line['adjoint'] = adjoint_leg(a.adjoint_n, a.adjoint_steps, a.adjoint_resident)
        progress('adjoint leg: gradient / forward %s' % line['grad_over_forward'])
    if not a.no_cpu:
        line['cpu_baseline'] = cpu_baseline(prm, files, a.cpu_steps)
        progress('cpu baseline done')
    print(json.dumps(line))


if __name__ == '__main__':
    main()
