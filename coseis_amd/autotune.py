"""
Tile shapes of the fused kernels, picked by measurement.

The interior kernel of a batch has 8-wave blocks (one launch for all
scenarios); what varies per scenario is the number of planes a block
marches, in the interior (long chains amortise the two planes of start-up
loads) and in the boundary lists (short chains finish with the rest of the
step).  Scenarios are grouped by grid size (bucket of the longest edge) and
each bucket gets the pair that gives the highest node-updates/s of the whole
batch, found by coordinate descent: one bucket at a time, every candidate,
keep the best.  Each trial advances the batch by `steps` steps.
"""
BUCKET_EDGES = (0, 72, 108, 1 << 30)
CANDIDATES = ((16, 8), (32, 8), (64, 16), (128, 16))
DEFAULT = (128, 16)


def bucket(shape):
    """Index of the grid-size bucket of a mesh shape (nx, ny, nz, ...)."""
    m = max(shape[:3])
    for b in range(len(BUCKET_EDGES) - 1):
        if BUCKET_EDGES[b] <= m < BUCKET_EDGES[b + 1]:
            return b
    return len(BUCKET_EDGES) - 2


def apply(batch, shapes, choice):
    for q, s in enumerate(shapes):
        batch.set_tiles(q, *choice[bucket(s)])
    batch.replan()


def tune_batch(batch, shapes, steps=10, candidates=CANDIDATES):
    """Pick (planes, planes_boundary) per bucket for a batch (solver.Batch)
    whose scenarios have the given shapes.  Returns (choice, node-updates/s
    of the best, trials [(bucket, candidate, rate)]); the batch is left with
    the choice applied."""
    buckets = sorted(set(bucket(s) for s in shapes))
    choice = {b: DEFAULT for b in buckets}

    def rate():
        ms, nu = batch.bench(steps)
        return nu / (ms * 1e-3) if ms > 0 else 0.0

    apply(batch, shapes, choice)
    batch.bench(2)  # settle
    best = rate()
    trials = [(None, DEFAULT, best)]
    for b in buckets:
        for cand in candidates:
            if cand == choice[b]:
                continue
            old = choice[b]
            choice[b] = cand
            apply(batch, shapes, choice)
            r = rate()
            trials.append((b, cand, r))
            if r > best:
                best = r
            else:
                choice[b] = old
    apply(batch, shapes, choice)
    return choice, best, trials
