"""Launch-option autotuning for batches of SuBSENSE streams.

The batch's launch options (grid scaling of the fix-point and cooperative match phases per stream,
depth of the sequential match phase, sample loads in flight per lane, number of full-grid fix-point
launches) change the order of evaluation but never the results, so trials run on live frames.

`OnlineTuner` steps a batch and keeps it on the fastest candidate as the content drifts: every
`period` steps it runs one trial round (one step per candidate, in turn, so drift between frames
affects every candidate alike), times each trial step's GPU work with HIP events, and switches to the
candidate with the lowest median over the last `window` rounds. The best depth of the sequential
match phase depends on how far the models have settled (a scene's first frames favour the full scan,
settled models a short one), which is why the trials never stop.

`tune_batch` is the one-shot form: trial rounds over a fixed number of warm-up steps.
"""
import numpy as np

# fix_px_per_block / cont_px_per_block scale each stream's grid with its pixel count, so streams of
# every resolution in the batch get a grid sized to their worklist
CANDIDATES = [
    {"fix_px_per_block": 8192, "fixpoint_launches": 2, "seq_limit": 8, "cont_px_per_block": 2048, "fetch_depth": 4},
    {"fix_px_per_block": 8192, "fixpoint_launches": 2, "seq_limit": 255, "fetch_depth": 4},
    {"fix_px_per_block": 8192, "fixpoint_launches": 2, "seq_limit": 6, "cont_px_per_block": 2048, "fetch_depth": 4},
    {"fix_px_per_block": 8192, "fixpoint_launches": 2, "seq_limit": 12, "cont_px_per_block": 2048, "fetch_depth": 4},
    {"fix_px_per_block": 8192, "fixpoint_launches": 3, "seq_limit": 8, "cont_px_per_block": 2048, "fetch_depth": 4},
    {"fix_px_per_block": 8192, "fixpoint_launches": 2, "seq_limit": 8, "cont_px_per_block": 2048, "fetch_depth": 2},
]


class OnlineTuner:
    """Wraps batch.apply_device; starts on candidates[0] and re-evaluates every `period` steps."""

    def __init__(self, batch, candidates=None, period=64, window=4, log=None):
        self.batch = batch
        self.cands = list(candidates or CANDIDATES)
        self.period = max(int(period), len(self.cands))
        self.window = int(window)
        self.base = batch.options()
        self.hist = [[] for _ in self.cands]
        self.best = 0
        self.rounds = 0
        self.n = 0
        self.log = log
        batch.set_options(**dict(self.base, **self.cands[0]))

    def apply(self, images, masks):
        pos = self.n % self.period
        self.n += 1
        if pos >= len(self.cands):
            self.batch.apply_device(images, masks)
            return
        if pos == 0:
            self.batch.set_timing("step")
        self.batch.set_options(**dict(self.base, **self.cands[pos]))
        self.batch.apply_device(images, masks)
        ms, _ = self.batch.take_timing()
        self.hist[pos] = (self.hist[pos] + [ms])[-self.window:]
        if pos == len(self.cands) - 1:
            self.batch.set_timing("")
            med = [float(np.median(h)) for h in self.hist]
            self.best = int(np.argmin(med))
            self.rounds += 1
            self.batch.set_options(**dict(self.base, **self.cands[self.best]))
            if self.log:
                self.log("autotune round %d: %s -> best %d" % (self.rounds, ", ".join("%.3f" % m for m in med),
                                                                 self.best))

    def options(self):
        """The options the batch is left on (the current best candidate)."""
        return self.batch.options()

    def medians(self):
        return [float(np.median(h)) if h else None for h in self.hist]


def tune_batch(batch, frame_ptrs, n_steps, candidates=None, log=None):
    """Steps `batch` through frames 0..n_steps-1 (frame_ptrs(t) -> (image pointers, mask pointers)),
    using every step as a trial of the candidates in turn; leaves the best options set and returns them."""
    cands = list(candidates or CANDIDATES)
    tuner = OnlineTuner(batch, cands, period=len(cands), window=max(1, n_steps // len(cands)), log=log)
    rounds = n_steps // len(cands)
    for t in range(rounds * len(cands)):
        tuner.apply(*frame_ptrs(t))
    for t in range(rounds * len(cands), n_steps):
        batch.apply_device(*frame_ptrs(t))
    batch.synchronize()
    return batch.options()
