"""Launch-option autotuner for batches of SuBSENSE streams.

The batch's launch options (grid scaling of the fix-point and cooperative match phases per stream,
depth of the sequential match phase, sample loads in flight per lane, number of full-grid fix-point
launches) change the order of
evaluation but never the results, so the tuner can run its trials on live frames: it steps the
batch through its warm-up frames, cycling over the candidate option sets (one step each, in turn,
so content drift between frames affects every candidate alike), times each step's GPU work with
HIP events, and keeps the candidate with the lowest median step time for the remaining steps.
"""
import numpy as np

# fix_px_per_block / cont_px_per_block scale each stream's grid with its pixel count, so streams of
# every resolution in the batch get a grid sized to their worklist
CANDIDATES = [
    {"fix_px_per_block": 8192, "fixpoint_launches": 2, "seq_limit": 255, "fetch_depth": 4},
    {"fix_px_per_block": 8192, "fixpoint_launches": 2, "seq_limit": 255, "fetch_depth": 2},
    {"fix_px_per_block": 8192, "fixpoint_launches": 3, "seq_limit": 255, "fetch_depth": 4},
    {"fix_px_per_block": 8192, "fixpoint_launches": 2, "seq_limit": 8, "cont_px_per_block": 2048, "fetch_depth": 4},
    {"fix_px_per_block": 8192, "fixpoint_launches": 2, "seq_limit": 12, "cont_px_per_block": 2048, "fetch_depth": 4},
    {"fix_px_per_block": 8192, "fixpoint_launches": 2, "seq_limit": 16, "cont_px_per_block": 2048, "fetch_depth": 4},
]


def tune_batch(batch, frame_ptrs, n_steps, candidates=None, log=None):
    """Steps `batch` through frames 0..n_steps-1 (frame_ptrs(t) -> (image pointers, mask pointers)),
    using the first steps as timed trials of the candidates; leaves the best options set and returns them."""
    cands = candidates or CANDIDATES
    base = batch.options()
    rounds = n_steps // len(cands)
    times = [[] for _ in cands]
    t = 0
    if rounds > 0:
        batch.set_timing("step")
        for _ in range(rounds):
            for i, c in enumerate(cands):
                batch.set_options(**dict(base, **c))
                batch.apply_device(*frame_ptrs(t))
                t += 1
                ms, _ = batch.take_timing()
                times[i].append(ms)
        batch.set_timing("")
        med = [float(np.median(x)) for x in times]
        best = int(np.argmin(med))
        if log:
            log("autotune: " + ", ".join("%s -> %.3f ms" % (c, m) for c, m in zip(cands, med)))
        batch.set_options(**dict(base, **cands[best]))
    for t in range(t, n_steps):
        batch.apply_device(*frame_ptrs(t))
    batch.synchronize()
    return batch.options()
