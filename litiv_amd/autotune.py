"""Launch-option autotuning for batches of SuBSENSE streams.

The batch's launch options change the order of evaluation but never the results, so trials run on
live frames. Two kinds of option are tuned:

* shared by every stream (one kernel variant per launch): sample loads in flight per lane of the
  sequential match (`fetch_depth`), full-grid fix-point launches (`fixpoint_launches`), number of
  member groups stepped concurrently on their own streams (`groups`);
* per resolution class (member overrides, `Batch.set_stream_options`): samples scanned by the
  sequential match phase before the dense continuation (`seq_limit`) and pixels per block of the
  continuation phase's grid-stride loop (`cont_px_per_block`) -- the launch shape each resolution
  gets.

`OnlineTuner` steps a batch and runs one trial round every `period` steps (one step per trial
configuration, timed with HIP events), cycling through a round over the shared candidates and one
round per (class, option). A switch needs a gain above `margin` (a fraction of the step time), so
noise does not flip options that make no difference. A whole cycle without a change doubles the period (up to `max_period`)
and any change brings it back, so trials cost little once the choices have settled. A shared round keeps the candidate with the lowest median step time over
the last `window` rounds. A class round pits the neighbouring values of the class's current choice
against it on adjacent frames and switches when the median of the paired differences favours a
neighbour. The trials never stop, because the best values move with the content (a scene's first
frames favour a full sequential scan, settled models a short one).

`tune_batch` is the one-shot form: shared-option rounds over a fixed number of warm-up steps.
"""
import numpy as np

SHARED_CANDIDATES = [
    {"fixpoint_launches": 2, "fetch_depth": 4, "groups": 3},
    {"fixpoint_launches": 2, "fetch_depth": 4, "groups": 2},
    {"fixpoint_launches": 2, "fetch_depth": 4, "groups": 1},
    {"fixpoint_launches": 2, "fetch_depth": 4, "groups": 4},
    {"fixpoint_launches": 3, "fetch_depth": 4, "groups": 3},
    {"fixpoint_launches": 2, "fetch_depth": 2, "groups": 3},
]
CLASS_CHOICES = {"seq_limit": [4, 6, 8, 12, 16, 24, 255], "cont_px_per_block": [512, 1024, 2048, 4096, 8192]}
CLASS_START = {"seq_limit": 12, "cont_px_per_block": 2048}

# older name kept for callers of tune_batch
CANDIDATES = SHARED_CANDIDATES


class OnlineTuner:
    """Wraps batch.apply_device. classes: [(key, member indices)], default batch.shape_classes()."""

    def __init__(self, batch, shared=None, classes=None, period=32, window=4, max_period=256, margin=0.01,
                 log=None):
        self.batch = batch
        self.shared = list(shared or SHARED_CANDIDATES)
        self.classes = list(classes if classes is not None else batch.shape_classes())
        self.base_period = self.period = int(period)
        self.max_period = max(int(max_period), self.period)
        self.changed = False
        self.next_round = 0
        self.margin = float(margin)
        self.window = int(window)
        self.log = log
        self.base = batch.options()
        self.shared_hist = [[] for _ in self.shared]
        self.shared_best = 0
        self.choice = [dict(CLASS_START) for _ in self.classes]
        self.deltas = {}
        # the rounds cycle: shared, then every (class, option)
        self.plan = [("shared", None, None)] + [("class", ci, opt) for ci in range(len(self.classes))
                                                for opt in CLASS_CHOICES]
        self.round_kind = 0
        self.rounds = 0
        self.n = 0
        self.trial = None  # list of (apply_config, record) for the running round
        self.pos = 0
        batch.set_options(**dict(self.base, **self.shared[0]))
        for ci in range(len(self.classes)):
            self._set_class(ci, self.choice[ci])

    def _set_class(self, ci, values):
        for i in self.classes[ci][1]:
            self.batch.set_stream_options(i, **values)

    def _start_round(self):
        kind, ci, opt = self.plan[self.round_kind % len(self.plan)]
        self.round_kind += 1
        if kind == "shared":
            cfgs = []
            for k, cand in enumerate(self.shared):
                cfgs.append((lambda c=cand: self.batch.set_options(**dict(self.base, **c, use_graph=0)), ("shared", k)))
            self.trial = cfgs
        else:
            vals = CLASS_CHOICES[opt]
            cur = self.choice[ci][opt]
            j = vals.index(cur) if cur in vals else 0
            alts = [vals[j2] for j2 in (j - 1, j + 1) if 0 <= j2 < len(vals)]
            cfgs = [(lambda: self._set_class(ci, self.choice[ci]), ("base", ci, opt))]
            for a in alts:
                cfgs.append((lambda a=a: self._set_class(ci, dict(self.choice[ci], **{opt: a})), ("alt", ci, opt, a)))
            self.trial = cfgs
        self.pos = 0
        self.times = {}
        # trial steps are launched directly: each configuration re-plans the batch, and a graph would
        # be captured again for every one of them
        self.batch.set_options(use_graph=0)
        self.batch.set_timing("step")

    def _finish_round(self):
        self.batch.set_timing("")
        self.rounds += 1
        kinds = [rec[0] for _, rec in self.trial]
        if kinds[0] == "shared":
            for (_, rec) in self.trial:
                k = rec[1]
                self.shared_hist[k] = (self.shared_hist[k] + [self.times[rec]])[-self.window:]
            med = [float(np.median(h)) if h else float("inf") for h in self.shared_hist]
            best = int(np.argmin(med))
            if med[best] > med[self.shared_best] * (1.0 - self.margin):
                best = self.shared_best
            self.changed |= best != self.shared_best
            self.shared_best = best
            msg = "shared %s -> %d" % (", ".join("%.3f" % m for m in med), self.shared_best)
        else:
            _, ci, opt = self.trial[0][1]
            base_t = self.times[self.trial[0][1]]
            best_a, best_d = None, -self.margin * base_t
            for (_, rec) in self.trial[1:]:
                a = rec[3]
                key = (ci, opt, a)
                self.deltas[key] = (self.deltas.get(key, []) + [self.times[rec] - base_t])[-self.window:]
                if len(self.deltas[key]) >= 2:
                    d = float(np.median(self.deltas[key]))
                    if d < best_d:
                        best_a, best_d = a, d
            if best_a is not None:
                self.changed = True
                self.choice[ci][opt] = best_a
                for key in [k for k in self.deltas if k[0] == ci and k[1] == opt]:
                    del self.deltas[key]
            self._set_class(ci, self.choice[ci])
            msg = "class %s %s -> %s" % (self.classes[ci][0], opt, self.choice[ci][opt])
        self.batch.set_options(**dict(self.base, **self.shared[self.shared_best]))
        if self.round_kind % len(self.plan) == 0:  # a whole cycle done
            self.period = self.base_period if self.changed else min(2 * self.period, self.max_period)
            self.changed = False
        self.next_round = self.n + self.period
        if self.log:
            self.log("autotune round %d: %s (next in %d steps)" % (self.rounds, msg, self.period))

    def apply(self, images, masks):
        if self.trial is None and self.n >= self.next_round:
            self._start_round()
        self.n += 1
        if self.trial is None:
            self.batch.apply_device(images, masks)
            return
        setup, rec = self.trial[self.pos]
        setup()
        self.batch.apply_device(images, masks)
        ms, _ = self.batch.take_timing()
        self.times[rec] = ms
        self.pos += 1
        if self.pos == len(self.trial):
            self._finish_round()
            self.trial = None

    def options(self):
        """The shared options the batch is left on."""
        return self.batch.options()

    def apply_choices(self, batch):
        """Sets the current choices (shared options and per-class overrides) on another batch of the
        same members, e.g. a fresh one for a measurement pass."""
        batch.set_options(**dict(self.base, **self.shared[self.shared_best]))
        for ci, (_, idx) in enumerate(self.classes):
            for i in idx:
                batch.set_stream_options(i, **self.choice[ci])

    def class_choices(self):
        return [{"shape": list(key), "streams": len(idx), **self.choice[ci]}
                for ci, (key, idx) in enumerate(self.classes)]

    def shared_medians(self):
        return [float(np.median(h)) if h else None for h in self.shared_hist]


def tune_batch(batch, frame_ptrs, n_steps, candidates=None, log=None):
    """Steps `batch` through frames 0..n_steps-1 (frame_ptrs(t) -> (image pointers, mask pointers)),
    using every step as a trial of the shared candidates in turn; leaves the best options set and
    returns them."""
    cands = list(candidates or SHARED_CANDIDATES)
    tuner = OnlineTuner(batch, cands, classes=[], period=0, max_period=0, window=max(1, n_steps // len(cands)),
                        log=log)
    rounds = n_steps // len(cands)
    for t in range(rounds * len(cands)):
        tuner.apply(*frame_ptrs(t))
    for t in range(rounds * len(cands), n_steps):
        batch.apply_device(*frame_ptrs(t))
    batch.synchronize()
    return batch.options()
