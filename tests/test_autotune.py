"""CPU tests of the launch-option tuner's control logic against a simulated batch whose step time
depends on the options in a known way (no GPU: the batch is a stand-in with the same methods)."""
import random

from litiv_amd.autotune import CLASS_CHOICES, OnlineTuner, tune_batch


class SimBatch:
    """Step time = 8 ms + penalties: fetch depth 2 costs 0.2 ms, 2 / 3 / 4 groups save 0.15 / 0.2 / 0.3 ms;
    class 0's seq_limit is best at 8,
    class 1's at 16 (0.12 ms, 1.5%, per step of distance in the choice list, above the tuner's 1%
    switching margin); cont_px_per_block has no effect; Gaussian noise of 0.01 ms."""

    def __init__(self, seed=0):
        self.o = {"use_graph": 1, "seq_limit": 12, "cont_blocks": 1024, "cont_px_per_block": 2048,
                  "fixpoint_launches": 2, "fixpoint_blocks": 256, "fix_px_per_block": 8192, "fetch_depth": 4,
                  "groups": 1}
        self.so = {i: {"seq_limit": 0, "cont_px_per_block": 0, "fix_px_per_block": 0} for i in range(4)}
        self.rng = random.Random(seed)
        self.timing = ""
        self.graph_steps = self.direct_steps = 0
        self.last = None

    def options(self):
        return dict(self.o)

    def set_options(self, **kw):
        assert set(kw) <= set(self.o)
        self.o.update(kw)

    def set_stream_options(self, i, **kw):
        self.so[i].update(kw)

    def shape_classes(self):
        return [((1080, 1920, 3), [0, 1]), ((240, 320, 3), [2, 3])]

    def set_timing(self, stage):
        self.timing = stage

    def apply_device(self, images, masks):
        if self.o["use_graph"]:
            self.graph_steps += 1
        else:
            self.direct_steps += 1
        lims = CLASS_CHOICES["seq_limit"]
        t = 8.0 + (0.2 if self.o["fetch_depth"] == 2 else 0.0) - {1: 0.0, 2: 0.15, 3: 0.2, 4: 0.3}[self.o["groups"]]
        for members, best in (((0, 1), 8), ((2, 3), 16)):
            sl = self.so[members[0]]["seq_limit"] or self.o["seq_limit"]
            t += 0.12 * abs(lims.index(sl) - lims.index(best))
        self.last = t + self.rng.gauss(0.0, 0.01)

    def take_timing(self):
        assert self.timing == "step"
        return self.last, 1

    def synchronize(self):
        pass


def test_online_tuner_finds_per_class_optimum_and_backs_off():
    b = SimBatch()
    t = OnlineTuner(b, period=8, max_period=128)
    for _ in range(3000):
        t.apply(None, None)
    ch = {tuple(c["shape"]): c["seq_limit"] for c in t.class_choices()}
    assert ch[(1080, 1920, 3)] == 8 and ch[(240, 320, 3)] == 16
    assert b.so[0]["seq_limit"] == 8 and b.so[2]["seq_limit"] == 16  # applied to the members
    assert b.o["fetch_depth"] == 4 and b.o["groups"] == 4 and b.o["use_graph"] == 1  # shared best, graph on
    assert t.period == 128                                           # settled: trials back off
    assert b.direct_steps < 0.1 * (b.direct_steps + b.graph_steps)  # trials are a small share of steps


def test_apply_choices_copies_to_another_batch():
    b = SimBatch(1)
    t = OnlineTuner(b, period=8)
    for _ in range(1500):
        t.apply(None, None)
    other = SimBatch(2)
    t.apply_choices(other)
    assert other.so == b.so and other.o["fetch_depth"] == b.o["fetch_depth"]


def test_tune_batch_picks_the_fastest_shared_candidate():
    b = SimBatch(3)
    opts = tune_batch(b, lambda k: (None, None), 30)
    assert opts["fetch_depth"] == 4 and opts["use_graph"] == 1
