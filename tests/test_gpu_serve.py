"""GPU tests of the serving loop: streams admitted against a device-memory budget, frames submitted
in irregular bursts and batched with whatever else is waiting (continuous batching), contexts reused
after a stream closes. Every frame's mask must equal the sequential oracle's for that stream alone."""
import numpy as np
import pytest

from litiv_amd import Server, synth
from oracle_lib import SUBSENSE, Oracle

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


class Stream:
    """Frames on the device, one device mask per frame, and the oracle's masks for the same stream."""

    def __init__(self, torch, sid, w, h, c, n):
        cfg = synth.preset(synth.QVGA, sid, width=w, height=h, channels=c, n_objects=2, obj_min=6, obj_max=w // 5)
        seq = synth.Sequence(cfg)
        self.frames = [seq.frame(t) for t in range(n)]
        self.d = [torch.from_numpy(np.ascontiguousarray(f)).cuda() for f in self.frames]
        self.m = [torch.zeros((h, w), dtype=torch.uint8, device="cuda") for _ in range(n)]
        self.params = {"stream_id": 100 + sid, "rng_seed": 0xB65E + sid}
        o = Oracle(SUBSENSE, **self.params)
        o.initialize(self.frames[0])
        self.expected = [o.apply(f) for f in self.frames]
        self.w, self.h, self.c = w, h, c
        self.next = 0


def _open(srv, s):
    return srv.open(s.d[0].data_ptr(), s.w, s.h, s.c, **s.params)


@pytest.mark.parametrize("groups", [1, 3])
def test_serving_mix_matches_each_stream_alone(groups):
    torch = _torch()
    shapes = [(96, 72, 3), (160, 120, 1), (128, 96, 3), (64, 48, 3), (64, 48, 3)]
    streams = [Stream(torch, i, w, h, c, 14) for i, (w, h, c) in enumerate(shapes)]
    torch.cuda.synchronize()
    sizes = [Server.model_bytes(s.w, s.h, s.c) for s in streams]
    # room for the first three models only
    srv = Server(hbm_budget=sizes[0] + sizes[1] + sizes[2] + sizes[3] // 2, max_inflight=2, groups=groups)
    ids = {}
    for i in range(3):
        ids[i] = _open(srv, streams[i])
        assert ids[i] is not None
    assert _open(srv, streams[3]) is None  # does not fit: waits
    rng = np.random.default_rng(7)
    checked = {i: 0 for i in range(len(streams))}
    by_id = {}

    def submit_some(i, k):
        s = streams[i]
        for _ in range(k):
            if s.next >= len(s.frames):
                return
            t = s.next
            srv.submit(ids[i], s.d[t].data_ptr(), s.m[t].data_ptr(), 0.0, tag=i * 1000 + t)
            s.next += 1

    def collect():
        srv.pump()
        for tag, sid, status in srv.poll():
            assert status == 0
            i, t = int(tag) // 1000, int(tag) % 1000
            assert by_id.get(sid, i) == i
            torch.cuda.synchronize()
            got = streams[i].m[t].cpu().numpy()
            assert np.array_equal(got, streams[i].expected[t]), "stream %d frame %d: %d px differ" % (
                i, t, int((got != streams[i].expected[t]).sum()))
            checked[i] += 1

    for i in range(3):
        by_id[ids[i]] = i
    # frame 0 was the initialization frame; apply() sequences start at frame 0 as in the oracle
    while any(streams[i].next < len(streams[i].frames) for i in range(3)):
        for i in range(3):
            submit_some(i, int(rng.integers(0, 4)))
        collect()
    srv.drain()
    collect()
    assert all(checked[i] == len(streams[i].frames) for i in range(3))
    st = srv.stats()
    assert st["frames_done"] == sum(checked.values()) and st["steps"] < st["frames_done"]  # frames were batched
    # stream 0 leaves; stream 3 (other geometry) fits once stream 0's idle context is evicted;
    # stream 4 (stream 3's geometry) then reuses stream 3's context, with nothing stale in it
    srv.close(ids[0])
    srv.drain()
    ids[3] = _open(srv, streams[3])
    assert ids[3] is not None
    by_id = {ids[1]: 1, ids[2]: 2, ids[3]: 3}
    for t in range(len(streams[3].frames)):
        submit_some(3, 1)
        if t % 3 == 0:
            collect()
    srv.drain()
    collect()
    assert checked[3] == len(streams[3].frames)
    srv.close(ids[3])
    srv.drain()
    assert srv.stats()["pooled_contexts"] == 1
    ids[4] = _open(srv, streams[4])
    assert srv.stats()["pooled_contexts"] == 0
    by_id = {ids[1]: 1, ids[2]: 2, ids[4]: 4}
    submit_some(4, 14)
    srv.drain()
    collect()
    assert checked[4] == len(streams[4].frames)
    # state of a served stream equals the oracle's after its sequence
    o = Oracle(SUBSENSE, **streams[4].params)
    o.initialize(streams[4].frames[0])
    for f in streams[4].frames:
        o.apply(f)
    member = srv.member(ids[4])
    for name in ("samples_color", "samples_desc", "R", "T", "last_fg"):
        assert np.array_equal(member.state(name), o.state(name)), name
    srv.close_server()


def test_server_rejects_bad_calls():
    from litiv_amd.bgs import LvbgsError
    torch = _torch()
    srv = Server(hbm_budget=1 << 30)
    f = torch.zeros((48, 64, 3), dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    with pytest.raises(LvbgsError):
        srv.submit(3, f.data_ptr())  # unknown stream
    sid = srv.open(f.data_ptr(), 64, 48, 3)
    srv.close(sid)
    with pytest.raises(LvbgsError):
        srv.submit(sid, f.data_ptr())  # closing stream
    assert Server(hbm_budget=1024).open(f.data_ptr(), 64, 48, 3) is None
    srv.drain()
    srv.close_server()
