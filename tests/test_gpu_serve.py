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


def _serve_oracle(kind, k, n, params):
    seq = synth.Sequence(synth.preset(kind, k))
    o = Oracle(SUBSENSE, **params)
    o.initialize(seq.frame(0))
    return [o.apply(seq.frame(t)) for t in range(n)]


def test_serving_full_size_resolutions_with_admission_churn():
    """Streams of the serving workload's four resolutions (320x240, 640x480, 1280x720 twice,
    1920x1080 RGB) at full size. The budget holds four of the five models: the fifth waits, is
    admitted when the first stream closes (evicting its pooled context), and every frame of every
    stream is checked against that stream's oracle (run in host threads)."""
    from concurrent.futures import ThreadPoolExecutor
    torch = _torch()
    want = [(320, 240), (640, 480), (1280, 720), (1280, 720), (1920, 1080)]
    picks, k = [], 0
    while len(picks) < len(want):
        cfg = synth.preset(synth.SERVE, k)
        if (cfg.width, cfg.height) == want[len(picks)]:
            picks.append(k)
        k += 1
    n = 8
    params = [{"stream_id": 2000 + k} for k in picks]
    pool = ThreadPoolExecutor(max_workers=len(picks))
    futs = [pool.submit(_serve_oracle, synth.SERVE, k, n, p) for k, p in zip(picks, params)]
    seqs = [synth.Sequence(synth.preset(synth.SERVE, k)) for k in picks]
    bufs = [torch.empty(n * s.frame_bytes, dtype=torch.uint8, device="cuda") for s in seqs]
    masks = [torch.empty((n, s.cfg.height, s.cfg.width), dtype=torch.uint8, device="cuda") for s in seqs]
    torch.cuda.synchronize()
    for s, b in zip(seqs, bufs):
        s.frames_device(0, n, b.data_ptr())
    torch.cuda.synchronize()
    sizes = [Server.model_bytes(s.cfg.width, s.cfg.height, 3) for s in seqs]
    # order of opening: the 1080p stream first, then 320x240, 640x480, 1280x720; the second 1280x720
    # stream does not fit until the 320x240 one has left and its pooled context is evicted
    order = [4, 0, 1, 2, 3]
    budget = sizes[4] + sizes[1] + sizes[2] + sizes[3] + sizes[0] // 2
    srv = Server(hbm_budget=budget, max_inflight=2, groups=2, seq_limit=8)
    ids = {}
    for i in order[:4]:
        ids[i] = srv.open(bufs[i].data_ptr(), seqs[i].cfg.width, seqs[i].cfg.height, 3, **params[i])
        assert ids[i] is not None
    assert srv.open(bufs[3].data_ptr(), seqs[3].cfg.width, seqs[3].cfg.height, 3, **params[3]) is None
    nxt = {i: 0 for i in range(len(picks))}
    status = {}

    def submit(i, count):
        for _ in range(count):
            t = nxt[i]
            if t >= n:
                return
            srv.submit(ids[i], bufs[i].data_ptr() + t * seqs[i].frame_bytes, masks[i][t].data_ptr(), 0.0,
                       tag=i * 1000 + t)
            nxt[i] += 1

    def collect():
        srv.pump()
        for tag, sid, st in srv.poll():
            status[int(tag)] = int(st)

    rng = np.random.default_rng(3)
    while nxt[0] < n:
        for i in order[:4]:
            submit(i, int(rng.integers(1, 3)))
        collect()
    srv.close(ids[0])
    srv.drain()
    collect()
    ids[3] = srv.open(bufs[3].data_ptr(), seqs[3].cfg.width, seqs[3].cfg.height, 3, **params[3])
    assert ids[3] is not None
    while any(nxt[i] < n for i in (1, 2, 3, 4)):
        for i in (1, 2, 3, 4):
            submit(i, int(rng.integers(0, 3)))
        collect()
    srv.drain()
    collect()
    assert len(status) == n * len(picks) and all(v == 0 for v in status.values())
    torch.cuda.synchronize()
    for i in range(len(picks)):
        exp = futs[i].result()
        got = masks[i].cpu().numpy()
        for t in range(n):
            assert np.array_equal(got[t], exp[t]), "%dx%d stream frame %d: %d px differ" % (
                seqs[i].cfg.width, seqs[i].cfg.height, t, int((got[t] != exp[t]).sum()))
    pool.shutdown()
    srv.close_server()


def test_device_error_is_reported_for_one_stream_only():
    """The fix-point cap set to 0 from the host (no full-grid fix-point launches): a stream whose raster-
    order fix point has work stops with LVBGS_ERR_DEVICE (every kernel exits normally); a flat, static
    stream in the same steps (every sample equals the frame, every pixel is background in every frame,
    so no decision ever differs from its guess and the fix point has no work) keeps its exact results.
    The failed stream's later frames complete with the same status, further submits are refused, and
    after close (which acknowledges the failure) its pooled context serves a new stream exactly."""
    from litiv_amd.bgs import ERR_DEVICE, LvbgsError
    torch = _torch()
    n = 10
    moving = Stream(torch, 0, 96, 72, 3, n)
    still = Stream(torch, 1, 80, 64, 3, n)
    f0 = np.empty((64, 80, 3), np.uint8)
    f0[...] = (90, 120, 150)
    still.frames = [f0] * n
    still.d = [torch.from_numpy(f0).cuda()] * n
    o = Oracle(SUBSENSE, **still.params)
    o.initialize(f0)
    still.expected = [o.apply(f0) for _ in range(n)]
    torch.cuda.synchronize()
    srv = Server(hbm_budget=1 << 30, max_inflight=2, fixpoint_launches=0)
    srv.debug_fixpoint_cap(0)
    sm, ss = _open(srv, moving), _open(srv, still)
    for t in range(n):
        srv.submit(sm, moving.d[t].data_ptr(), moving.m[t].data_ptr(), 0.0, tag=t)
        srv.submit(ss, still.d[t].data_ptr(), still.m[t].data_ptr(), 0.0, tag=100 + t)
    srv.drain()
    done = srv.poll()
    st_m = {int(tag): int(s) for tag, sid, s in done if sid == sm}
    st_s = {int(tag) - 100: int(s) for tag, sid, s in done if sid == ss}
    assert sorted(st_m) == list(range(n)) and sorted(st_s) == list(range(n))
    assert all(v == 0 for v in st_s.values())
    bad = [t for t in range(n) if st_m[t] != 0]
    assert bad, "the moving stream never needed the fix point"
    assert all(st_m[t] == ERR_DEVICE for t in bad) and bad == list(range(bad[0], n))
    for t in range(bad[0]):
        assert np.array_equal(moving.m[t].cpu().numpy(), moving.expected[t]), t
    for t in range(n):
        assert np.array_equal(still.m[t].cpu().numpy(), still.expected[t]), t
        assert not still.expected[t].any()
    with pytest.raises(LvbgsError) as e:
        srv.submit(sm, moving.d[0].data_ptr(), moving.m[0].data_ptr())
    assert e.value.code == ERR_DEVICE
    srv.close(sm)
    srv.drain()
    assert srv.stats()["pooled_contexts"] == 1
    # default cap again: a new stream of the same geometry reuses the pooled context, flags cleared
    srv.debug_fixpoint_cap(1 << 16)
    again = Stream(torch, 0, 96, 72, 3, n)
    torch.cuda.synchronize()
    sa = _open(srv, again)
    assert srv.stats()["pooled_contexts"] == 0
    for t in range(n):
        srv.submit(sa, again.d[t].data_ptr(), again.m[t].data_ptr(), 0.0, tag=200 + t)
    srv.drain()
    assert all(int(s) == 0 for _, _, s in srv.poll())
    for t in range(n):
        assert np.array_equal(again.m[t].cpu().numpy(), again.expected[t]), t
    srv.close_server()


def test_host_ingest_matches_device_frames():
    """Frames submitted from host memory (lvbgs_server_submit_host) give the oracle's masks, mixed in
    the same steps with a stream fed device frames: from pinned memory and from pageable memory, both
    with padded rows (copied into the stream's ingest slot on the server's copy stream; a step takes
    a frame once its copy has completed). A stream whose slots are all busy refuses more until frames
    complete."""
    from litiv_amd.bgs import PinnedBuffer
    torch = _torch()
    n = 12
    hs = Stream(torch, 5, 128, 96, 3, n)       # pinned host memory
    ps = Stream(torch, 7, 80, 64, 3, n)        # pageable host memory
    ds = Stream(torch, 6, 96, 72, 3, n)        # device frames
    torch.cuda.synchronize()
    ring = 3
    srv = Server(hbm_budget=1 << 30, max_inflight=2, host_ring=ring)
    assert srv.stream_bytes(128, 96, 3) >= Server.model_bytes(128, 96, 3) + ring * 128 * 96 * 3
    row, stride = 128 * 3, 128 * 3 + 8        # padded rows
    fb = stride * 96
    pinned = PinnedBuffer(n * fb)
    for t in range(n):
        pinned.array[t * fb:(t + 1) * fb].reshape(96, stride)[:, :row] = hs.frames[t].reshape(96, row)
    prow, pstride = 80 * 3, 80 * 3 + 12
    pageable = [np.zeros((64, pstride), np.uint8) for _ in range(n)]
    for t in range(n):
        pageable[t][:, :prow] = ps.frames[t].reshape(64, prow)
    sh = srv.open_host(hs.frames[0], **hs.params)
    sp = srv.open_host(ps.frames[0], **ps.params)
    sd = _open(srv, ds)
    statuses, t_h, t_p, t_d = {}, 0, 0, 0
    # no step runs before a pump: the ring fills and the next host frame is refused
    for t_h in range(ring):
        assert srv.submit_host(sh, pinned.ptr + t_h * fb, hs.m[t_h].data_ptr(), 0.0, tag=t_h, stride=stride)
    t_h = ring
    assert not srv.submit_host(sh, pinned.ptr + t_h * fb, hs.m[t_h].data_ptr(), 0.0, tag=t_h, stride=stride)
    while t_h < n or t_p < n or t_d < n:
        if t_h < n and srv.submit_host(sh, pinned.ptr + t_h * fb, hs.m[t_h].data_ptr(), 0.0, tag=t_h, stride=stride):
            t_h += 1
        if t_p < n and srv.submit_host(sp, pageable[t_p].ctypes.data, ps.m[t_p].data_ptr(), 0.0, tag=200 + t_p,
                                       stride=pstride):
            t_p += 1
        if t_d < n:
            srv.submit(sd, ds.d[t_d].data_ptr(), ds.m[t_d].data_ptr(), 0.0, tag=100 + t_d)
            t_d += 1
        srv.pump()
        for tag, sid, st in srv.poll():
            statuses[int(tag)] = int(st)
    srv.drain()
    for tag, sid, st in srv.poll():
        statuses[int(tag)] = int(st)
    assert len(statuses) == 3 * n and all(v == 0 for v in statuses.values())
    torch.cuda.synchronize()
    for t in range(n):
        assert np.array_equal(hs.m[t].cpu().numpy(), hs.expected[t]), "pinned host frame %d" % t
        assert np.array_equal(ps.m[t].cpu().numpy(), ps.expected[t]), "pageable host frame %d" % t
        assert np.array_equal(ds.m[t].cpu().numpy(), ds.expected[t]), "device frame %d" % t
    srv.close_server()
    pinned.free()
