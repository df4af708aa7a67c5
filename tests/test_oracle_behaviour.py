"""Behaviour of the CPU oracle on scenes whose answer is known without running it (CPU only): a static
scene stays background, an object far from the background colour is foreground, results are
deterministic for a key and depend on it, lr = inf freezes the sample model, and the reference's
contract errors are reported."""
import numpy as np
import pytest

from oracle_lib import LOBSTER, SUBSENSE, Oracle


def _flat(h=60, w=80, c=3, v=(90, 120, 150)):
    img = np.empty((h, w, c), np.uint8) if c == 3 else np.empty((h, w), np.uint8)
    img[...] = v if c == 3 else v[0]
    return img


@pytest.mark.parametrize("algo,c", [(SUBSENSE, 3), (SUBSENSE, 1), (LOBSTER, 3), (LOBSTER, 1)])
def test_static_scene_stays_background(algo, c):
    o = Oracle(algo)
    f = _flat(c=c)
    o.initialize(f)
    for _ in range(8):
        assert not o.apply(f).any()


@pytest.mark.parametrize("algo", [SUBSENSE, LOBSTER])
def test_distinct_object_is_foreground(algo):
    o = Oracle(algo)
    bg = _flat()
    o.initialize(bg)
    for _ in range(3):
        o.apply(bg)
    f = bg.copy()
    f[20:40, 30:55] = (230, 20, 40)
    m = o.apply(f)
    assert (m[24:36, 34:51] == 255).all()          # object interior
    assert not m[:, :20].any() and not m[48:].any()  # far from the object


def test_deterministic_and_keyed():
    rng = np.random.default_rng(0)
    base = rng.integers(60, 200, (48, 64, 3), dtype=np.uint8)
    frames = [np.clip(base.astype(int) + rng.integers(-12, 13, base.shape), 0, 255).astype(np.uint8)
              for _ in range(6)]

    def run(**kw):
        o = Oracle(SUBSENSE, **kw)
        o.initialize(frames[0])
        masks = [o.apply(f) for f in frames]
        return masks, o.state("samples_color")

    m1, s1 = run()
    m2, s2 = run()
    assert all(np.array_equal(a, b) for a, b in zip(m1, m2)) and np.array_equal(s1, s2)
    _, s3 = run(stream_id=5)
    assert not np.array_equal(s1, s3)


def test_infinite_learning_rate_freezes_samples():
    rng = np.random.default_rng(1)
    f0 = rng.integers(0, 256, (40, 56, 3), dtype=np.uint8)
    o = Oracle(SUBSENSE)
    o.initialize(f0)
    before = o.state("samples_color").copy(), o.state("samples_desc").copy()
    for _ in range(4):
        o.apply(rng.integers(0, 256, f0.shape, dtype=np.uint8), float("inf"))
    assert np.array_equal(o.state("samples_color"), before[0])
    assert np.array_equal(o.state("samples_desc"), before[1])


def test_contract_errors():
    o = Oracle(SUBSENSE)
    with pytest.raises(RuntimeError):
        o.apply(_flat())                      # not initialized
    lo = Oracle(LOBSTER)
    lo.initialize(_flat())
    with pytest.raises(RuntimeError):
        lo.apply(_flat(), 0.0)                # LOBSTER needs a positive learning rate
    with pytest.raises(RuntimeError):
        lo.apply(_flat(h=30))                 # size mismatch
