"""Host-side logic of the Python surface that runs before any GPU call:
mixed-frame state lists, propagate_n_body_long's planet states and batch
size.  kete handles every state of propagate_n_body on its own
(src/kete/rust/propagation.rs:105-140), so a list mixing frames or centers
is split into one batch per (frame, center) and reassembled in order; the
planet states of propagate_n_body_long are rotated to equatorial when given
in the ecliptic frame (kete uses them as given, propagation/mod.rs:186-190,
i.e. assumes equatorial); batches are limited to 55 objects (kete: any)."""
import numpy as np
import pytest

import kete_amd as ka

OBL = 0.40909280422232897


def states(frames_centers):
    out = []
    for k, (f, c) in enumerate(frames_centers):
        out.append(ka.State(f"obj{k}", 2460000.5, np.array([1.0 + k, 0.1, 0.2]),
                            np.array([0.0, 0.017, 0.001 * k]), f, c))
    return out


def test_mixed_frames_and_centers_grouped_in_order():
    E, Q = ka.Frames.Ecliptic, ka.Frames.Equatorial
    sts = states([(E, 10), (Q, 10), (E, 10), (E, 0), (Q, 10)])
    seen = []

    def fn(batch, idx):
        seen.append((batch.frame, batch.center_id, list(idx)))
        xyz = batch.xyz.copy()
        xyz[0] += 100.0                     # mark the result
        return ka.StateArray(batch.jd, xyz, batch.frame, batch.center_id, batch.desigs)

    out = ka._per_frame_group(sts, fn)
    assert seen == [(E, 10, [0, 2]), (Q, 10, [1, 4]), (E, 0, [3])]
    assert [s.desig for s in out] == [s.desig for s in sts]
    for a, b in zip(out, sts):
        assert a.frame == b.frame and a.center_id == b.center_id
        assert a.pos[0] == b.pos[0] + 100.0 and np.array_equal(a.vel, b.vel)


def test_uniform_list_is_one_batch():
    sts = states([(ka.Frames.Ecliptic, 10)] * 4)
    calls = []
    ka._per_frame_group(sts, lambda b, i: calls.append(len(b)) or b)
    assert calls == [4]


def test_planet_states_rotated_from_ecliptic():
    c, s = np.cos(OBL), np.sin(OBL)
    rng = np.random.default_rng(1)
    xyz = rng.normal(size=(6, 9))
    ecl = ka.StateArray(np.full(9, 2460000.5), xyz, ka.Frames.Ecliptic, 10)
    pl = ka._planet_array(ecl, 2460000.5)
    for k in range(9):
        for o in (0, 3):
            x, y, z = xyz[o:o + 3, k]
            assert np.allclose(pl[k, o:o + 3], [x, c * y - s * z, s * y + c * z], rtol=0,
                               atol=1e-15)
    equ = ka.StateArray(np.full(9, 2460000.5), xyz, ka.Frames.Equatorial, 10)
    assert np.array_equal(ka._planet_array(equ, 2460000.5), xyz.T)
    with pytest.raises(ValueError):
        ka._planet_array(ka.StateArray(np.full(8, 2460000.5), xyz[:, :8]), 2460000.5)
    with pytest.raises(ValueError):
        ka._planet_array(equ, 2460001.5)


@pytest.mark.parametrize("bad", [0, 56, 100])
def test_long_batch_size_limit(bad):
    sts = states([(ka.Frames.Ecliptic, 10)] * 3)
    with pytest.raises(ValueError, match="batch_size"):
        ka.propagate_n_body_long(sts, 2460010.5, batch_size=bad)
