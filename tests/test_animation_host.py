"""Host logic of keyframe animation (animation.py:55-180 of the reference): keyframe
interpolation, loop modes, the axis-angle matrix and the pivots, pinned by closed forms and by
the reference's float32 / float64 arithmetic written out term by term."""

import math

import numpy as np
import pytest

from implicitvelocityprojection_amd.animation import Keyframes, local_time, parse_tracks, rotation_matrix


def _lerp_vec3(keys, t):
    return Keyframes(keys, vector=True).at(t)


def _lerp_scalar(keys, t):
    return Keyframes(keys, vector=False).at(t)

KEYS = [{"t": 1.0, "value": [0.0, 0.0, 0.0]}, {"t": 2.0, "value": [1.0, -2.0, 4.0]},
        {"t": 4.0, "value": [1.0, 0.0, 0.0]}]


def test_keyframe_interpolation():
    np.testing.assert_array_equal(_lerp_vec3(KEYS, 0.0), [0, 0, 0])          # before the first key
    np.testing.assert_array_equal(_lerp_vec3(KEYS, 9.0), [1, 0, 0])          # after the last
    np.testing.assert_allclose(_lerp_vec3(KEYS, 1.5), [0.5, -1.0, 2.0], rtol=1e-7)
    np.testing.assert_allclose(_lerp_vec3(KEYS, 3.0), [1.0, -1.0, 2.0], rtol=1e-7)
    assert _lerp_vec3(KEYS, 1.5).dtype == np.float32
    np.testing.assert_array_equal(_lerp_vec3([], 1.0), [0, 0, 0])
    sk = [{"t": 0.0, "value": 0.0}, {"t": 2.0, "value": 90.0}]
    assert _lerp_scalar(sk, 0.5) == pytest.approx(22.5) and _lerp_scalar(sk, 5.0) == 90.0 and _lerp_scalar([], 1) == 0


def test_loop_modes():
    assert local_time(7.3, "none", 1.0, 3.0) == 7.3
    assert local_time(0.5, "cycle", 1.0, 3.0) == 1.0                      # before the loop starts
    assert local_time(4.5, "cycle", 1.0, 3.0) == pytest.approx(2.5)       # 1 + (3.5 mod 2)
    assert local_time(4.5, "pingpong", 1.0, 3.0) == pytest.approx(1.5)    # phase 3.5: going back
    assert local_time(2.5, "pingpong", 1.0, 3.0) == pytest.approx(2.5)
    assert local_time(2.5, "cycle", 3.0, 3.0) == 2.5                      # empty loop: no remap


def test_rotation_matrix():
    rng = np.random.default_rng(0)
    for _ in range(10):
        a = rng.normal(size=3)
        a /= np.linalg.norm(a)
        ang = rng.uniform(-math.pi, math.pi)
        R = rotation_matrix(a, ang).astype(np.float64)
        np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-6)
        assert np.linalg.det(R) == pytest.approx(1.0, abs=1e-6)
        np.testing.assert_allclose(R @ a, a, atol=1e-6)                    # the axis stays
        assert np.trace(R) == pytest.approx(1 + 2 * math.cos(ang), abs=1e-6)
    # +90 degrees about y: x -> -z
    np.testing.assert_allclose(rotation_matrix([0, 1, 0], math.pi / 2) @ [1, 0, 0], [0, 0, -1], atol=1e-7)


def test_tracks_and_pivots():
    pts = np.array([[0.0, 0.0, 0.0], [2.0, 0.0, 0.0], [1.0, 3.0, 0.0]], np.float32)
    obj = {"translation": [5.0, 5.0, 5.0], "voxelizedPoints": pts,
           "Animation": {"tracks": [
               {"channels": {"translation": KEYS}, "pivot": "object_restCOM", "rotation_axis": "z"},
               {"channels": {"rotation_angle_deg": [{"t": 0.0, "value": 0.0}, {"t": 1.0, "value": 180.0}]},
                "pivot": "world", "pivot_world": [1.0, 2.0, 3.0], "rotation_axis": [0, 0, 2],
                "loop": {"mode": "pingpong", "start": 0.0, "end": 1.0}},
               {"channels": {}}]}}
    tr = parse_tracks(7, obj)
    assert [t.object_id for t in tr] == [7, 7, 7]
    np.testing.assert_allclose(tr[0].pivot, [1.0, 1.0, 0.0])               # mean rest position
    np.testing.assert_array_equal(tr[1].pivot, [1.0, 2.0, 3.0])
    np.testing.assert_array_equal(tr[2].pivot, [5.0, 5.0, 5.0])            # default: the translation
    np.testing.assert_allclose(tr[1].rotation_axis, [0, 0, 1])            # normalised
    R, T = tr[1].pose(1.5)                                      # pingpong: back at 0.5 -> 90 deg
    np.testing.assert_allclose(R @ [1, 0, 0], [0, 1, 0], atol=1e-7)
    np.testing.assert_array_equal(T, [0, 0, 0])
    assert parse_tracks(1, {"translation": [0, 0, 0]}) == []


def test_arithmetic_follows_the_reference_bitwise():
    """The float32 vector interpolation (v0 + alpha (v1 - v0), alpha rounded once), the float64
    scalar interpolation and every Rodrigues entry formed in float64 and rounded once."""
    rng = np.random.default_rng(3)
    keys = [{"t": 0.0, "value": [0.1, 0.2, 0.3]}, {"t": 0.7, "value": [1.3, -2.1, 0.9]},
            {"t": 2.0, "value": [0.4, 0.4, -3.3]}]
    for t in rng.uniform(-0.5, 2.5, 200).tolist() + [0.7, 2.0, 0.0]:
        got = _lerp_vec3(keys, t)
        if t <= 0.0 or t >= 2.0:
            continue
        k = 0 if t <= 0.7 else 1
        t0, t1 = keys[k]["t"], keys[k + 1]["t"]
        v0 = np.array(keys[k]["value"], np.float32)
        v1 = np.array(keys[k + 1]["value"], np.float32)
        alpha = np.float32((t - t0) / (t1 - t0))
        want = (v0 + alpha * (v1 - v0)).astype(np.float32)
        assert got.dtype == np.float32 and np.array_equal(got.view(np.uint32), want.view(np.uint32)), t
    sk = [{"t": 0, "value": 10}, {"t": 3, "value": -35.5}]
    for t in rng.uniform(0.0, 3.0, 50).tolist():
        assert _lerp_scalar(sk, t) == 10 + ((t - 0) / (3 - 0)) * (-35.5 - 10)
    for _ in range(50):
        a = rng.normal(size=3).astype(np.float32)
        a /= np.linalg.norm(a)
        ang = float(rng.uniform(-math.pi, math.pi))
        c, s_, t_ = math.cos(ang), math.sin(ang), 1.0 - math.cos(ang)
        x, y, z = (float(v) for v in a)
        want = np.array([[t_ * x * x + c, t_ * x * y - s_ * z, t_ * x * z + s_ * y],
                         [t_ * x * y + s_ * z, t_ * y * y + c, t_ * y * z - s_ * x],
                         [t_ * x * z - s_ * y, t_ * y * z + s_ * x, t_ * z * z + c]]).astype(np.float32)
        got = rotation_matrix(a, ang)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
