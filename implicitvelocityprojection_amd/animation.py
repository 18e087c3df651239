"""Keyframe animation of static rigid bodies, behind the interface of the reference's
``AnimationEngine`` (scripts/sim/utils/animation.py): ``build(scene_data)``, ``has_tracks()``,
``apply(t, dt)``, ``clear_kinematic()``.

Each scene object's ``"Animation": {"tracks": [...]}`` block becomes one ``Track`` per entry: two
keyframe channels (a translation vector and a rotation angle in degrees), a loop mode, a pivot and
a rotation axis.  The pose at time t is evaluated on the host; ``ivp_rigid_transform`` then
places the object's particles at ``pivot + T + R (x0 - pivot)`` on the GPU.

Poses must equal the reference's bit for bit (the particles of an animated boundary enter the
neighbour rows), so three pieces of arithmetic follow animation.py exactly: the vector channel
interpolates in float32 (``v0 + alpha (v1 - v0)``, alpha in float64 rounded once), the angle
channel in float64, and the rotation matrix forms every entry in float64 from the axis as stored
and rounds it to float32 once (animation.py:55-118).

Animated objects must be static rigid bodies (material rigid, not dynamic): their particles are
boundary particles whose positions change between steps.  An animated fluid or dynamic body would
move particles between the dynamic and static sets every step; it raises
``NotImplementedError``.
"""

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _native as N

_NAMED_AXES = {"x": (1.0, 0.0, 0.0), "y": (0.0, 1.0, 0.0), "z": (0.0, 0.0, 1.0)}


class Keyframes:
    """A piecewise-linear channel over [{"t": time, "value": v}, ...] (animation.py:55-86).

    Before the first key and after the last one the channel holds the end value.  In between, t
    belongs to the first listed segment [t_k, t_k+1] that contains it, so a time that falls on an
    interior key is evaluated at the end of the segment before it.  Vector channels hold float32
    triples; scalar channels keep the JSON numbers and interpolate in float64.  An empty channel
    reads as zero."""

    def __init__(self, keys, vector):
        self.vector = bool(vector)
        keys = list(keys or [])
        self.times = np.array([float(k["t"]) for k in keys], np.float64)
        if self.vector:
            self.values = np.array([k["value"] for k in keys], np.float32).reshape(-1, 3)
        else:
            self.values = [k["value"] for k in keys]

    def _zero(self):
        return np.zeros(3, np.float32) if self.vector else 0.0

    def _held(self, k):
        return self.values[k].copy() if self.vector else float(self.values[k])

    def at(self, t):
        if self.times.size == 0:
            return self._zero()
        if t <= self.times[0]:
            return self._held(0)
        if t >= self.times[-1]:
            return self._held(-1)
        contains = (self.times[:-1] <= t) & (t <= self.times[1:])
        if not contains.any():
            return self._held(-1)
        k = int(np.argmax(contains))
        alpha = (t - self.times[k]) / (self.times[k + 1] - self.times[k])
        v0, v1 = self.values[k], self.values[k + 1]
        if self.vector:
            return v0 + np.float32(alpha) * (v1 - v0)
        return v0 + alpha * (v1 - v0)


def local_time(t, mode, start, end):
    """The track's clock (animation.py:89-105): "none" or an empty interval leave t unchanged;
    otherwise the clock waits at `start` until then, after which "cycle" repeats [start, end) and
    "pingpong" runs it forwards then backwards (another mode keeps t)."""
    if mode == "none" or end <= start:
        return t
    if t < start:
        return start
    span = end - start
    if mode == "cycle":
        return start + (t - start) % span
    if mode == "pingpong":
        phase = (t - start) % (2.0 * span)
        return start + phase if phase < span else end - (phase - span)
    return t


def rotation_matrix(axis, angle_rad):
    """Rodrigues' rotation about a unit axis (animation.py:108-118):
    R = cos I + sin [a]x + (1 - cos) a a^T, the entries formed in float64 -- (1 - cos) a_i a_j as
    ((1 - cos) a_i) a_j with i <= j -- and rounded to float32 once."""
    c, s = np.cos(angle_rad), np.sin(angle_rad)
    a = np.asarray(axis).astype(np.float64)
    ta = (1.0 - c) * a
    outer = ta[:, None] * a[None, :]
    R = np.triu(outer) + np.triu(outer, 1).T
    R[np.diag_indices(3)] += c
    R += s * np.array([[0.0, -a[2], a[1]], [a[2], 0.0, -a[0]], [-a[1], a[0], 0.0]])
    return R.astype(np.float32)


def _unit_axis(spec):
    """A rotation axis given as "x" / "y" / "z" or a vector, normalised in float32."""
    v = np.array(_NAMED_AXES[spec.lower()] if isinstance(spec, str) else spec, np.float32)
    norm = np.linalg.norm(v)
    if norm > 1e-12:
        v /= norm
    return v


def _pivot(spec, obj, cfg):
    """A track's pivot (animation.py:121-130, 157-160): a vector as given; "world" reads the
    track's pivot_world (or pivotWorld); any other name is the object's rest centre of mass, the
    float32 mean of its voxelised or generated points.  Without those, the object's
    translation."""
    default = obj.get("translation", [0, 0, 0])
    if not isinstance(spec, str):
        return spec
    if spec == "world":
        return cfg.get("pivot_world", cfg.get("pivotWorld", default))
    pts = obj.get("voxelizedPoints", obj.get("generatedPositions"))
    if pts is None or len(pts) == 0:
        return default
    return np.asarray(pts, np.float32).mean(axis=0)


@dataclass
class Track:
    """One animation track of one scene object."""
    object_id: int
    translation: Keyframes
    angle_deg: Keyframes
    pivot: np.ndarray
    rotation_axis: np.ndarray
    loop_mode: str = "none"
    loop_start: float = 0.0
    loop_end: float = 0.0

    @classmethod
    def from_config(cls, object_id, obj, cfg):
        channels = cfg.get("channels", {})
        loop = cfg.get("loop", {})
        return cls(object_id=int(object_id),
                   translation=Keyframes(channels.get("translation"), vector=True),
                   angle_deg=Keyframes(channels.get("rotation_angle_deg"), vector=False),
                   pivot=np.array(_pivot(cfg.get("pivot", obj.get("translation", [0, 0, 0])), obj, cfg), np.float32),
                   rotation_axis=_unit_axis(cfg.get("rotation_axis", "y")),
                   loop_mode=loop.get("mode", "none"), loop_start=float(loop.get("start", 0.0)),
                   loop_end=float(loop.get("end", 0.0)))

    def pose(self, t):
        """(R, T) at time t (animation.py:174-180): x = pivot + T + R (x0 - pivot)."""
        tl = local_time(t, self.loop_mode, self.loop_start, self.loop_end)
        angle = float(self.angle_deg.at(tl))
        return rotation_matrix(self.rotation_axis, angle * np.pi / 180.0), self.translation.at(tl)


def parse_tracks(object_id, obj):
    """The tracks of one scene object ([] without an "Animation" block)."""
    return [Track.from_config(object_id, obj, cfg) for cfg in (obj.get("Animation") or {}).get("tracks", [])]


class AnimationEngine:
    """Tracks of every object of a scene, applied to the particle store at each step."""

    def __init__(self, ps):
        self.ps = ps
        self.tracks = []

    def build(self, scene_data):
        self.tracks = [tr for oid, obj in scene_data.get("object_map", {}).items() for tr in parse_tracks(oid, obj)]

    def has_tracks(self):
        return bool(self.tracks)

    def _check_static_rigid(self, sel, object_id):
        n = self.ps.particle_num
        if bool(((self.ps.is_dynamic[:n] != 0) & sel).any()) or bool(((self.ps.material[:n] != 0) & sel).any()):
            raise NotImplementedError(f"animated object {object_id} is not a static rigid body; only static "
                                      "rigid bodies can be animated in this build")

    def apply(self, t, dt):
        """Place every animated object at its pose at time t; True when some particle moved."""
        ps = self.ps
        n = ps.particle_num
        lib = N.load()
        stream = ctypes.c_void_p(torch.cuda.current_stream(ps.device).cuda_stream)
        vec3 = lambda a: (ctypes.c_float * 3)(*np.asarray(a, np.float32).reshape(3).tolist())
        moved = False
        for tr in self.tracks:
            sel = ps.object_id[:n] == tr.object_id
            if not bool(sel.any()):
                continue
            self._check_static_rigid(sel, tr.object_id)
            R, T = tr.pose(t)
            N.check(lib.ivp_rigid_transform(ctypes.c_void_p(ps.x0.data_ptr()), ctypes.c_void_p(ps.x.data_ptr()),
                                            ctypes.c_void_p(ps.object_id.data_ptr()), n, int(tr.object_id),
                                            vec3(tr.pivot), (ctypes.c_float * 9)(*R.reshape(9).tolist()), vec3(T),
                                            stream), "ivp_rigid_transform")
            ps.is_kinematic[:n][sel] = 1
            moved = True
        return moved

    def clear_kinematic(self):
        """The kinematic flags back to 0 after the step (animated bodies are static rigid here, so
        is_dynamic is unchanged)."""
        self.ps.is_kinematic.zero_()
