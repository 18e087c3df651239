"""Keyframe animation of rigid bodies, mirroring the reference's ``AnimationEngine``
(scripts/sim/utils/animation.py:1-229): tracks parsed from each scene object's ``"Animation"``
block, evaluated on the host (linear keyframes, loop modes, axis-angle rotation about a pivot),
applied on the GPU by ``ivp_rigid_transform`` to the object's rest positions.

Animated objects must be static rigid bodies (material rigid, not dynamic): their particles are
boundary particles whose positions change between steps.  An animated fluid or dynamic body
would move particles between the dynamic and static sets every step; it raises
``NotImplementedError``.
"""

import ctypes

import numpy as np
import torch

from . import _native as N


def _lerp_vec3(keys, t):
    """animation.py:55-70: linear interpolation of {"t", "value": [x, y, z]} keyframes, float32."""
    if not keys:
        return np.zeros(3, dtype=np.float32)
    if t <= keys[0]["t"]:
        return np.array(keys[0]["value"], dtype=np.float32)
    if t >= keys[-1]["t"]:
        return np.array(keys[-1]["value"], dtype=np.float32)
    for i in range(len(keys) - 1):
        t0, t1 = keys[i]["t"], keys[i + 1]["t"]
        if t0 <= t <= t1:
            alpha = (t - t0) / (t1 - t0)
            v0 = np.array(keys[i]["value"], dtype=np.float32)
            v1 = np.array(keys[i + 1]["value"], dtype=np.float32)
            return v0 + alpha * (v1 - v0)
    return np.array(keys[-1]["value"], dtype=np.float32)


def _lerp_scalar(keys, t):
    """animation.py:73-86"""
    if not keys:
        return 0.0
    if t <= keys[0]["t"]:
        return float(keys[0]["value"])
    if t >= keys[-1]["t"]:
        return float(keys[-1]["value"])
    for i in range(len(keys) - 1):
        t0, t1 = keys[i]["t"], keys[i + 1]["t"]
        if t0 <= t <= t1:
            alpha = (t - t0) / (t1 - t0)
            return keys[i]["value"] + alpha * (keys[i + 1]["value"] - keys[i]["value"])
    return float(keys[-1]["value"])


def remap_time(t, loop_mode, loop_start, loop_end):
    """animation.py:89-105: "none", "cycle" or "pingpong" over [loop_start, loop_end]."""
    if loop_mode == "none" or loop_end <= loop_start:
        return t
    period = loop_end - loop_start
    if t < loop_start:
        return loop_start
    elapsed = t - loop_start
    if loop_mode == "cycle":
        return loop_start + (elapsed % period)
    if loop_mode == "pingpong":
        phase = elapsed % (2.0 * period)
        return loop_start + phase if phase < period else loop_end - (phase - period)
    return t


def axis_angle_matrix(axis, angle_rad):
    """animation.py:108-118: Rodrigues' rotation, float32."""
    c, s = np.cos(angle_rad), np.sin(angle_rad)
    t = 1.0 - c
    x, y, z = axis
    return np.array([[t * x * x + c, t * x * y - s * z, t * x * z + s * y],
                     [t * x * y + s * z, t * y * y + c, t * y * z - s * x],
                     [t * x * z - s * y, t * y * z + s * x, t * z * z + c]], dtype=np.float32)


def _named_pivot(name, obj, track_cfg):
    """animation.py:121-130: "world" -> the track's pivot_world; otherwise the mean rest position
    of the object's particles; the object's translation when it has none."""
    fallback = obj.get("translation", [0, 0, 0])
    if name == "world":
        return track_cfg.get("pivot_world", track_cfg.get("pivotWorld", fallback))
    pts = obj.get("voxelizedPoints", obj.get("generatedPositions"))
    if pts is not None and len(pts) > 0:
        return np.asarray(pts, dtype=np.float32).mean(axis=0)
    return fallback


def _axis(axis_cfg):
    if isinstance(axis_cfg, str):
        return {"x": [1, 0, 0], "y": [0, 1, 0], "z": [0, 0, 1]}[axis_cfg.lower()]
    return list(axis_cfg)


class AnimationTrack:
    def __init__(self, object_id, channels, pivot, rotation_axis, loop_mode="none", loop_start=0.0, loop_end=0.0):
        self.object_id = object_id
        self.channels = channels
        self.pivot = np.array(pivot, dtype=np.float32)
        self.rotation_axis = np.array(rotation_axis, dtype=np.float32)
        n = np.linalg.norm(self.rotation_axis)
        if n > 1e-12:
            self.rotation_axis /= n
        self.loop_mode = loop_mode
        self.loop_start = loop_start
        self.loop_end = loop_end


def parse_tracks(obj_id, obj):
    """animation.py:150-171: the tracks of one scene object ([] without an "Animation" block)."""
    tracks = []
    for tc in (obj.get("Animation") or {}).get("tracks", []):
        loop = tc.get("loop", {})
        pivot = tc.get("pivot", obj.get("translation", [0, 0, 0]))
        if isinstance(pivot, str):
            pivot = _named_pivot(pivot, obj, tc)
        tracks.append(AnimationTrack(int(obj_id), tc.get("channels", {}), pivot, _axis(tc.get("rotation_axis", "y")),
                                     loop.get("mode", "none"), float(loop.get("start", 0.0)),
                                     float(loop.get("end", 0.0))))
    return tracks


def track_transform(track, t):
    """animation.py:174-180: (R, translation) at time t."""
    tl = remap_time(t, track.loop_mode, track.loop_start, track.loop_end)
    trans = _lerp_vec3(track.channels.get("translation", []), tl)
    angle = _lerp_scalar(track.channels.get("rotation_angle_deg", []), tl)
    return axis_angle_matrix(track.rotation_axis, float(angle) * np.pi / 180.0), trans


class AnimationEngine:
    def __init__(self, ps):
        self.ps = ps
        self.tracks = []

    def build(self, scene_data):
        self.tracks = []
        for obj_id, obj in scene_data.get("object_map", {}).items():
            self.tracks.extend(parse_tracks(obj_id, obj))

    def has_tracks(self):
        return len(self.tracks) > 0

    def apply(self, t, dt):
        """Move every animated object to its pose at time t: x = pivot + T + R (x0 - pivot).
        Returns True when some particle belongs to an animated object."""
        ps = self.ps
        n = ps.particle_num
        lib = N.load()
        stream = ctypes.c_void_p(torch.cuda.current_stream(ps.device).cuda_stream)
        moved = False
        for tr in self.tracks:
            sel = ps.object_id[:n] == tr.object_id
            if not bool(sel.any()):
                continue
            if bool(((ps.is_dynamic[:n] != 0) & sel).any()) or bool(((ps.material[:n] != 0) & sel).any()):
                raise NotImplementedError(f"animated object {tr.object_id} is not a static rigid body; only static "
                                          "rigid bodies can be animated in this build")
            R, T = track_transform(tr, t)
            f3 = lambda a: (ctypes.c_float * 3)(*[float(x) for x in np.asarray(a, np.float32).reshape(3)])
            Rc = (ctypes.c_float * 9)(*[float(x) for x in R.reshape(9)])
            N.check(lib.ivp_rigid_transform(ctypes.c_void_p(ps.x0.data_ptr()), ctypes.c_void_p(ps.x.data_ptr()),
                                            ctypes.c_void_p(ps.object_id.data_ptr()), n, int(tr.object_id),
                                            f3(tr.pivot), Rc, f3(T), stream), "ivp_rigid_transform")
            ps.is_kinematic[:n][sel] = 1
            moved = True
        return moved

    def clear_kinematic(self):
        """animation.py:36-52: kinematic flags back to 0 (animated bodies are static rigid here, so
        is_dynamic is unchanged)."""
        self.ps.is_kinematic.zero_()
