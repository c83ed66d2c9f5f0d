"""Scenario files in the reference simulator's JSON and tipsy formats.

from_json / to_json (src/NBody/Simulation/Simulation.cpp:154-228): a
"particles" array whose entries may hold "position", "velocity", "mass",
"acceleration" (each vector as {"x", "y", "z"}), "passive", "color" and
"sphere".  A particle with a mass is active (a source of gravity); a particle
gets an Acceleration (it is moved by the solver) if the file gives one, or
if "passive" is absent or true (Simulation.cpp:208-211).  A particle without
a position takes no part in the force pass.  Colours and spheres are kept
for round trips and ignored by the physics.
"""

from __future__ import annotations

import json
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Scenario:
    position: np.ndarray            # (n, 3) float32 (zeros where absent)
    velocity: np.ndarray            # (n, 3) float32 (zeros where absent)
    mass: np.ndarray                # (n,) float32 (zeros where absent)
    has_position: np.ndarray        # (n,) bool
    has_velocity: np.ndarray        # (n,) bool
    active: np.ndarray              # (n,) bool: has a Mass
    passive: np.ndarray             # (n,) bool: has an Acceleration
    acceleration: np.ndarray        # (n, 3) float32 initial Acceleration
    extra: list = field(default_factory=list)  # per particle: colour / sphere entries

    @property
    def n(self) -> int:
        return len(self.mass)


def _vec(d):
    return [float(d["x"]), float(d["y"]), float(d["z"])]


def loads(text: str) -> Scenario:
    return from_dict(json.loads(text))


def load(path: str) -> Scenario:
    with open(path) as f:
        return from_dict(json.load(f))


def from_dict(j: dict) -> Scenario:
    parts = j["particles"]
    n = len(parts)
    pos = np.zeros((n, 3), np.float32)
    vel = np.zeros((n, 3), np.float32)
    acc = np.zeros((n, 3), np.float32)
    m = np.zeros(n, np.float32)
    hp, hv, act, pas = (np.zeros(n, bool) for _ in range(4))
    extra = []
    for i, p in enumerate(parts):
        if "position" in p:
            pos[i] = _vec(p["position"])
            hp[i] = True
        if "velocity" in p:
            vel[i] = _vec(p["velocity"])
            hv[i] = True
        if "mass" in p:
            m[i] = float(p["mass"])
            act[i] = True
        if "acceleration" in p:
            acc[i] = _vec(p["acceleration"])
            pas[i] = True
        elif "passive" not in p or bool(p["passive"]):
            pas[i] = True
        extra.append({k: p[k] for k in ("color", "sphere") if k in p})
    return Scenario(pos, vel, m, hp, hv, act, pas, acc, extra)


def to_dict(s: Scenario) -> dict:
    parts = []
    for i in range(s.n):
        e = {}
        if s.has_position[i]:
            e["position"] = dict(zip("xyz", map(float, s.position[i])))
        if s.has_velocity[i]:
            e["velocity"] = dict(zip("xyz", map(float, s.velocity[i])))
        if s.active[i]:
            e["mass"] = float(s.mass[i])
        if s.passive[i]:
            e["acceleration"] = dict(zip("xyz", map(float, s.acceleration[i])))
        else:
            e["passive"] = False
        e.update(s.extra[i] if i < len(s.extra) else {})
        parts.append(e)
    return {"particles": parts}


def dumps(s: Scenario) -> str:
    return json.dumps(to_dict(s), indent=4)


def save(s: Scenario, path: str):
    with open(path, "w") as f:
        f.write(dumps(s))


# ------------------------------------------------------------------ tipsy
# from_tipsy (src/NBody/Simulation/Simulation.cpp:273-395).  Header: f64
# time, then u32 n, ndim, ngas, ndark, nstar (28 bytes), optionally padded to
# 32.  Records are FP32: gas 12 words (mass, pos3, vel3, rho, temp, h, metals,
# phi), dark matter 9 (mass, pos3, vel3, eps, phi), stars 11 (mass, pos3,
# vel3, metals, tform, eps, phi).  A file whose ndim is outside 1..3 is read
# byte-swapped.  Loaded: the gas particles only when there are no stars, never
# the dark matter, always the stars; each gets its mass, position, velocity,
# a zero acceleration (so it is active and passive), a sphere of radius
# cbrt(mass), white (gas) or |v| / ||v|| per component (stars) as colour.

_GAS_WORDS, _DARK_WORDS, _STAR_WORDS = 12, 9, 11


class TipsyError(ValueError):
    """The file's size does not match its header (the reference logs this and loads nothing)."""


def tipsy_header(data: bytes):
    """(time, n, ndim, ngas, ndark, nstar, big_endian) of a tipsy image."""
    if len(data) < 28:
        raise TipsyError("file shorter than a tipsy header")
    for order in ("<", ">"):
        t = np.frombuffer(data[:8], order + "f8")[0]
        n, ndim, ng, nd, ns = np.frombuffer(data[8:28], order + "u4")
        if 1 <= ndim <= 3:
            return float(t), int(n), int(ndim), int(ng), int(nd), int(ns), order == ">"
    raise TipsyError("tipsy header has no plausible dimension count in either byte order")


def from_tipsy(data: bytes) -> Scenario:
    t, n, ndim, ng, nd, ns, big = tipsy_header(data)
    body = 48 * ng + 36 * nd + 44 * ns
    if len(data) == 32 + body:
        off = 32
    elif len(data) == 28 + body:
        off = 28
    else:
        raise TipsyError(f"file size {len(data)} does not match the header ({ng} gas, {nd} dark, {ns} stars)")
    f32 = np.dtype(">f4" if big else "<f4")
    gas = np.frombuffer(data, f32, ng * _GAS_WORDS, off).reshape(ng, _GAS_WORDS)
    off += 48 * ng
    off += 36 * nd  # dark matter is skipped
    stars = np.frombuffer(data, f32, ns * _STAR_WORDS, off).reshape(ns, _STAR_WORDS)
    parts = [stars] if ns > 0 else [gas, stars]
    rec = np.concatenate([p[:, :7] for p in parts]).astype(np.float32) if parts else np.zeros((0, 7), np.float32)
    k = len(rec)
    m = rec[:, 0].copy()
    pos = rec[:, 1:4].copy()
    vel = rec[:, 4:7].copy()
    yes = np.ones(k, bool)
    extra = []
    n_gas_loaded = 0 if ns > 0 else ng
    for i in range(k):
        e = {"sphere": {"radius": float(np.cbrt(np.float32(m[i])))}}
        if i < n_gas_loaded:
            e["color"] = {"r": 1.0, "g": 1.0, "b": 1.0}
        else:
            v = vel[i].astype(np.float32)
            mag = np.float32(np.sqrt(np.float32(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])))
            # colorByVelocity (Simulation.cpp:264-271): red |vx|, blue |vy|, green |vz|, each / |v|
            with np.errstate(divide="ignore", invalid="ignore"):
                e["color"] = {"r": float(abs(v[0]) / mag), "g": float(abs(v[2]) / mag), "b": float(abs(v[1]) / mag)}
        extra.append(e)
    return Scenario(pos, vel, m, yes.copy(), yes.copy(), yes.copy(), yes.copy(), np.zeros((k, 3), np.float32), extra)


def load_tipsy(path: str) -> Scenario:
    with open(path, "rb") as f:
        return from_tipsy(f.read())


def accelerations(nbg, ctx, s: Scenario, params) -> np.ndarray:
    """ActiveTreeSolver::updateAccelerations on a scenario
    (Solvers/ActiveTreeSolver.h:33-72): the tree holds the active particles
    (position and mass) in file order; every particle with a position and an
    Acceleration gets the walk's result, others keep zero."""
    out = np.zeros((s.n, 3), np.float32)
    src = np.nonzero(s.active & s.has_position)[0]
    tgt = s.passive & s.has_position
    if len(src) == 0 or not tgt.any():
        return out
    b = nbg.Bodies(s.position[src, 0].copy(), s.position[src, 1].copy(), s.position[src, 2].copy(),
                   s.mass[src].copy(), np.zeros(len(src), np.float32), np.zeros(len(src), np.float32),
                   np.zeros(len(src), np.float32))
    sysm = nbg.System(ctx, params, len(src))
    sysm.set_state(b)
    own = sysm.update_accelerations()
    both = tgt[src]
    out[src[both]] = own[both]
    only = np.nonzero(tgt & ~(s.active & s.has_position))[0]
    if len(only):
        out[only] = sysm.accelerations_at(s.position[only, 0], s.position[only, 1], s.position[only, 2])
    sysm.close()
    return out
