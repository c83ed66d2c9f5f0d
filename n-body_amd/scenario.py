"""Scenario files in the reference simulator's JSON format.

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
