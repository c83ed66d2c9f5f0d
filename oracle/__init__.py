"""ORACLE — TEST INFRASTRUCTURE ONLY.

ctypes wrapper of liboracle.so, the CPU restatement of the reference's
tree-code force pass (oracle/oracle.cpp).  Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg may import this package, and only as the
checker or the timed CPU baseline, never as a product path.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# ORACLE_LIB names another build of oracle.cpp (the mutation tests load
# builds with a planted mistake through it).
LIB_PATH = os.environ.get("ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")

_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        ip, fp = _i32p, _f32p
        sig = {
            "orc_bbox": (ctypes.c_int, [ctypes.c_int64, fp, fp, fp, fp]),
            "orc_morton": (ctypes.c_int, [ctypes.c_int64, fp, fp, fp, ctypes.c_int, fp, _u64p]),
            "orc_expand_bits_magic": (ctypes.c_uint64, [ctypes.c_uint64, ctypes.c_int]),
            "orc_expand_bits": (ctypes.c_uint64, [ctypes.c_uint64, ctypes.c_int]),
            "orc_sort": (ctypes.c_int, [ctypes.c_int64, _u64p, _u32p]),
            "orc_tree_new": (_vp, [ip, fp, ctypes.c_int64, fp, fp, fp, fp]),
            "orc_tree_free": (None, [_vp]),
            "orc_tree_num_nodes": (ctypes.c_int64, [_vp]),
            "orc_tree_max_depth": (ctypes.c_int, [_vp]),
            "orc_tree_nodes": (ctypes.c_int, [_vp, _i64p, _i64p, _i32p, _i32p, fp, fp]),
            "orc_tree_parents": (ctypes.c_int, [_vp, _i32p]),
            "orc_tree_quads": (ctypes.c_int, [_vp, fp]),
            "orc_fmm_accel": (ctypes.c_int, [_vp, fp, fp, fp, _u64p, _f64p, ctypes.c_int32]),
            "orc_fmm_starts": (ctypes.c_int64, [_vp, _i32p, ctypes.c_int64]),
            "orc_fmm_sample": (ctypes.c_int64, [_vp, ctypes.c_int32, ctypes.c_int32]),
            "orc_fmm_p2l": (ctypes.c_int, [fp, ctypes.c_float, fp, fp, ctypes.c_float, ctypes.c_float, fp]),
            "orc_node_particle_quad": (ctypes.c_int, [fp, ctypes.c_float, fp, fp, ctypes.c_float, ctypes.c_float,
                                                      fp]),
            "orc_fmm_recommend": (ctypes.c_int, [fp, fp, fp, ctypes.c_float]),
            "orc_fmm_particle_far": (ctypes.c_int, [fp, fp, ctypes.c_float]),
            "orc_tree_order": (ctypes.c_int, [_vp, _u32p, _u64p]),
            "orc_tree_bbox": (ctypes.c_int, [_vp, fp]),
            "orc_tree_accel": (ctypes.c_int, [_vp, ctypes.c_int64, fp, fp, fp, fp, fp, fp, _u64p, _u64p,
                                              ctypes.c_int]),
            "orc_tree_accel_f64": (ctypes.c_int, [_vp, ctypes.c_int64, fp, fp, fp, _f64p, _f64p, _f64p, _u64p,
                                                  _u64p, ctypes.c_int]),
            "orc_step": (ctypes.c_int, [ip, fp, ctypes.c_int64, fp, fp, fp, fp, fp, fp, fp, fp, fp, fp,
                                        ctypes.c_int, ctypes.c_int]),
            "orc_direct_f32": (ctypes.c_int, [ctypes.c_int64, fp, fp, fp, fp, ctypes.c_int64, fp, fp, fp,
                                              ctypes.c_float, ctypes.c_float, fp, fp, fp]),
            "orc_direct_f64": (ctypes.c_int, [ctypes.c_int64, fp, fp, fp, fp, ctypes.c_int64, fp, fp, fp,
                                              ctypes.c_double, ctypes.c_double, _f64p, _f64p, _f64p]),
            "orc_potential_f64": (ctypes.c_double, [ctypes.c_int64, fp, fp, fp, fp, ctypes.c_double,
                                                    ctypes.c_double]),
            "orc_particle_particle": (ctypes.c_int, [fp, ctypes.c_float, fp, ctypes.c_float, ctypes.c_float, fp]),
            "orc_boxes_intersect": (ctypes.c_int, [fp, fp]),
            "orc_com": (ctypes.c_int, [ctypes.c_int64, fp, fp, fp, fp, fp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a, t=_f32p):
    return None if a is None else a.ctypes.data_as(t)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _iparams(solver, leaf, max_depth, bits):
    return np.array([solver, leaf, max_depth, bits], np.int32)


def _fparams(theta, G, eps2, dt):
    return np.array([theta, G, eps2, dt], np.float32)


def bbox(x, y, z):
    out = np.zeros(6, np.float32)
    lib().orc_bbox(len(x), _p(_f32(x)), _p(_f32(y)), _p(_f32(z)), _p(out))
    return out


def morton(x, y, z, bits, bb):
    k = np.zeros(len(x), np.uint64)
    rc = lib().orc_morton(len(x), _p(_f32(x)), _p(_f32(y)), _p(_f32(z)), bits, _p(_f32(bb)), _p(k, _u64p))
    assert rc == 0
    return k


def sort(keys):
    keys = np.ascontiguousarray(keys, np.uint64)
    perm = np.zeros(len(keys), np.uint32)
    lib().orc_sort(len(keys), _p(keys, _u64p), _p(perm, _u32p))
    return perm


class Tree:
    """A reference tree over bodies: solver 0 octree, 1 LBVH, 3 / 4 their
    quadrupole variants (nodes also carry second moments), 5 the octree of
    the implicit FMM (fmm_accel)."""

    def __init__(self, solver, x, y, z, m, theta=0.5, leaf=16, max_depth=32, bits=63, G=1.0, eps2=1e-5, dt=1e-3):
        self._arrs = [_f32(a) for a in (x, y, z, m)]
        self.solver = solver
        self.h = lib().orc_tree_new(_p(_iparams(solver, leaf, max_depth, bits), _i32p),
                                    _p(_fparams(theta, G, eps2, dt)), len(x), *[_p(a) for a in self._arrs])
        if not self.h:
            raise ValueError("bad oracle parameters")
        self.n = len(x)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_tree_free(self.h)
            self.h = None

    def nodes(self):
        k = lib().orc_tree_num_nodes(self.h)
        first = np.zeros(k, np.int64)
        count = np.zeros(k, np.int64)
        leaf = np.zeros(k, np.int32)
        depth = np.zeros(k, np.int32)
        com4 = np.zeros((k, 4), np.float32)
        geo6 = np.zeros((k, 6), np.float32)
        lib().orc_tree_nodes(self.h, _p(first, _i64p), _p(count, _i64p), _p(leaf, _i32p), _p(depth, _i32p),
                             _p(com4), _p(geo6))
        parent = np.zeros(k, np.int32)
        lib().orc_tree_parents(self.h, _p(parent, _i32p))
        return dict(first=first, count=count, leaf=leaf.astype(bool), depth=depth, com4=com4, geo6=geo6,
                    parent=parent)

    def quads(self):
        """Per-node second moments (xx, xy, xz, yy, yz, zz), pre-order."""
        k = lib().orc_tree_num_nodes(self.h)
        q = np.zeros((k, 6), np.float32)
        lib().orc_tree_quads(self.h, _p(q))
        return q

    def fmm_accel(self, counts=False, threads=0):
        """Implicit FMM accelerations of every body (original order) and the
        counts (pp, np, pn, nn)."""
        a = [np.zeros(self.n, np.float32) for _ in range(3)]
        c = np.zeros(4, np.uint64)
        rc = lib().orc_fmm_accel(self.h, *[_p(v) for v in a], _p(c, _u64p), None, int(threads))
        if rc != 0:
            raise ValueError("not an FMM tree")
        acc = np.stack(a, axis=1)
        return (acc, tuple(int(v) for v in c)) if counts else acc

    def fmm_mass_received(self):
        """Per body, the FP64 sum of the masses of every source the implicit
        FMM's traversal reaches it with (each source once: the total mass)."""
        out = np.zeros(self.n, np.float64)
        c = np.zeros(4, np.uint64)
        if lib().orc_fmm_accel(self.h, None, None, None, _p(c, _u64p), _p(out, _f64p), 0) != 0:
            raise ValueError("not an FMM tree")
        return out

    def fmm_sample(self, k, threads=0):
        """Runs the implicit FMM's traversals from its first k starting nodes
        (a timing sample) on `threads` threads (0: all); returns the number of
        bodies they hold."""
        r = lib().orc_fmm_sample(self.h, int(k), int(threads))
        if r < 0:
            raise ValueError("not an FMM tree")
        return int(r)

    def fmm_starts(self):
        out = np.zeros(512, np.int32)
        k = lib().orc_fmm_starts(self.h, _p(out, _i32p), 512)
        return out[:k]

    def max_depth(self):
        return lib().orc_tree_max_depth(self.h)

    def order(self):
        perm = np.zeros(self.n, np.uint32)
        keys = np.zeros(self.n, np.uint64)
        lib().orc_tree_order(self.h, _p(perm, _u32p), _p(keys, _u64p))
        return keys, perm

    def bbox(self):
        out = np.zeros(6, np.float32)
        lib().orc_tree_bbox(self.h, _p(out))
        return out

    def accel(self, tx, ty, tz, threads=0, counts=False):
        tx, ty, tz = _f32(tx), _f32(ty), _f32(tz)
        nt = len(tx)
        a = [np.zeros(nt, np.float32) for _ in range(3)]
        pp = np.zeros(nt, np.uint64)
        npc = np.zeros(nt, np.uint64)
        lib().orc_tree_accel(self.h, nt, _p(tx), _p(ty), _p(tz), *[_p(v) for v in a], _p(pp, _u64p),
                             _p(npc, _u64p), threads)
        acc = np.stack(a, axis=1)
        return (acc, pp, npc) if counts else acc


    def accel_f64(self, tx, ty, tz, threads=0, counts=False):
        """The monopole walk's interaction set (the same FP32 opening
        decisions as accel) evaluated and summed in FP64: the exact value of
        the reference's FP32 result for each target, to measure rounding."""
        tx, ty, tz = _f32(tx), _f32(ty), _f32(tz)
        nt = len(tx)
        a = [np.zeros(nt, np.float64) for _ in range(3)]
        pp = np.zeros(nt, np.uint64)
        npc = np.zeros(nt, np.uint64)
        rc = lib().orc_tree_accel_f64(self.h, nt, _p(tx), _p(ty), _p(tz), *[_p(v, _f64p) for v in a],
                                      _p(pp, _u64p), _p(npc, _u64p), threads)
        if rc != 0:
            raise ValueError("accel_f64 needs a monopole Barnes-Hut or linear-BVH tree")
        acc = np.stack(a, axis=1)
        return (acc, pp, npc) if counts else acc


def step(solver, bodies, nsteps, theta=0.5, leaf=16, max_depth=32, bits=63, G=1.0, eps2=1e-5, dt=1e-3, threads=0):
    """Solver::step repeated nsteps times; returns the new (x,y,z,vx,vy,vz) and last accelerations."""
    x, y, z, m, vx, vy, vz = [np.array(a, np.float32, copy=True) for a in bodies]
    n = len(x)
    ax, ay, az = (np.zeros(n, np.float32) for _ in range(3))
    rc = lib().orc_step(_p(_iparams(solver, leaf, max_depth, bits), _i32p), _p(_fparams(theta, G, eps2, dt)), n,
                        _p(x), _p(y), _p(z), _p(m), _p(vx), _p(vy), _p(vz), _p(ax), _p(ay), _p(az), nsteps, threads)
    assert rc == 0
    return (x, y, z, vx, vy, vz), np.stack([ax, ay, az], axis=1)


def direct_f32(sx, sy, sz, sm, tx, ty, tz, G, eps2):
    nt = len(tx)
    a = [np.zeros(nt, np.float32) for _ in range(3)]
    lib().orc_direct_f32(len(sx), *[_p(_f32(v)) for v in (sx, sy, sz, sm)], nt,
                         *[_p(_f32(v)) for v in (tx, ty, tz)], G, eps2, *[_p(v) for v in a])
    return np.stack(a, axis=1)


def direct_f64(sx, sy, sz, sm, tx, ty, tz, G, eps2):
    nt = len(tx)
    a = [np.zeros(nt, np.float64) for _ in range(3)]
    lib().orc_direct_f64(len(sx), *[_p(_f32(v)) for v in (sx, sy, sz, sm)], nt,
                         *[_p(_f32(v)) for v in (tx, ty, tz)], G, eps2, *[_p(v, _f64p) for v in a])
    return np.stack(a, axis=1)


def potential_f64(x, y, z, m, G, eps2):
    return lib().orc_potential_f64(len(x), *[_p(_f32(v)) for v in (x, y, z, m)], G, eps2)


def particle_particle(a3, m, p3, G, eps2):
    out = np.zeros(3, np.float32)
    lib().orc_particle_particle(_p(_f32(a3)), m, _p(_f32(p3)), G, eps2, _p(out))
    return out


def node_particle_quad(c3, M, q6, p3, G, eps2):
    """Gravity::nodeParticle<2> for one node (COM c3, mass M, moment q6) and target p3."""
    out = np.zeros(3, np.float32)
    lib().orc_node_particle_quad(_p(_f32(c3)), M, _p(_f32(q6)), _p(_f32(p3)), G, eps2, _p(out))
    return out


def fmm_p2l(s3, m, c3, p3, G, eps2):
    """One source's local field at c3 evaluated at p3 (implicit FMM expansion)."""
    out = np.zeros(3, np.float32)
    lib().orc_fmm_p2l(_p(_f32(s3)), m, _p(_f32(c3)), _p(_f32(p3)), G, eps2, _p(out))
    return out


def fmm_recommend(a_cell4, a_com3, p_cell4, theta):
    """The implicit FMM's node-node criterion: 0 approximate, else
    1 (descend the active node) | 2 (descend the passive node).  Cells are
    (centre x, y, z, side)."""
    return int(lib().orc_fmm_recommend(_p(_f32(a_cell4)), _p(_f32(a_com3)), _p(_f32(p_cell4)), theta))


def fmm_particle_far(x3, p_cell4, theta):
    """The implicit FMM's particle-node criterion (body far from a passive cell)."""
    return bool(lib().orc_fmm_particle_far(_p(_f32(x3)), _p(_f32(p_cell4)), theta))


def boxes_intersect(a6, b6):
    return bool(lib().orc_boxes_intersect(_p(_f32(a6)), _p(_f32(b6))))


def com(x, y, z, m):
    out = np.zeros(4, np.float32)
    lib().orc_com(len(x), *[_p(_f32(v)) for v in (x, y, z, m)], _p(out))
    return out
