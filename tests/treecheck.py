"""Independent checks of tree structure, shared by the oracle and GPU tests.

These do not restate the build algorithms; they verify the properties that
define the reference's trees:
  * LBVH (src/NBody/Simulation/Solvers/Trees/LinearBVH.h:47-76,114-119): the
    root covers every body; a node is a leaf iff it has <= L bodies or equal
    first/last keys; an internal node has exactly two children that split
    its range at the first key with the highest differing bit set.
  * Octree (Trees/Octree.h:33-61,100-147,185-205): levels 0 and 1 always
    split; deeper nodes split iff > L bodies and depth < 2 + maxDepth; each
    body of a child lies on the child's side of every parent centre plane
    (p[d] < c[d] for the lower half); child k's centre is the parent's plus
    (+-(side/2)/2) per axis in FP32 with k = 4x + 2y + z; every body lies in
    its cell's box [c - side/2, c + side/2).
"""

from __future__ import annotations

import numpy as np


def lbvh_check(keys_sorted: np.ndarray, first, count, leaf, L: int):
    n = len(keys_sorted)
    nodes = {}
    for f, c, lf in zip(first.tolist(), count.tolist(), np.asarray(leaf).tolist()):
        assert (f, c) not in nodes, f"duplicate node range {(f, c)}"
        nodes[(f, c)] = bool(lf)
    if n == 0:
        assert not nodes
        return 0
    assert (0, n) in nodes, "root must cover all bodies"
    seen = 0
    stack = [(0, n)]
    while stack:
        f, c = stack.pop()
        seen += 1
        k0, k1 = int(keys_sorted[f]), int(keys_sorted[f + c - 1])
        should_split = c > L and k0 != k1
        assert nodes[(f, c)] == (not should_split), f"leaf rule violated at {(f, c)}"
        if not should_split:
            continue
        msb = (k0 ^ k1).bit_length() - 1
        seg = keys_sorted[f:f + c].astype(object)
        bits = np.array([(int(k) >> msb) & 1 for k in seg])
        s = int(np.argmax(bits == 1))
        assert s > 0 and bits[:s].sum() == 0 and bits[s:].all(), "keys not partitioned by the split bit"
        left, right = (f, s), (f + s, c - s)
        assert left in nodes and right in nodes, f"children of {(f, c)} missing"
        stack += [right, left]
    assert seen == len(nodes), "nodes not reachable from the root"
    return seen


def summaries_check(pos_sorted: np.ndarray, mass_sorted: np.ndarray, first, count, com4, geo6=None,
                    rel=2e-5):
    """COM / mass close to their FP64 definitions; bbox exactly min/max."""
    for i, (f, c) in enumerate(zip(first.tolist(), count.tolist())):
        if c == 0:
            continue
        p = pos_sorted[f:f + c].astype(np.float64)
        m = mass_sorted[f:f + c].astype(np.float64)
        M = m.sum()
        assert abs(com4[i, 3] - M) <= rel * M + 1e-30, f"mass of node {i}"
        if M > 0:
            C = (p * m[:, None]).sum(0) / M
            ext = max(np.abs(p).max(), 1e-30)
            assert np.all(np.abs(com4[i, :3] - C) <= rel * ext), f"COM of node {i}"
        if geo6 is not None:
            assert np.array_equal(geo6[i, :3], pos_sorted[f:f + c].min(0)), f"bbox min of node {i}"
            assert np.array_equal(geo6[i, 3:], pos_sorted[f:f + c].max(0)), f"bbox max of node {i}"


def preorder_parents(depth: np.ndarray) -> np.ndarray:
    parent = np.full(len(depth), -1, np.int64)
    stack = []
    for i, d in enumerate(depth.tolist()):
        while stack and depth[stack[-1]] >= d:
            stack.pop()
        parent[i] = stack[-1] if stack else -1
        stack.append(i)
    return parent


def octree_check(pos: np.ndarray, perm: np.ndarray, nodes: dict, L: int, max_depth: int):
    """pos: original-order positions; perm: leaf body order; nodes: with
    explicit parent links (index 0 is the root)."""
    depth = nodes["depth"]
    parent = nodes["parent"]
    first, count, leaf, geo = nodes["first"], nodes["count"], nodes["leaf"], nodes["geo6"]
    n = len(pos)
    assert count[0] == n and depth[0] == 0 and parent[0] == -1
    nkids = np.bincount(parent[1:], minlength=len(depth))
    for i in range(len(depth)):
        d, c = int(depth[i]), int(count[i])
        if d <= 1:
            assert not leaf[i], "levels 0 and 1 are always split"
        elif c > 0:
            assert bool(leaf[i]) == (not (c > L and d < 2 + max_depth)), f"split rule at node {i}"
        # a split creates all 8 children, empty ones included (Trees/Octree.h:100-147)
        assert nkids[i] == (0 if leaf[i] else 8), f"node {i} has {nkids[i]} children"
        if i > 0:
            p = int(parent[i])
            assert depth[p] == d - 1, f"depth of node {i}"
            assert geo[i, 3] == np.float32(geo[p, 3] / np.float32(2)), f"side of node {i}"
        if i > 0 and c > 0:
            pc = geo[parent[i], :3]
            cc = geo[i, :3]
            body = pos[perm[first[i]:first[i] + c]]
            for dim in range(3):
                lower = cc[dim] < pc[dim]
                side = body[:, dim] < pc[dim]
                assert np.all(side == lower), f"node {i}: body on the wrong side of its parent's plane"
    # Child centres: OctreeNode::split (Trees/Octree.h:113-146) places child
    # k (= 4x + 2y + z, the children array order) at the parent centre plus
    # (+-o, +-o, +-o), o = (side / 2) / 2 in FP32, + on the axes whose bit is
    # set.  The children of a node appear in k order (pre-order node ids).
    f32 = np.float32
    kids = {}
    for i in range(1, len(depth)):
        kids.setdefault(int(parent[i]), []).append(i)
    for pnode, ch in kids.items():
        o = f32(f32(geo[pnode, 3] / f32(2)) / f32(2))
        pc = geo[pnode, :3].astype(f32)
        for k, c in enumerate(ch):
            bits = ((k >> 2) & 1, (k >> 1) & 1, k & 1)
            want = np.array([f32(pc[d] + (o if bits[d] else -o)) for d in range(3)], f32)
            assert np.array_equal(geo[c, :3], want), f"centre of child {k} of node {pnode}"
    # Containment: every body of a cell lies in the cell [c - s/2, c + s/2)
    # (OctreeNode::boundingBox, Octree.h:92-98), up to the rounding of the
    # FP32 centres (a body on the parent's plane side can sit half an ulp of
    # the centre outside its child's nominal box).
    for i in range(len(depth)):
        c = int(count[i])
        if c == 0:
            continue
        body = pos[perm[first[i]:first[i] + c]].astype(np.float64)
        ctr = geo[i, :3].astype(np.float64)
        h = float(geo[i, 3]) / 2.0
        slack = np.spacing(np.abs(geo[i, :3]).astype(f32) + f32(geo[i, 3])).astype(np.float64)
        assert np.all(body >= ctr - h - slack) and np.all(body < ctr + h + slack), f"body outside cell {i}"
    # a parent's count is the sum of its children's
    sums = np.bincount(parent[1:], weights=count[1:], minlength=len(depth))
    inner = ~leaf.astype(bool)
    assert np.array_equal(sums[inner].astype(np.int64), count[inner]), "child counts must sum to the parent's"
    # leaves partition the bodies
    lc = count[leaf.astype(bool)].sum()
    assert lc == n, "leaves must hold every body once"
    assert sorted(perm.tolist()) == list(range(n))


def com_restatement(first, count, leaf, parent, depth, pos_sorted, mass_sorted, octree):
    """CenterOfMassSummary::summarize (Physics/Summaries/CenterOfMassSummary.h:
    30-55) restated in numpy float32, independently of the oracle: a leaf
    sums M += m and C = C + m p over its bodies in order, an internal node
    sums M += M_c and C = C + com_c M_c over its children in order (pre-order
    ids), then com = C / M.  Octree nodes that are never summarized keep
    the default zero COM and mass: empty nodes below level 1 (NodeBase::refine
    returns on them, Trees/NodeBase.h:57-75); levels 0 and 1 are always split
    and summarized (Trees/Tree.h:172-216), so an empty one gets 0 / 0.
    Returns com4 (x, y, z, M) per node."""
    f32 = np.float32
    k = len(first)
    M = np.zeros(k, f32)
    C = np.zeros((k, 3), f32)
    summarized = np.ones(k, bool)
    if octree:
        summarized = (count > 0) | (depth <= 1)
    lf = np.flatnonzero(leaf.astype(bool) & summarized & (count > 0))
    cnt = count[lf]
    for j in range(int(cnt.max()) if len(lf) else 0):
        a = lf[cnt > j]
        idx = first[a] + j
        m = mass_sorted[idx].astype(f32)
        M[a] = M[a] + m
        C[a] = C[a] + pos_sorted[idx].astype(f32) * m[:, None]
    kids = {}
    for i in range(1, k):
        kids.setdefault(int(parent[i]), []).append(i)
    inner = [i for i in range(k) if not leaf[i] and summarized[i]]
    com = np.zeros((k, 3), f32)
    with np.errstate(divide="ignore", invalid="ignore"):
        done = np.zeros(k, bool)
        for i in lf:
            com[i] = C[i] / M[i]
            done[i] = True
        for i in sorted(inner, key=lambda v: -int(depth[v])):  # children before parents
            for c in kids.get(i, []):
                M[i] = M[i] + M[c]
                C[i] = C[i] + com[c] * M[c]
            com[i] = C[i] / M[i]
            done[i] = True
    out = np.zeros((k, 4), f32)
    out[:, :3] = com
    out[:, 3] = M
    return out


def same_bits(a, b):
    """Element-wise equality of float arrays with NaN equal to NaN."""
    a, b = np.asarray(a), np.asarray(b)
    return bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))
