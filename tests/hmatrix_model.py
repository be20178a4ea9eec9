"""CPU restatement of the hierarchical path's cluster tree and block tree.

Test infrastructure: it fixes, index for index, the tree order and block
lists the library must produce (the rules are the library's own; the
reference builds its trees with HMatrices.jl, which is not available here,
so the cluster-tree indexing is pinned against this restatement).

Cluster tree: a node covering tree positions [lo, hi) splits when it holds
more than `leaf` points, along the longest side of its tight bounding box at
the midpoint (stable partition, coordinate < midpoint to the left; an empty
side falls back to an even count split).  Block tree, depth first from
(root, root): pairs with dist > near_radius and min(diam) < eta * dist are
low-rank, pairs of two leaves are dense, otherwise the non-leaf side(s) split.
"""
import numpy as np


def cluster_tree(points, leaf):
    """points: (n, 3).  Returns perm and the node list [(lo, hi, left, right, bmin, bmax)]."""
    n = points.shape[0]
    perm = np.arange(n)
    nodes = []

    def box(lo, hi):
        p = points[perm[lo:hi]]
        return p.min(axis=0), p.max(axis=0)

    b0 = box(0, n)
    nodes.append([0, n, -1, -1, b0[0], b0[1]])
    stack = [0]
    while stack:
        nid = stack.pop()
        lo, hi = nodes[nid][0], nodes[nid][1]
        if hi - lo <= leaf:
            continue
        bmin, bmax = nodes[nid][4], nodes[nid][5]
        ext = bmax - bmin
        axis = int(np.argmax(ext))  # first maximum
        mid = 0.5 * (bmin[axis] + bmax[axis])
        seg = perm[lo:hi]
        left = seg[points[seg, axis] < mid]
        right = seg[~(points[seg, axis] < mid)]
        perm[lo:hi] = np.concatenate([left, right])
        split = lo + len(left)
        if split in (lo, hi):
            split = lo + (hi - lo) // 2
        li = len(nodes)
        for a, b in ((lo, split), (split, hi)):
            bb = box(a, b)
            nodes.append([a, b, -1, -1, bb[0], bb[1]])
        nodes[nid][2], nodes[nid][3] = li, li + 1
        stack += [li + 1, li]
    return perm, nodes


def block_tree(nodes, eta, near_radius):
    """Returns (dense, lowrank) lists of (row node, col node)."""
    def diam(a):
        return float(np.sqrt(((a[5] - a[4]) ** 2).sum()))

    def dist(a, b):
        g = np.maximum(0.0, np.maximum(b[4] - a[5], a[4] - b[5]))
        return float(np.sqrt((g ** 2).sum()))

    dense, low = [], []
    stack = [(0, 0)]
    while stack:
        i, j = stack.pop()
        a, b = nodes[i], nodes[j]
        d = dist(a, b)
        if d > near_radius and min(diam(a), diam(b)) < eta * d:
            low.append((i, j))
            continue
        la, lb = a[2] < 0, b[2] < 0
        if la and lb:
            dense.append((i, j))
        elif la:
            stack += [(i, b[3]), (i, b[2])]
        elif lb:
            stack += [(a[3], j), (a[2], j)]
        else:
            stack += [(a[3], b[3]), (a[3], b[2]), (a[2], b[3]), (a[2], b[2])]
    return dense, low


def plan(points, leaf=64, eta=2.0, near_radius=0.0):
    """Same outputs as bemhip.hmatrix_plan: perm, node count, block rows."""
    perm, nodes = cluster_tree(np.asarray(points, dtype=np.float64), leaf)
    dense, low = block_tree(nodes, eta, near_radius)
    rows = [(0, nodes[i][0], nodes[i][1] - nodes[i][0], nodes[j][0], nodes[j][1] - nodes[j][0]) for i, j in dense]
    rows += [(1, nodes[i][0], nodes[i][1] - nodes[i][0], nodes[j][0], nodes[j][1] - nodes[j][0]) for i, j in low]
    return perm, len(nodes), np.array(rows, dtype=np.int64).reshape(-1, 5)
