"""Independent cell reference: Qhull (scipy.spatial.HalfspaceIntersection).

A Voronoi / Laguerre cell clipped to a box is the intersection of the box
half-spaces with one half-space per other seed; Qhull computes it without any
of the oracle's code (no security radius, no dual triangles, no expansion
arithmetic), so agreement pins the oracle's cell algebra.
"""
import numpy as np
from scipy.spatial import ConvexHull, HalfspaceIntersection


def qhull_cell(x, i, w=None, box=(0, 0, 0, 1, 1, 1)):
    """Volume, centroid, {neighbour: facet area} of cell i by Qhull."""
    n = x.shape[0]
    pi = x[i]
    wi = 0.0 if w is None else w[i]
    H, ids = [], []
    for j in range(n):
        if j == i:
            continue
        e = x[j] - pi
        d2 = e @ e
        wj = 0.0 if w is None else w[j]
        # local coordinates y = x - pi: e.y <= (d2 + wi - wj)/2
        H.append(np.r_[e, -(d2 + wi - wj) / 2])
        ids.append(j)
    for a in range(3):
        v = np.zeros(3)
        v[a] = -1
        H.append(np.r_[v, -(pi[a] - box[a])])
        ids.append(-1)
        v = np.zeros(3)
        v[a] = 1
        H.append(np.r_[v, -(box[3 + a] - pi[a])])
        ids.append(-1)
    H = np.array(H)
    # an interior point: Chebyshev centre by a tiny LP (scipy.optimize.linprog)
    from scipy.optimize import linprog
    norm = np.linalg.norm(H[:, :3], axis=1)
    res = linprog(np.r_[0, 0, 0, -1.0], A_ub=np.c_[H[:, :3], norm], b_ub=-H[:, 3],
                  bounds=[(None, None)] * 3 + [(0, None)], method="highs")
    if res.status != 0 or res.x[3] <= 1e-14:
        return 0.0, None, {}
    hs = HalfspaceIntersection(H, res.x[:3])
    V = hs.intersections
    hull = ConvexHull(V)
    vol = hull.volume
    c0 = V.mean(axis=0)
    mg = np.zeros(3)
    for s in hull.simplices:
        a, b, c = V[s]
        tv = abs(np.dot(np.cross(a - c0, b - c0), c - c0)) / 6
        mg += tv * (a + b + c + c0) / 4
    cen = mg / vol + pi
    # facets: halfspaces with >= 3 incident vertices
    inc = {}
    for vi, fs in enumerate(hs.dual_facets):
        for f in fs:
            inc.setdefault(f, []).append(vi)
    areas = {}
    for f, vs in inc.items():
        if len(vs) < 3 or ids[f] < 0:
            continue
        P = V[vs]
        nrm = H[f, :3] / np.linalg.norm(H[f, :3])
        c = P.mean(axis=0)
        u = P[0] - c
        u -= nrm * (u @ nrm)
        u /= np.linalg.norm(u)
        v = np.cross(nrm, u)
        ang = np.arctan2((P - c) @ v, (P - c) @ u)
        P = P[np.argsort(ang)]
        a = 0.0
        for k in range(1, len(P) - 1):
            a += 0.5 * np.linalg.norm(np.cross(P[k] - P[0], P[k + 1] - P[0]))
        if a > 0:
            areas[ids[f]] = areas.get(ids[f], 0.0) + a
    return vol, cen, areas
