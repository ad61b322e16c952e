"""FP64 reference of the FIRE force preconditioner (test infrastructure).

TEST INFRASTRUCTURE ONLY: tests/ compare the device's preconditioned force
with this restatement.  The preconditioner is this repository's design (the
reference integrates the raw force, MolWorld_sp3.h:2400-2440); its device
definition is firecore_amd/csrc/precond.h.

    F~_a = F_a + sum_{stars k containing a} g_k * P_k(F)_a + g_c(a) * P_c(F)_a

P_G(F) is the orthogonal projection (Euclidean metric, unit masses) of the
force field restricted to the atom group G onto the 6-dimensional space of
rigid motions of G (translations and infinitesimal rotations about the
centroid): P_G(F)_a = t + w x (r_a - c), t = mean force, w = I^-1 L.  A star
is a node and its bonded neighbours; c(a) is the connected component of a.

Pinned in tests/test_precond.py against an independent construction (least
squares onto an explicit rigid-motion basis), idempotence, self-adjointness
and positive definiteness of the preconditioner, not by retyping this formula.
"""
import numpy as np


def rigid_project(x, F, reg=0.0):
    """t + w x (x - c) for one group (x, F: (n,3)).  reg adds reg*tr(I) to I
    for groups of three or more atoms (the device uses reg=1e-6); pairs use
    the exact projection onto the rotations perpendicular to their axis."""
    n = len(x)
    c = x.mean(0)
    r = x - c
    t = F.mean(0)
    L = np.cross(r, F).sum(0)
    if n == 1:
        return np.tile(t, (1, 1))
    if n == 2:  # rotations about the pair's axis are not rigid motions of it
        w = L / (r * r).sum()
        return t + np.cross(w, r)
    I = (r * r).sum() * np.eye(3) - r.T @ r
    I = I + (reg * np.trace(I) + 1e-300) * np.eye(3)
    w = np.linalg.solve(I, L)
    return t + np.cross(w, r)


def rigid_project_lstsq(x, F):
    """The same projection built from an explicit basis of the 6 rigid
    motions of the group (independent construction)."""
    n = len(x)
    c = x.mean(0)
    r = x - c
    B = np.zeros((3 * n, 6))
    for k in range(3):
        B[k::3, k] = 1.0
    for k, e in enumerate(np.eye(3)):
        B[:, 3 + k] = np.cross(e, r).reshape(-1)
    coef, *_ = np.linalg.lstsq(B, F.reshape(-1), rcond=None)
    return (B @ coef).reshape(n, 3)


def components(natoms, neighs, maxc=4):
    """Connected components of the bonded graph, indices >= maxc folded into
    maxc-1 (the device's rule)."""
    parent = list(range(natoms))

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a

    for i in range(natoms):
        for j in neighs[i]:
            if 0 <= j < natoms:
                a, b = find(i), find(int(j))
                if a != b:
                    parent[max(a, b)] = min(a, b)
    ids, out = {}, np.zeros(natoms, np.int64)
    for i in range(natoms):
        r = find(i)
        if r not in ids:
            ids[r] = len(ids)
        out[i] = min(ids[r], maxc - 1)
    return out


def precondition(x, F, neighs, nnode, gstar, grigid, comp=None, reg=1e-6):
    """Preconditioned atom forces.  x, F: (natoms,3); neighs (natoms,4);
    gstar: scalar or per-node array; grigid: scalar or per-component array."""
    N = len(x)
    if comp is None:
        comp = components(N, neighs)
    gs = np.broadcast_to(np.asarray(gstar, float), (nnode,))
    gr = np.broadcast_to(np.asarray(grigid, float), (int(comp.max()) + 1,))
    out = F.copy()
    for k in range(nnode):
        S = [k] + [int(j) for j in neighs[k] if j >= 0]
        out[S] += gs[k] * rigid_project(x[S], F[S], reg)
    for c in range(int(comp.max()) + 1):
        S = np.where(comp == c)[0]
        out[S] += gr[c] * rigid_project(x[S], F[S], reg)
    return out
