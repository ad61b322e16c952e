"""FP64 reference of the learned displacement proposal (test infrastructure).

TEST INFRASTRUCTURE ONLY: used by tests/ to check the device's BF16-MFMA
forward/backward kernels.  The model itself is this repository's design (the
reference has no learned component); its definition lives in the product
(firecore_amd/csrc/mlp_device.h) and is restated here in numpy/float64.

Per atom i of a system (bonded graph = the neighbour lists):
  inputs   element one-hot e_i (C,N,O,H), force F_i, pi-vector p_i (0 for
           caps), GridFF force G_i; Fb_i = mean of F_j over bonded j,
           Fbb_i = mean of Fb_j over bonded j
  x_i (16) = [e_i(4), log(|F|+1e-4), Fz/|F|, log(|G|+1e-4), Gz/|G|, F^.G^,
              F^.p, p_z, log(|Fb|+1e-4), F^.Fb^, Fb_z/|Fb|, deg/4, |p|]
           (rotation invariants; z is the substrate normal)
  h1 = silu(W1 x + b1)                      16 -> 64
  h2 = silu(W2 [h1_i ; sum_j h1_j] + b2)   128 -> 64
  h3 = silu(W3 [h2_i ; sum_j h2_j] + b3)   128 -> 64
  c  = W4 h3 + b4                           64 -> 4
  d_i = CSCALE * (c0 F_i + c1 Fb_i + c2 G_i + c3 Fbb_i)   (displacement, A)
The displacement is a learned combination of equivariant vectors, so the map
commutes with rotations about the substrate normal (and with all rotations in
vacuum).  Weights are stored as bf16 on the device; pass `bf16=True` to round
weights and layer inputs the way the device does.
"""
import numpy as np

NIN, NH, NOUT = 16, 64, 4
CSCALE = 1.0  # A^2/eV: coefficient units of the head
EPS = 1e-4

PARAM_SHAPES = [("W1", (NH, NIN)), ("b1", (NH,)), ("W2", (NH, 2 * NH)), ("b2", (NH,)),
                ("W3", (NH, 2 * NH)), ("b3", (NH,)), ("W4", (NOUT, NH)), ("b4", (NOUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in PARAM_SHAPES)


def unpack(theta):
    out, o = {}, 0
    for name, sh in PARAM_SHAPES:
        n = int(np.prod(sh))
        out[name] = theta[o:o + n].reshape(sh)
        o += n
    return out


def pack(P):
    return np.concatenate([P[name].reshape(-1) for name, _ in PARAM_SHAPES])


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    P = {}
    for name, sh in PARAM_SHAPES:
        if name.startswith("W"):
            P[name] = rng.normal(scale=np.sqrt(2.0 / sh[1]), size=sh) * (0.1 if name == "W4" else 1.0)
        else:
            P[name] = np.zeros(sh)
    return pack(P)


def bf16(a):
    """Round-to-nearest-even to bfloat16, returned as float64."""
    a32 = np.asarray(a, np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def neighbour_lists(natoms, neighs):
    """Bonded neighbours of every atom (symmetric lists from neighs)."""
    return [[int(j) for j in neighs[i] if j >= 0] for i in range(natoms)]


def features(elem, F, pi, G, neighs):
    """x (N,16) and the equivariant basis (N,4,3)."""
    N = len(elem)
    nb = neighbour_lists(N, neighs)
    Fb = np.array([F[l].mean(0) if l else np.zeros(3) for l in nb])
    Fbb = np.array([Fb[l].mean(0) if l else np.zeros(3) for l in nb])
    nF = np.linalg.norm(F, axis=1)
    nG = np.linalg.norm(G, axis=1)
    nFb = np.linalg.norm(Fb, axis=1)
    np_ = np.linalg.norm(pi, axis=1)
    Fh = F / (nF[:, None] + 1e-12)
    Gh = G / (nG[:, None] + 1e-12)
    Fbh = Fb / (nFb[:, None] + 1e-12)
    x = np.zeros((N, NIN))
    x[np.arange(N), np.asarray(elem)] = 1.0
    x[:, 4] = np.log(nF + EPS)
    x[:, 5] = Fh[:, 2]
    x[:, 6] = np.log(nG + EPS)
    x[:, 7] = Gh[:, 2]
    x[:, 8] = (Fh * Gh).sum(1)
    x[:, 9] = (Fh * pi).sum(1)
    x[:, 10] = pi[:, 2]
    x[:, 11] = np.log(nFb + EPS)
    x[:, 12] = (Fh * Fbh).sum(1)
    x[:, 13] = Fbh[:, 2]
    x[:, 14] = np.array([len(l) for l in nb]) / 4.0
    x[:, 15] = np_
    basis = np.stack([F, Fb, G, Fbb], 1)
    return x, basis, nb


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def aggregate(h, nb):
    return np.array([h[l].sum(0) if l else np.zeros(h.shape[1]) for l in nb])


def forward(theta, x, nb, bf=False):
    P = unpack(theta)
    r = bf16 if bf else (lambda a: a)
    cache = {}
    a0 = r(x)
    z1 = a0 @ r(P["W1"]).T + P["b1"]
    h1 = silu(z1)
    a1 = r(np.concatenate([h1, aggregate(h1, nb)], 1))
    z2 = a1 @ r(P["W2"]).T + P["b2"]
    h2 = silu(z2)
    a2 = r(np.concatenate([h2, aggregate(h2, nb)], 1))
    z3 = a2 @ r(P["W3"]).T + P["b3"]
    h3 = silu(z3)
    a3 = r(h3)
    c = a3 @ r(P["W4"]).T + P["b4"]
    cache.update(a0=a0, z1=z1, a1=a1, z2=z2, a2=a2, z3=z3, a3=a3)
    return c, cache


def displacement(c, basis):
    return CSCALE * np.einsum("nk,nkd->nd", c, basis)


def loss(theta, x, basis, nb, target, w, bf=False):
    """Weighted squared error sum_i w_i |d_i - target_i|^2."""
    c, _ = forward(theta, x, nb, bf)
    d = displacement(c, basis)
    return float((w[:, None] * (d - target) ** 2).sum())


def grad(theta, x, basis, nb, target, w, bf=False):
    """Analytic gradient of `loss` (layer inputs rounded as in forward when bf)."""
    P = unpack(theta)
    r = bf16 if bf else (lambda a: a)
    c, K = forward(theta, x, nb, bf)
    d = displacement(c, basis)
    dd = 2.0 * w[:, None] * (d - target)                 # dL/dd
    dc = CSCALE * np.einsum("nd,nkd->nk", dd, basis)      # dL/dc
    G = {}
    G["W4"] = dc.T @ K["a3"]
    G["b4"] = dc.sum(0)
    dh3 = dc @ r(P["W4"])
    dz3 = dh3 * dsilu(K["z3"])
    G["W3"] = dz3.T @ K["a2"]
    G["b3"] = dz3.sum(0)
    da2 = dz3 @ r(P["W3"])
    dh2 = da2[:, :NH] + aggregate(da2[:, NH:], nb)       # sum over neighbours is self-adjoint (symmetric graph)
    dz2 = dh2 * dsilu(K["z2"])
    G["W2"] = dz2.T @ K["a1"]
    G["b2"] = dz2.sum(0)
    da1 = dz2 @ r(P["W2"])
    dh1 = da1[:, :NH] + aggregate(da1[:, NH:], nb)
    dz1 = dh1 * dsilu(K["z1"])
    G["W1"] = dz1.T @ K["a0"]
    G["b1"] = dz1.sum(0)
    return pack(G)
