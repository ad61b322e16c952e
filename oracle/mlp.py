"""FP64 reference of the proposal network (test infrastructure).

TEST INFRASTRUCTURE ONLY: tests/ check the device's BF16-MFMA forward and
backward kernels and its loss gradient against this restatement.  The network
is this repository's design (the reference has no learned component); its
device definition is firecore_amd/csrc/proposal.h (forward, features) and
firecore_amd/csrc/proposal_train.hip (loss, backward).

Per atom i (bonded graph = the neighbour lists):
  x_i (16) = [onehot(elem)(4), log(|F|+1e-4), F^_z, log(|G|+1e-4), G^_z, F^.G^,
              (F^.p)^2, p_z^2, log(|Fb|+1e-4), F^.Fb^, deg/4, is_node, 0.1*h]
             F exact force, G its substrate (GridFF) part, p pi-vector (0 for
             caps), Fb mean force of the bonded neighbours, h height above the
             grid origin
  h1 = silu(W1 x + b1);  h2 = silu(W2 [h1_i; sum_j h1_j] + b2)
  h3 = silu(W3 [h2_i; sum_j h2_j] + b3);  o = W4 h3 + b4     (o: 2 per atom)
  star gain of node k   g_k = gstar * sigmoid(o_k0)
  molecule gain of c    g_c = grigid * sigmoid(mean_{i in c} o_i1)
With bf=True the forward pass rounds, as the device does, the inputs, the
dense weights, every stored activation and every neighbour sum to bfloat16.
"""
import numpy as np

NIN, NH, NOUT = 16, 64, 2
PARAM_SHAPES = [("W1", (NH, NIN)), ("b1", (NH,)), ("W2", (NH, 2 * NH)), ("b2", (NH,)),
                ("W3", (NH, 2 * NH)), ("b3", (NH,)), ("W4", (NOUT, NH)), ("b4", (NOUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in PARAM_SHAPES)


def unpack(theta):
    out, o = {}, 0
    for name, sh in PARAM_SHAPES:
        n = int(np.prod(sh))
        out[name] = np.asarray(theta[o:o + n], np.float64).reshape(sh)
        o += n
    return out


def pack(P):
    return np.concatenate([np.asarray(P[name], np.float64).reshape(-1) for name, _ in PARAM_SHAPES])


def init_params(seed=0, head_scale=0.1):
    rng = np.random.default_rng(seed)
    P = {}
    for name, sh in PARAM_SHAPES:
        if name.startswith("W"):
            P[name] = rng.normal(scale=np.sqrt(2.0 / sh[1]), size=sh) * (head_scale if name == "W4" else 1.0)
        else:
            P[name] = np.zeros(sh)
    return pack(P)


def bf16(a):
    """Round to nearest even bfloat16, returned as float64 (finite inputs)."""
    a32 = np.ascontiguousarray(np.asarray(a, np.float64).astype(np.float32))
    u = a32.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def neighbour_lists(natoms, neighs):
    return [[int(j) for j in neighs[i] if j >= 0] for i in range(natoms)]


def features(elem, F, G, pi, neighs, nnode, height):
    """x (N,16) from exact atom forces F, substrate forces G, pi-vectors (N,3;
    zero rows for caps) and heights (N,)."""
    N = len(elem)
    nb = neighbour_lists(N, neighs)
    Fb = np.array([F[l].mean(0) if l else np.zeros(3) for l in nb])
    nF = np.linalg.norm(F, axis=1)
    nG = np.linalg.norm(G, axis=1)
    nFb = np.linalg.norm(Fb, axis=1)
    Fh = F / (nF[:, None] + 1e-12)
    Gh = G / (nG[:, None] + 1e-12)
    Fbh = Fb / (nFb[:, None] + 1e-12)
    x = np.zeros((N, NIN))
    x[np.arange(N), np.asarray(elem)] = 1.0
    x[:, 4] = np.log(nF + 1e-4)
    x[:, 5] = Fh[:, 2]
    x[:, 6] = np.log(nG + 1e-4)
    x[:, 7] = Gh[:, 2]
    x[:, 8] = (Fh * Gh).sum(1)
    x[:, 9] = (Fh * pi).sum(1) ** 2
    x[:, 10] = pi[:, 2] ** 2
    x[:, 11] = np.log(nFb + 1e-4)
    x[:, 12] = (Fh * Fbh).sum(1)
    x[:, 13] = np.array([len(l) for l in nb]) / 4.0
    x[:, 14] = (np.arange(N) < nnode).astype(float)
    x[:, 15] = 0.1 * np.asarray(height)
    return x


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


def aggregate(h, nb):
    return np.array([h[l].sum(0) if l else np.zeros(h.shape[1]) for l in nb])


def forward(theta, x, nb, bf=False):
    """Outputs o (N,2) and the cache for backward."""
    P = unpack(theta)
    r = bf16 if bf else (lambda a: a)
    a0 = r(x)
    z1 = a0 @ r(P["W1"]).T + P["b1"]
    h1 = r(silu(z1))
    a1 = np.concatenate([h1, r(aggregate(h1, nb))], 1)
    z2 = a1 @ r(P["W2"]).T + P["b2"]
    h2 = r(silu(z2))
    a2 = np.concatenate([h2, r(aggregate(h2, nb))], 1)
    z3 = a2 @ r(P["W3"]).T + P["b3"]
    h3 = r(silu(z3))
    o = h3 @ P["W4"].T + P["b4"]
    return o, dict(a0=a0, z1=z1, a1=a1, z2=z2, a2=a2, z3=z3, h3=h3)


def backward(theta, cache, dout, nb, bf=False):
    """Parameter gradient for dloss/do = dout (N,2); with bf=True the dense
    weights used in the backward products are the bf16 ones (as on device)."""
    P = unpack(theta)
    r = bf16 if bf else (lambda a: a)
    K = cache
    G = {}
    G["W4"] = dout.T @ K["h3"]
    G["b4"] = dout.sum(0)
    dz3 = (dout @ P["W4"]) * dsilu(K["z3"])
    G["W3"] = dz3.T @ K["a2"]
    G["b3"] = dz3.sum(0)
    W3 = r(P["W3"])
    dh2 = dz3 @ W3[:, :NH] + aggregate(dz3, nb) @ W3[:, NH:]
    dz2 = dh2 * dsilu(K["z2"])
    G["W2"] = dz2.T @ K["a1"]
    G["b2"] = dz2.sum(0)
    W2 = r(P["W2"])
    dh1 = dz2 @ W2[:, :NH] + aggregate(dz2, nb) @ W2[:, NH:]
    dz1 = dh1 * dsilu(K["z1"])
    G["W1"] = dz1.T @ K["a0"]
    G["b1"] = dz1.sum(0)
    return pack(G)


def gains(o, nnode, comp, gstar, grigid):
    """(star gains per node, molecule gains per component index)."""
    gs = gstar * sigmoid(o[:nnode, 0])
    nc = int(comp.max()) + 1
    gc = np.array([grigid * sigmoid(o[comp == c, 1].mean()) for c in range(nc)])
    return gs, gc


def dgains_do(o, nnode, comp, gstar, grigid, dgs, dgc):
    """Chain rule from dloss/dgains to dloss/do (N,2)."""
    N = len(o)
    d = np.zeros((N, 2))
    s = sigmoid(o[:nnode, 0])
    d[:nnode, 0] = dgs * gstar * s * (1 - s)
    for c in range(len(dgc)):
        m = comp == c
        sc = sigmoid(o[m, 1].mean())
        d[m, 1] = dgc[c] * grigid * sc * (1 - sc) / m.sum()
    return d
