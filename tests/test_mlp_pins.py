"""Independent pins of the proposal-network oracle (oracle/mlp.py), CPU only.

* mlp.forward / mlp.backward against a separate torch float64 definition of the
  same network (torch.nn.functional.silu, index_add_ aggregation over the bonded
  graph) and torch autograd for every parameter block;
* mlp.forward(bf=True) against the same torch network with torch's own
  bfloat16 rounding at the places the device rounds;
* mlp.gains / mlp.dgains_do (the chain rule the device loss gradient is
  compared against) against central finite differences of the gains.
"""
import numpy as np
import pytest

from oracle import mlp

torch = pytest.importorskip("torch")


def _graph(N, rng):
    neighs = -np.ones((N, 4), int)
    for i in range(1, N):
        j = int(rng.integers(0, i))
        while (neighs[j] >= 0).sum() >= 4:
            j = int(rng.integers(0, i))
        neighs[i][list(neighs[i]).index(-1)] = j
        neighs[j][list(neighs[j]).index(-1)] = i
    return neighs


def _torch_net(theta, x, neighs, bf=False):
    """The network written directly in torch (float64): returns (o, params)."""
    th = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    shapes = dict(mlp.PARAM_SHAPES)
    P, o = {}, 0
    for name, _ in mlp.PARAM_SHAPES:
        n = int(np.prod(shapes[name]))
        P[name] = th[o:o + n].reshape(shapes[name])
        o += n
    rb = (lambda t: t.to(torch.bfloat16).to(torch.float64)) if bf else (lambda t: t)
    src, dst = [], []
    for i, row in enumerate(neighs):
        for j in row:
            if j >= 0:
                src.append(int(j))
                dst.append(i)
    src, dst = torch.tensor(src), torch.tensor(dst)

    def agg(h):
        return torch.zeros_like(h).index_add_(0, dst, h[src])

    silu = torch.nn.functional.silu
    a0 = rb(torch.tensor(x, dtype=torch.float64))
    h1 = rb(silu(a0 @ rb(P["W1"]).T + P["b1"]))
    h2 = rb(silu(torch.cat([h1, rb(agg(h1))], 1) @ rb(P["W2"]).T + P["b2"]))
    h3 = rb(silu(torch.cat([h2, rb(agg(h2))], 1) @ rb(P["W3"]).T + P["b3"]))
    out = h3 @ P["W4"].T + P["b4"]
    return out, th


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_forward_and_backward_match_torch_autograd(seed):
    rng = np.random.default_rng(seed)
    N = 37
    neighs = _graph(N, rng)
    nb = mlp.neighbour_lists(N, neighs)
    x = rng.normal(size=(N, mlp.NIN))
    theta = mlp.init_params(seed + 10, head_scale=1.0)
    theta[mlp.NPARAM - 2:] = rng.normal(size=2)  # nonzero head biases
    o, cache = mlp.forward(theta, x, nb)
    ot, th = _torch_net(theta, x, neighs)
    assert np.abs(o - ot.detach().numpy()).max() < 1e-12 * (1 + np.abs(o).max())
    w = rng.normal(size=(N, 2))
    (ot * torch.tensor(w)).sum().backward()
    g_ref = th.grad.numpy()
    g = mlp.backward(theta, cache, w, nb)
    o_ = 0
    for name, sh in mlp.PARAM_SHAPES:
        n = int(np.prod(sh))
        a, b = g[o_:o_ + n], g_ref[o_:o_ + n]
        assert np.abs(a - b).max() <= 1e-11 * (1 + np.abs(b).max()), name
        o_ += n


def test_bf16_forward_matches_torch_rounding():
    rng = np.random.default_rng(5)
    N = 29
    neighs = _graph(N, rng)
    nb = mlp.neighbour_lists(N, neighs)
    x = rng.normal(size=(N, mlp.NIN)) * 3.0
    theta = mlp.init_params(8, head_scale=1.0)
    o, _ = mlp.forward(theta, x, nb, bf=True)
    ot, _ = _torch_net(theta, x, neighs, bf=True)
    # same roundings at the same places: agreement to FP64 summation order
    assert np.abs(o - ot.detach().numpy()).max() < 1e-10 * (1 + np.abs(o).max())
    # and the roundings matter (a forgotten one would show up at this size)
    o32, _ = mlp.forward(theta, x, nb, bf=False)
    assert np.abs(o - o32).max() > 1e-4


@pytest.mark.parametrize("ncomp", [1, 3])
def test_dgains_do_matches_finite_differences(ncomp):
    rng = np.random.default_rng(11 + ncomp)
    N, M = 23, 9
    gstar, grigid = 3.0, 127.0
    o = rng.normal(size=(N, 2)) * 1.5
    comp = np.sort(rng.integers(0, ncomp, N))
    comp[:ncomp] = np.arange(ncomp)
    comp = np.sort(comp)
    dgs = rng.normal(size=M)
    dgc = rng.normal(size=ncomp)

    def L(oo):
        gs, gc = mlp.gains(oo, M, comp, gstar, grigid)
        return (gs * dgs).sum() + (gc * dgc).sum()

    d = mlp.dgains_do(o, M, comp, gstar, grigid, dgs, dgc)
    h = 1e-6
    for i in range(N):
        for k in range(2):
            e = np.zeros_like(o)
            e[i, k] = h
            fd = (L(o + e) - L(o - e)) / (2 * h)
            assert abs(fd - d[i, k]) <= 1e-7 * (1 + abs(fd)), (i, k, fd, d[i, k])
    # caps carry no star gain
    assert np.all(d[M:, 0] == 0.0)


def test_gains_closed_form():
    """sigmoid(0) = 1/2: zero outputs give half of each ceiling."""
    N, M = 10, 4
    gs, gc = mlp.gains(np.zeros((N, 2)), M, np.zeros(N, np.int64), 3.0, 127.0)
    assert np.allclose(gs, 1.5) and np.allclose(gc, 63.5)
