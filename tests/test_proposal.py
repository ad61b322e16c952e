"""Proposal network: FP64 oracle pins (CPU) and device parity (GPU).

Pins of the oracle (oracle/mlp.py): its analytic backward pass against
central finite differences in FP64, its bfloat16 rounding against torch's,
rotation invariance of the features about the substrate normal, and the
training-loss gradient (dE/dgain = -eta F(x').P(F0)) against finite
differences of the FP64 force-field energy.  The device forward, backward and
loss gradient are then compared with the oracle on the same inputs.
"""
import numpy as np
import pytest

import oracle
from oracle import mlp
from oracle import precond as pc
from firecore_amd import native

GSTAR, GRIGID = 3.0, 63.0


def _chain_graph(N, rng):
    neighs = -np.ones((N, 4), int)
    for i in range(1, N):
        j = int(rng.integers(0, i))
        while (neighs[j] >= 0).sum() >= 4:
            j = int(rng.integers(0, i))
        neighs[i][list(neighs[i]).index(-1)] = j
        neighs[j][list(neighs[j]).index(-1)] = i
    return neighs


def test_oracle_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    N = 14
    neighs = _chain_graph(N, rng)
    nb = mlp.neighbour_lists(N, neighs)
    x = rng.normal(size=(N, mlp.NIN))
    th = mlp.init_params(1, head_scale=1.0)
    w = rng.normal(size=(N, 2))
    o, cache = mlp.forward(th, x, nb)
    g = mlp.backward(th, cache, w, nb)
    for p in rng.choice(mlp.NPARAM, 40, replace=False):
        e = np.zeros(mlp.NPARAM)
        e[p] = 1e-6
        lp = (mlp.forward(th + e, x, nb)[0] * w).sum()
        lm = (mlp.forward(th - e, x, nb)[0] * w).sum()
        fd = (lp - lm) / 2e-6
        assert abs(fd - g[p]) <= 1e-6 * (1 + abs(fd)), (p, fd, g[p])


def test_training_start_is_the_constant_gain_setting():
    """train_proposal's initial parameters start every gain at the constant
    setting (g0 = 2 per star, 63 per molecule, within 10 %) under the ceilings: the head's
    biases are the logits of g0/gmax, its weights are small."""
    from firecore_amd import train_proposal as tp
    th = tp.initial_params(3, tp.GSTAR, tp.GRIGID).astype(np.float64)
    rng = np.random.default_rng(0)
    N = 30
    x = rng.normal(size=(N, 16))
    nb = [[j for j in (i - 1, i + 1) if 0 <= j < N] for i in range(N)]
    o, _ = mlp.forward(th, x, nb)
    g_node, g_comp = mlp.gains(o, 10, np.zeros(N, np.int64), tp.GSTAR, tp.GRIGID)
    assert np.allclose(g_node, tp.G0_STAR, rtol=0.1) and np.allclose(g_comp[:1], tp.G0_RIGID, rtol=0.1)


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(2)
    v = np.concatenate([rng.normal(size=2000) * 10.0 ** rng.integers(-6, 6, 2000),
                        # exact ties between two bf16 neighbours
                        (np.arange(1, 200) * 2 + 1) * 2.0 ** -8 + 1.0])
    ref = torch.tensor(v.astype(np.float32)).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(mlp.bf16(v), ref)


def test_features_invariant_under_rotation_about_z():
    rng = np.random.default_rng(3)
    N, M = 12, 6
    neighs = _chain_graph(N, rng)
    F, G = rng.normal(size=(N, 3)), rng.normal(size=(N, 3))
    pi = np.zeros((N, 3))
    pi[:M] = rng.normal(size=(M, 3))
    pi[:M] /= np.linalg.norm(pi[:M], axis=1, keepdims=True)
    elem = rng.integers(0, 4, N)
    h = rng.uniform(2, 5, N)
    a = 0.7
    R = np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]])
    x1 = mlp.features(elem, F, G, pi, neighs, M, h)
    x2 = mlp.features(elem, F @ R.T, G @ R.T, pi @ R.T, neighs, M, h)
    assert np.abs(x1 - x2).max() < 1e-12


def test_loss_gradient_matches_finite_differences(orc, nacl_grid):
    """d/dg of E(x0 + eta*P(g)F0) = -eta F(x').P_g(F0), FP64 force field."""
    m = native.gen_workload_system(1, 3, 4242, 16, 40, 0.1, nacl_grid)
    N, M = m["natoms"], m["nnode"]
    eta = 0.02
    f0, _, _, E0 = orc.eval(m, m["apos"], grid=nacl_grid)
    x0, F0 = m["apos"][:N], f0[:N]
    comp = pc.components(N, m["neighs"])
    gs = np.full(M, 1.5)
    gc = np.array([20.0])

    def energy(gs_, gc_):
        Ft = pc.precondition(x0, F0, m["neighs"], M, gs_, gc_, comp)
        pos = m["apos"].copy()
        pos[:N] += eta * Ft
        f, _, _, E = orc.eval(m, pos, grid=nacl_grid)
        return E, f[:N]

    E1, F1 = energy(gs, gc)
    for k in [0, M // 2, M - 1]:
        S = [k] + [int(j) for j in m["neighs"][k] if j >= 0]
        ana = -eta * (F1[S] * pc.rigid_project(x0[S], F0[S], 1e-6)).sum()
        h = 1e-4
        gp, gm = gs.copy(), gs.copy()
        gp[k] += h
        gm[k] -= h
        fd = (energy(gp, gc)[0] - energy(gm, gc)[0]) / (2 * h)
        assert abs(fd - ana) <= 1e-6 + 1e-4 * abs(ana), (k, fd, ana)
    ana = -eta * (F1 * pc.rigid_project(x0, F0, 1e-6)).sum()
    fd = (energy(gs, gc + 1e-3)[0] - energy(gs, gc - 1e-3)[0]) / 2e-3
    assert abs(fd - ana) <= 1e-6 + 1e-4 * abs(ana), (fd, ana)


# ----------------------------------------------------------------------------- GPU
def _setup_batch(nsys, seed, nmin, nmax, theta_seed=5):
    native.set_opt()
    native.init_workload(1, nsys, seed, nmin, nmax, 0.1, "synth:NaCl", 0.1)
    th = mlp.init_params(theta_seed, head_scale=1.0)
    native.set_proposal_weights(th.astype(np.float32))
    native.set_precond(2, GSTAR, GRIGID)
    # the device keeps an fp32 master copy; the oracle rounds the dense weights to bf16 itself
    return th.astype(np.float32).astype(np.float64)


def _systems(B):
    for s in range(B.nsys):
        N, M, a0, n0, v0 = (int(v) for v in B.info[s, :5])
        yield s, N, M, a0, n0, v0


@pytest.mark.gpu
def test_forward_matches_oracle():
    th = _setup_batch(10, 9100, 16, 200)
    try:
        out, feats = native.proposal_forward()
        B = native.Batch()
        worst = 0.0
        for s, N, M, a0, n0, v0 in _systems(B):
            nb = mlp.neighbour_lists(N, B.neighs[a0:a0 + N])
            o_ref, _ = mlp.forward(th, feats[a0:a0 + N].astype(np.float64), nb, bf=True)
            err = np.abs(out[a0:a0 + N] - o_ref).max() / (np.abs(o_ref).max() + 1e-3)
            worst = max(worst, err)
        # bf16 activations: an input that rounds to the other neighbour moves an
        # output by ~2^-8 relative; FP32 accumulation order adds ~1e-6
        assert worst < 2e-2, worst
    finally:
        native.set_precond(0)


@pytest.mark.gpu
def test_features_match_oracle(nacl_grid, orc):
    _setup_batch(6, 9200, 16, 120)
    try:
        _, feats = native.proposal_forward()
        native.eval_precond()  # downloads the exact forces of the same geometry into the host view
        B = native.Batch()
        g = native.grid_info()
        for s, N, M, a0, n0, v0 in _systems(B):
            pos = B.atoms[v0:v0 + N + M, :3].astype(np.float64)
            F = B.aforces[v0:v0 + N, :3].astype(np.float64)
            G = np.zeros((N, 3))
            for i in range(N):
                plq = B.PLQs[a0 + i, :3].astype(np.float64)
                _, G[i] = orc.grid_sample(nacl_grid, pos[i], plq)
            pi = np.zeros((N, 3))
            pi[:M] = pos[N:]
            x = mlp.features(B.elem[a0:a0 + N], F, G, pi, B.neighs[a0:a0 + N], M, pos[:N, 2] - g["pos0"][2])
            # the device feeds bf16(x) to the matrix cores
            assert np.abs(feats[a0:a0 + N] - x).max() < 1e-2 * (1 + np.abs(x).max())
    finally:
        native.set_precond(0)


@pytest.mark.gpu
def test_backward_matches_oracle():
    th = _setup_batch(8, 9300, 16, 180)
    try:
        out, feats = native.proposal_forward()
        B = native.Batch()
        rng = np.random.default_rng(7)
        dout = rng.normal(size=out.shape).astype(np.float32)
        g_dev = native.proposal_backward(dout).astype(np.float64)
        g_ref = np.zeros(mlp.NPARAM)
        for s, N, M, a0, n0, v0 in _systems(B):
            nb = mlp.neighbour_lists(N, B.neighs[a0:a0 + N])
            _, cache = mlp.forward(th, feats[a0:a0 + N].astype(np.float64), nb, bf=True)
            g_ref += mlp.backward(th, cache, dout[a0:a0 + N].astype(np.float64), nb, bf=True)
        g_ref /= B.nsys
        # blocks of the parameter vector: each compared in relative L2 norm
        o = 0
        for name, sh in mlp.PARAM_SHAPES:
            n = int(np.prod(sh))
            a, b = g_dev[o:o + n], g_ref[o:o + n]
            rel = np.linalg.norm(a - b) / (np.linalg.norm(b) + 1e-12)
            assert rel < 2e-2, (name, rel)
            o += n
    finally:
        native.set_precond(0)


@pytest.mark.gpu
def test_loss_gradient_matches_oracle(nacl_grid, orc):
    """Device dloss/doutputs and loss vs the FP64 oracle chain (forces,
    preconditioner, energy at the displaced geometry)."""
    _setup_batch(6, 9400, 16, 100)
    eta, dmax = 0.02, 0.3
    try:
        out, feats = native.proposal_forward()
        dout, loss = native.proposal_loss_grad(eta, dmax)
        B = native.Batch()
        for s, N, M, a0, n0, v0 in _systems(B):
            m = B.system(s)
            f0, _, _, E0 = orc.eval(m, m["apos"], grid=nacl_grid)
            x0, F0 = m["apos"][:N], f0[:N]
            comp = pc.components(N, m["neighs"])
            o = out[a0:a0 + N].astype(np.float64)
            gs, gc = mlp.gains(o, M, comp, GSTAR, GRIGID)
            Ft = pc.precondition(x0, F0, m["neighs"], M, gs, gc, comp)
            d = eta * Ft
            dl = np.linalg.norm(d, axis=1)
            d[dl > dmax] *= (dmax / dl[dl > dmax])[:, None]
            pos = m["apos"].copy()
            pos[:N] += d
            f1, _, _, E1 = orc.eval(m, pos, grid=nacl_grid)
            F1 = f1[:N]
            assert abs(loss[s] - (E1 - E0) / N) < 1e-4 + 1e-3 * abs((E1 - E0) / N)
            dgs = np.array([-eta / N * (F1[S] * pc.rigid_project(x0[S], F0[S], 1e-6)).sum()
                            for S in ([k] + [int(j) for j in m["neighs"][k] if j >= 0] for k in range(M))])
            dgc = np.array([-eta / N * (F1[comp == c] * pc.rigid_project(x0[comp == c], F0[comp == c], 1e-6)).sum()
                            for c in range(int(comp.max()) + 1)])
            ref = mlp.dgains_do(o, M, comp, GSTAR, GRIGID, dgs, dgc)
            scale = np.abs(ref).max(0) + 1e-9
            err = (np.abs(dout[a0:a0 + N] - ref) / scale).max()
            assert err < 1e-2, (s, err)
    finally:
        native.set_precond(0)


def _check_relaxed_against_oracle(orc, B, kind, seed0, nmin, nmax, systems):
    """Every listed system: converged on the device, and for the FP64 oracle
    at the device's final geometry (exact double inputs of the system) max|F|
    over all vecs (atoms and pi-vectors) < 1.01e-3 eV/A and |E_dev - E_oracle|
    < 1e-4 eV."""
    g = native.grid_info()
    worst_f = worst_e = 0.0
    for s in systems:
        s = int(s)
        assert B.status[s] == native.ST_CONVERGED, (s, int(B.status[s]))
        m = native.gen_workload_system(kind, s, seed0, nmin, nmax, 0.1, g)
        m["apos"] = B.positions(s)
        f, _, _, E = orc.eval(m, grid=g)
        fm = np.sqrt((f ** 2).sum(1)).max()
        assert fm < 1.01e-3, (s, m["natoms"], fm)
        assert abs(B.E[s] - E) < 1e-4, (s, m["natoms"], B.E[s], E)
        worst_f, worst_e = max(worst_f, fm), max(worst_e, abs(B.E[s] - E))
    return worst_f, worst_e


@pytest.mark.gpu
@pytest.mark.parametrize("inkernel", [1, 0])
def test_learned_relaxation_converges_to_oracle_minima(orc, inkernel):
    """Random network weights (gains stay within their ceilings): every system
    converges on the exact force, and at the relaxed geometry the FP64 oracle
    finds max|F| < 1.01e-3 eV/A over all vecs and the same energy to 1e-4 eV.
    Network inside the relaxation kernel, and evaluated between chunks of the
    relaxation (graph replay)."""
    _setup_batch(24, 9500, 16, 120)
    old = native.set_learned_inkernel(inkernel)
    try:
        native.set_force_limit(10.0)
        native.set_opt(0.1, 0.02, 0.2, 1.1, 0.5, 0.8, 5)
        native.relax(10000, 1e-3, 0, True, True)
        B = native.Batch()
        assert (np.array(B.status) == native.ST_CONVERGED).all()
        assert (native.last_chunks() > 0) == (inkernel == 0)
        _check_relaxed_against_oracle(orc, B, 1, 9500, 16, 120, range(24))
    finally:
        native.set_learned_inkernel(old)
        native.set_precond(0)
        native.set_force_limit(0.0)
        native.set_opt()


@pytest.mark.gpu
@pytest.mark.parametrize("inkernel", [2, 0])
def test_zero_network_reproduces_constant_gains(inkernel):
    """With every weight and bias zero the network outputs exactly 0, so its
    gains are exactly gstar/2 and grigid/2 (sigma(0) = 1/2): the learned form
    must then reproduce the constant-gain form bit for bit.  With a short
    period this checks every part of the proposal machinery (segment breaks and
    restarts with re-evaluated forces, parked velocities, restaged exclusion
    and REQ rows; chunk suspension and resumption) against a path that has none
    of it, over all four size classes."""
    native.set_opt(0.1, 0.02, 0.2, 1.1, 0.5, 0.8, 5)
    native.set_force_limit(10.0)
    native.init_workload(1, 48, 9700, 16, 256, 0.1, "synth:NaCl", 0.1)
    native.set_proposal_weights(np.zeros(native.proposal_nparam(), np.float32))
    old = native.set_learned_inkernel(inkernel)
    runs = []
    try:
        for mode, gs, gr in [(1, 2.0, 63.0), (2, 4.0, 126.0)]:
            native.set_precond(mode, gs, gr)
            native.set_proposal_every(8)
            native.relax(10000, 1e-3, 0, True, True)
            B = native.Batch()
            pos = np.concatenate([B.positions(s) for s in range(B.nsys)])
            runs.append((np.array(B.status), np.array(B.nsteps), np.array(B.E), pos))
        assert len(native.class_info()) == 4
        assert (native.last_chunks() > 0) == (inkernel == 0)
    finally:
        native.set_learned_inkernel(old)
        native.set_precond(0)
        native.set_force_limit(0.0)
        native.set_opt()
    (st0, ns0, E0, p0), (st1, ns1, E1, p1) = runs
    assert (st0 == native.ST_CONVERGED).all()
    assert (st1 == st0).all() and (ns1 == ns0).all()
    assert np.array_equal(E1, E0) and np.array_equal(p1, p0)


@pytest.mark.gpu
def test_training_lowers_the_loss():
    _setup_batch(64, 9600, 16, 128, theta_seed=11)
    try:
        native.set_opt(0.1, 0.02, 0.2, 1.1, 0.5, 0.8, 5)
        native.set_force_limit(10.0)
        native.set_precond(0)
        native.record_snapshots(50, 40)
        native.relax(2000, 1e-3, 0, True, False)
        native.set_precond(2, GSTAR, GRIGID)
        # the loss on a fixed set of snapshot draws (learning rate 0 leaves the
        # parameters unchanged) before and after training on other draws
        before = native.train_proposal(8, 0.0, 0.02, 0.3, seed=77).mean()
        hist = native.train_proposal(80, 3e-3, 0.02, 0.3, seed=1)
        after = native.train_proposal(8, 0.0, 0.02, 0.3, seed=77).mean()
        assert np.isfinite(hist).all()
        assert after < before, (before, after)
    finally:
        native.record_snapshots(1, 0)
        native.set_precond(0)
        native.set_force_limit(0.0)
        native.set_opt()


def _load_trained():
    import json
    import os
    here = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "firecore_amd")
    meta = json.load(open(os.path.join(here, "proposal_weights.json")))
    return np.load(os.path.join(here, "proposal_weights.npy")), meta


@pytest.mark.gpu
def test_trained_network_matches_oracle_in_every_class(orc):
    """The learned form as benchmarked (trained weights, their proposal period
    and gain ceilings, per-vec clamp 10 eV/A, set_opt defaults) on the first 96
    systems of the 4096-system GridFF workload (seeds 10001.., N ~ U[16,256]):
    all four workgroup classes run, every system converges, and the FP64 oracle
    confirms each relaxed geometry (max|F| < 1.01e-3 eV/A over all vecs,
    |dE| < 1e-4 eV)."""
    theta, meta = _load_trained()
    native.set_opt(0.1, 0.02, 0.2, 1.1, 0.5, 0.8, 5)
    native.set_force_limit(10.0)
    native.init_workload(1, 96, 10001, 16, 256, 0.1, "synth:NaCl", 0.1)
    native.set_proposal_weights(theta)
    native.set_proposal_every(int(meta["every"]))
    native.set_precond(2, meta["gstar"], meta["grigid"])
    try:
        native.relax(10000, 1e-3, 0, True, True)
        B = native.Batch()
        assert {int(c["block"]) for c in native.class_info()} == {64, 128, 192, 256}
        wf, we = _check_relaxed_against_oracle(orc, B, 1, 10001, 16, 256, range(B.nsys))
        print("trained network: worst oracle Fmax %.3e, worst |dE| %.3e" % (wf, we), flush=True)
    finally:
        native.set_precond(0)
        native.set_force_limit(0.0)
        native.set_opt()


@pytest.mark.gpu
def test_trained_network_on_molecule_pairs(orc):
    """The screening workload's systems (pairs of 8-64-atom molecules over the
    GridFF, intermolecular pairs included) relaxed by the learned form with the
    trained weights (period and ceilings of the benchmark, at most 3000 steps):
    each converged system is confirmed by the FP64 oracle (max|F| < 1.01e-3
    eV/A over all vecs, |dE| < 1e-4 eV)."""
    theta, meta = _load_trained()
    native.set_opt(0.1, 0.02, 0.2, 1.1, 0.5, 0.8, 5)
    native.set_force_limit(10.0)
    native.init_workload(2, 64, 20001, 8, 64, 0.1, "synth:NaCl", 0.1)
    native.set_proposal_weights(theta)
    native.set_proposal_every(int(meta["every"]))
    native.set_precond(2, meta["gstar"], meta["grigid"])
    try:
        native.relax(3000, 1e-3, 0, True, True)
        B = native.Batch()
        conv = np.where(np.array(B.status) == native.ST_CONVERGED)[0]
        assert len(conv) >= 0.9 * B.nsys, np.bincount(np.array(B.status), minlength=4)
        _check_relaxed_against_oracle(orc, B, 2, 20001, 8, 64, conv)
    finally:
        native.set_precond(0)
        native.set_force_limit(0.0)
        native.set_opt()


@pytest.mark.gpu
def test_inkernel_and_chunked_learned_agree():
    """Nonzero (trained) weights: the network inside the relaxation kernel and
    the network run between chunks see the same (positions, exact forces) pair
    at every proposal, so both reach the same minima; the first gains of the
    chunked path equal proposal_forward's at the start geometry."""
    theta, meta = _load_trained()
    native.set_opt(0.1, 0.02, 0.2, 1.1, 0.5, 0.8, 5)
    native.set_force_limit(10.0)
    native.init_workload(1, 32, 9800, 16, 200, 0.1, "synth:NaCl", 0.1)
    native.set_proposal_weights(theta)
    native.set_proposal_every(16)
    native.set_precond(2, meta["gstar"], meta["grigid"])
    res = []
    old = native.set_learned_inkernel(1)
    try:
        for ink in (1, 0):
            native.set_learned_inkernel(ink)
            native.relax(10000, 1e-3, 0, True, True)
            B = native.Batch()
            res.append((np.array(B.status), np.array(B.nsteps), np.array(B.E)))
    finally:
        native.set_learned_inkernel(old)
        native.set_precond(0)
        native.set_force_limit(0.0)
        native.set_opt()
    (st1, ns1, E1), (st0, ns0, E0) = res
    assert (st1 == native.ST_CONVERGED).all() and (st0 == native.ST_CONVERGED).all()
    # same minima (trajectories may part in the last bits of the features)
    assert np.median(np.abs(E1 - E0)) < 1e-4, np.abs(E1 - E0)
    assert abs(np.mean(ns1) - np.mean(ns0)) < 0.1 * np.mean(ns1), (ns1.mean(), ns0.mean())
