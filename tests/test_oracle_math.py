"""Closed-form and invariant checks of the FP64 oracle: they fail on a dropped
term, a wrong sign or a transposed index independently of the reference."""
import numpy as np
import pytest

from firecore_amd import native


def _energy(orc, m, apos, sw=(1, 1, 1, 1, 1), grid=None):
    return orc.eval(m, apos=apos, sw=sw, grid=grid)[3]


@pytest.mark.parametrize("sw", [(1, 1, 1, 1, 1), (1, 0, 0, 0, 0), (0, 1, 0, 0, 0), (0, 0, 1, 0, 0), (0, 0, 0, 1, 0),
                                (0, 0, 0, 0, 1), (0, 0, 0, 1, 1)])
def test_forces_are_minus_energy_gradient(orc, sw):
    """F = -dE/dx for atoms (central differences); for pi-vectors the
    assembled force is the tangential projection of -dE/dpi."""
    m = native.gen_molecule(20, 3, 0.1, 4)
    N, M = m["natoms"], m["nnode"]
    f = orc.eval(m, sw=sw)[0]
    h = 1e-6
    rng = np.random.default_rng(1)
    for i in list(range(N)) + list(rng.choice(np.arange(N, N + M), 3, replace=False)):
        g = np.zeros(3)
        for k in range(3):
            ap = m["apos"].copy()
            ap[i, k] += h
            Ep = _energy(orc, m, ap, sw)
            ap[i, k] -= 2 * h
            Em = _energy(orc, m, ap, sw)
            g[k] = -(Ep - Em) / (2 * h)
        if i >= N:  # pi: compare tangential parts
            p = m["apos"][i] / np.linalg.norm(m["apos"][i])
            g = g - p * p.dot(g)
        assert np.abs(g - f[i]).max() < 2e-5 * max(1.0, np.abs(f[i]).max()), (i, g, f[i])


def test_grid_forces_are_minus_energy_gradient(orc, nacl_grid):
    g = nacl_grid
    rng = np.random.default_rng(5)
    for _ in range(20):
        p = np.array([rng.uniform(-4, 4), rng.uniform(-4, 4), rng.uniform(1.5, 7.0)])
        q = np.array([rng.uniform(5, 40), rng.uniform(0.1, 2), rng.uniform(-0.5, 0.5), 0.0])
        E, f = orc.grid_sample(g, p, q)
        h = 1e-5
        for k in range(3):
            dp = np.zeros(3)
            dp[k] = h
            Ep = orc.grid_sample(g, p + dp, q)[0]
            Em = orc.grid_sample(g, p - dp, q)[0]
            assert abs(-(Ep - Em) / (2 * h) - f[k]) < 1e-6 * max(1, abs(f[k]))


def test_newton_third_law_vacuum(orc):
    """Energy is translation invariant in vacuum: atom forces sum to zero."""
    for n, s in [(24, 1), (80, 2)]:
        m = native.gen_molecule(n, s, 0.1, s)
        f = orc.eval(m)[0]
        assert np.abs(f[:n].sum(0)).max() < 1e-10 * max(1, np.abs(f).max())


def test_bspline_reproduces_constant_and_linear_fields(orc):
    """Cubic B-splines sum to one and reproduce linear coefficient fields
    exactly: constant -> E=c, F=0; c_i = a*x_i -> E=a*x, F=-a."""
    nx, ny, nz = 10, 12, 30
    step = np.array([0.3, 0.25, 0.2])
    pos0 = np.array([-1.0, -1.5, 0.0])
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    rng = np.random.default_rng(2)
    # constant field in the three channels
    C = np.zeros((nx, ny, nz, 3))
    C[...] = [0.7, -1.3, 2.1]
    g = dict(n=(nx, ny, nz), pos0=pos0, step=step, alpha=1.5, coefs=C)
    plq = np.array([1.5, 0.4, -0.3, 0])
    for _ in range(20):
        p = pos0 + step * [rng.uniform(0, nx), rng.uniform(0, ny), rng.uniform(4, nz - 6)]
        E, f = orc.grid_sample(g, p, plq)
        assert abs(E - plq[:3].dot([0.7, -1.3, 2.1])) < 1e-12
        assert np.abs(f).max() < 1e-12
    # linear in z (non-periodic axis): node value = a*(iz - 1)... the sample at u
    # reproduces a*u exactly for every interior u
    a = 0.37
    C = np.zeros((nx, ny, nz, 3))
    C[..., 0] = a * iz
    g["coefs"] = C
    for _ in range(20):
        u = np.array([rng.uniform(0, nx), rng.uniform(0, ny), rng.uniform(4, nz - 6)])
        E, f = orc.grid_sample(g, pos0 + step * u, np.array([1.0, 0, 0, 0]))
        assert abs(E - a * u[2]) < 1e-12
        assert abs(f[2] + a / step[2]) < 1e-11 and abs(f[0]) < 1e-12 and abs(f[1]) < 1e-12


def test_bspline_periodic_wrap(orc):
    """x/y are periodic: shifting by a full cell gives the same sample."""
    nx, ny, nz = 8, 9, 20
    step = np.array([0.5, 0.4, 0.25])
    rng = np.random.default_rng(3)
    g = dict(n=(nx, ny, nz), pos0=np.zeros(3), step=step, alpha=1.5, coefs=rng.normal(size=(nx, ny, nz, 3)))
    plq = np.array([0.3, 1.2, -0.7, 0])
    for _ in range(20):
        p = np.array([rng.uniform(-3, 6), rng.uniform(-3, 6), rng.uniform(1.0, 3.5)])
        E0, f0 = orc.grid_sample(g, p, plq)
        for sh in ([nx * step[0], 0, 0], [0, ny * step[1], 0], [-2 * nx * step[0], 3 * ny * step[1], 0]):
            E1, f1 = orc.grid_sample(g, p + np.array(sh), plq)
            assert abs(E1 - E0) < 1e-9 and np.abs(f1 - f0).max() < 1e-9


def test_bspline_brute_force_triple_sum(orc):
    """Interior sample == explicit 64-tap sum over basis products."""
    nx, ny, nz = 7, 6, 16
    step = np.array([0.3, 0.35, 0.2])
    rng = np.random.default_rng(4)
    C = rng.normal(size=(nx, ny, nz, 3))
    g = dict(n=(nx, ny, nz), pos0=np.zeros(3), step=step, alpha=1.5, coefs=C)
    plq = np.array([0.9, -0.2, 0.5, 0])

    def B(t):
        return np.array([(1 - t) ** 3, 3 * t ** 3 - 6 * t ** 2 + 4, -3 * t ** 3 + 3 * t ** 2 + 3 * t + 1, t ** 3]) / 6

    for _ in range(10):
        u = np.array([rng.uniform(0, nx), rng.uniform(0, ny), rng.uniform(4, nz - 6)])
        i = np.floor(u).astype(int)
        t = u - i
        E = 0.0
        for a in range(4):
            for b in range(4):
                for c in range(4):
                    E += B(t[0])[a] * B(t[1])[b] * B(t[2])[c] * C[(i[0] - 1 + a) % nx, (i[1] - 1 + b) % ny,
                                                                   i[2] - 1 + c].dot(plq[:3])
        assert abs(orc.grid_sample(g, u * step, plq)[0] - E) < 1e-12


def test_fire_first_step_closed_form(orc):
    """From rest the first FIRE move is p1 = p0 + dt^2 f0 (v1 = dt f0; cos_vf = 0
    keeps cv = 1-damp but v0 = 0), with dt passed as float."""
    m = native.gen_molecule(16, 2, 0.1, 3)
    N = m["natoms"]
    f0 = orc.eval(m)[0]
    fire = (0.05, 0.005, 0.1, 1.1, 0.5, 0.8, 3)
    n, apos, E, traj = orc.relax(m, fire, 1e-12, 2, 0, ntraj=1)
    dt = float(np.float32(0.05))
    expect = m["apos"][:N] + dt * dt * f0[:N]
    assert np.abs(traj[0][:N] - expect).max() < 1e-13


def test_fire_state_machine_on_quadratic():
    """FIRE_update_params on hand-traced inputs: velocity kill on cos<0, dt
    growth after minLastNeg positive steps."""
    # dt_max=0.1, dt_min=0.01, damp=0.1: five positive steps with minLastNeg=3
    dt, damp, lastNeg = 0.05, 0.1, 0
    seq = []
    for k in range(6):
        cos_pos = True
        if lastNeg > 3:
            dt = min(dt * 1.1, 0.1)
            damp *= 0.8
        lastNeg += 1
        seq.append((round(dt, 6), round(damp, 6)))
    assert seq[3] == (0.05, 0.1) and seq[4] == (0.055, 0.08) and seq[5] == (0.0605, 0.064)
