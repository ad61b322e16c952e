"""GPU parity of the HIP path against the FP64 oracle (through the C-ABI).

The oracle receives the exact double-precision inputs where the energy is
compared at relaxed geometries: the device stores parameters in FP32, and the
rounded (cos, sin) of the equilibrium angle is not unit-norm, which offsets
the reference's k*(1-cos) angle energy by k*(1-|cs0|) per angle (~2e-4 eV for
250-atom systems); the device uses the equivalent form k*csy^2/(1+csx).

Tolerances (FP32 device vs FP64 oracle):
  forces    max|dF| <= 1e-4 eV/A + 5e-6 * max|F|   (strained start geometries
            carry |F| up to ~10^3 eV/A; at relaxed geometries the bound is
            1e-4 eV/A)
  energies  |dE| <= 1e-4 eV + 1e-6 * |E|
  relaxed   oracle max|F| at the GPU's converged geometry < 1.01e-3 eV/A and
            |E_gpu - E_oracle| < 1e-4 eV at that geometry
"""
import numpy as np
import pytest

from conftest import FIRE_MULTI, FIRE_REF

pytestmark = pytest.mark.gpu


def P(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def nat():
    from firecore_amd import native
    native.lib()
    return native


def fmax(f):
    return np.sqrt((f ** 2).sum(1)).max()


def check_forces(native, orc, B, grid=None, sw=(1, 1, 1, 1, 1), systems=None):
    worst = 0.0
    for s in (range(B.nsys) if systems is None else systems):
        m = B.system(s)
        fo, _, _, Eo = orc.eval(m, sw=sw, grid=grid)
        fg = B.forces(s)
        d = np.abs(fg - fo).max()
        tol = 1e-4 + 5e-6 * np.abs(fo).max()
        assert d <= tol, ("force", s, m["natoms"], d, tol)
        assert abs(B.E[s] - Eo) <= 1e-4 + 1e-6 * abs(Eo), ("energy", s, B.E[s], Eo)
        worst = max(worst, d / tol)
    return worst


def test_forces_vacuum_uneven(nat, orc):
    P("case: vacuum uneven eval")
    mols = [nat.gen_molecule(n, s, 0.1, s + 100) for n, s in [(16, 1), (24, 2), (50, 3), (100, 4), (180, 5), (256, 6)]]
    nat.init_systems(mols)
    nat.eval_forces()
    check_forces(nat, orc, nat.Batch())


def test_forces_gridff(nat, orc):
    P("case: GridFF eval")
    nat.init_workload(1, 12, 10001, 16, 256, 0.1, "synth:NaCl", 0.1)
    nat.eval_forces()
    check_forces(nat, orc, nat.Batch(), grid=nat.grid_info())


@pytest.mark.parametrize("switch,sw", [(dict(Angles=-1), (1, 1, 1, 0, 1)), (dict(PiSigma=-1), (1, 1, 0, 1, 1)),
                                       (dict(PiPiI=-1), (1, 0, 1, 1, 1)), (dict(NonBonded=-1), (1, 1, 1, 1, 0))])
def test_forces_switches(nat, orc, switch, sw):
    from firecore_amd import MMFF_multi as mm
    P("case: switches", switch)
    mols = [nat.gen_molecule(n, s, 0.1, s) for n, s in [(30, 7), (90, 8)]]
    nat.init_systems(mols)
    mm.setSwitches_multi(**switch)
    try:
        nat.eval_forces()
        check_forces(nat, orc, nat.Batch(), sw=sw)
    finally:
        mm.setSwitches_multi(NonBonded=1, Angles=1, PiSigma=1, PiPiI=1)


def ring_molecule():
    """Cyclobutane-like 4-ring with H caps (1-3 pairs reached through two nodes)."""
    a = 1.5 / np.sqrt(2)
    C = np.array([[a, 0, 0], [0, a, 0], [-a, 0, 0], [0, -a, 0]])
    neighs = [[1, 3, 4, 5], [0, 2, 6, 7], [1, 3, 8, 9], [2, 0, 10, 11]]
    pos = list(C)
    for k in range(4):
        out = C[k] / np.linalg.norm(C[k])
        for sz in (1, -1):
            pos.append(C[k] + 1.09 * (0.6 * out + 0.8 * np.array([0, 0, sz])))
            neighs.append([k, -1, -1, -1])
    pis = [[0, 0, 1.0]] * 4
    apos = np.array(pos + pis, np.float64)
    apos[:12] += np.random.default_rng(0).normal(scale=0.05, size=(12, 3))
    REQs = np.array([[1.9255, np.sqrt(0.00455), 0.1, 0]] * 4 + [[1.443, np.sqrt(0.002), -0.05, 0]] * 8)
    apars = np.array([[np.cos(np.radians(109.5) / 2), np.sin(np.radians(109.5) / 2), 25.0, -0.33]] * 4)
    bLs = np.array([[1.5, 1.5, 1.09, 1.09]] * 4)
    bKs = np.array([[10.0, 10.0, 16.0, 16.0]] * 4)
    return dict(natoms=12, nnode=4, apos=apos, REQs=REQs, neighs=np.array(neighs, np.int32), apars=apars, bLs=bLs,
                bKs=bKs, Ksp=np.zeros((4, 4)), Kpp=np.zeros((4, 4)), elem=np.array([0] * 4 + [3] * 8, np.int32))


def test_forces_ring_corrections(nat, orc):
    P("case: ring topology")
    nat.init_systems([ring_molecule(), nat.gen_molecule(20, 3, 0.1, 4)])
    nat.eval_forces()
    check_forces(nat, orc, nat.Batch())


@pytest.mark.parametrize("path", [0, 1])
def test_gridff_trajectory_parity(nat, orc, path):
    """Six systems of the headline workload over the NaCl GridFF follow the
    FP64 trajectory for 40 FIRE steps (per-vec clamp 10 eV/A on both sides);
    on the persistent path the later steps read the per-atom window cache."""
    P("case: GridFF trajectory path", path)
    fire = tuple(FIRE_REF) + (10.0,)
    nat.set_opt(*FIRE_REF)
    nat.set_force_limit(10.0)
    try:
        nat.init_workload(1, 6, 10001, 16, 120, 0.1, "synth:NaCl", 0.1)
        nat.relax(40, 1e-12, path, True, True)
        B = nat.Batch()
        g = nat.grid_info()
        assert (B.nsteps == 40).all()
        for s in range(6):
            m = nat.gen_workload_system(1, s, 10001, 16, 120, 0.1, g)
            n, apos, E, traj = orc.relax(m, fire, 1e-12, 40, 0, grid=g, ntraj=40)
            d = np.abs(B.positions(s) - traj[-1]).max()
            assert d < 2e-4, (s, d)
    finally:
        nat.set_force_limit(0.0)


def test_forces_ring_corrections_persistent(nat, orc):
    """The persistent kernel's forces and energy at the start geometry (a
    zero-step relaxation) on the 4-ring, whose 1-3 pairs reached through two
    nodes go through the weighted pair corrections."""
    P("case: ring topology, persistent kernel")
    nat.set_force_limit(0.0)
    nat.init_systems([ring_molecule(), nat.gen_molecule(20, 3, 0.1, 4)])
    nat.relax(0, 1e-12, 0, True, True)
    B = nat.Batch()
    assert (B.status == nat.ST_MAXSTEP).all() and (B.nsteps == 0).all()
    check_forces(nat, orc, B)


@pytest.mark.parametrize("path", [0, 1])
def test_trajectory_parity_first_steps(nat, orc, path):
    """FP32 device FIRE follows the FP64 trajectory over the first 40 steps."""
    P("case: trajectory path", path)
    mols = [nat.gen_molecule(n, s, 0.1, s + 1) for n, s in [(24, 0), (60, 1), (120, 2)]]
    nat.set_force_limit(0.0)
    nat.set_opt(*FIRE_REF)
    nat.init_systems(mols)
    nat.relax(40, 1e-12, path, True, True)
    B = nat.Batch()
    assert (B.status == nat.ST_MAXSTEP).all() and (B.nsteps == 40).all()
    for s, m in enumerate(mols):
        n, apos, E, traj = orc.relax(m, FIRE_REF, 1e-12, 40, 0, ntraj=40)
        d = np.abs(B.positions(s) - traj[-1]).max()
        assert d < 2e-4, (s, d)


@pytest.mark.parametrize("path", [0, 1])
def test_hbond_trajectory_parity(nat, orc, path):
    """H-bond charges (REQ.w of opposite signs on donors and acceptors) select
    the kernels' H-bond variant of the pair term (getLJQH, Forces.h:341-360);
    its FIRE trajectory follows the FP64 oracle over the first 40 steps."""
    P("case: H-bond trajectory path", path)
    rng = np.random.default_rng(7)
    mols = [nat.gen_molecule(n, s, 0.1, s + 1) for n, s in [(24, 3), (70, 4), (150, 5)]]
    for m in mols:
        R = np.array(m["REQs"], dtype=np.float64).reshape(-1, 4)
        R[:, 3] = rng.choice([0.0, 0.5, -0.5], size=len(R))
        m["REQs"] = R
    nat.set_force_limit(0.0)
    nat.set_opt(*FIRE_REF)
    nat.init_systems(mols)
    nat.relax(40, 1e-12, path, True, True)
    B = nat.Batch()
    assert (B.status == nat.ST_MAXSTEP).all() and (B.nsteps == 40).all()
    for s, m in enumerate(mols):
        n, apos, E, traj = orc.relax(m, FIRE_REF, 1e-12, 40, 0, ntraj=40)
        d = np.abs(B.positions(s) - traj[-1]).max()
        assert d < 2e-4, (s, d)
        # the H-bond term moves the trajectory by far more than the tolerance
        m0 = dict(m, REQs=np.array(m["REQs"]) * np.array([1.0, 1.0, 1.0, 0.0]))
        _, _, _, traj0 = orc.relax(m0, FIRE_REF, 1e-12, 40, 0, ntraj=40)
        assert np.abs(traj0[-1] - traj[-1]).max() > 20 * 2e-4, s


def test_relax_config1_replicas(nat, orc):
    """1024 replicas of one 24-atom molecule, vacuum, FIRE to max|f| < 1e-3."""
    P("case: config-1 relax")
    nat.set_opt(*FIRE_MULTI)
    nat.set_force_limit(0.0)
    nat.init_workload(0, 1024, 0, 24, 24, 0.1, None, 0.1)
    nat.relax(5000, 1e-3, 0, True, True)
    B = nat.Batch()
    assert (B.status == nat.ST_CONVERGED).all(), np.bincount(B.status, minlength=4)
    assert (B.Fmax < 1e-3).all()
    tmpl = nat.gen_molecule(24, 0)  # exact (double) parameters shared by all replicas
    for s in range(0, 1024, 97):
        m = dict(tmpl)
        m["apos"] = B.positions(s)
        fo, _, _, Eo = orc.eval(m)
        assert fmax(fo) < 1.01e-3, (s, fmax(fo))
        assert abs(B.E[s] - Eo) < 1e-4, (s, B.E[s], Eo)


def test_relax_gridff_subset(nat, orc):
    """128 systems of the headline workload: converged systems are converged
    for the FP64 oracle too, and their energies agree at the GPU geometry."""
    P("case: config-2 subset relax")
    nat.set_opt(*FIRE_MULTI)
    nat.set_force_limit(10.0)
    nat.init_workload(1, 128, 10001, 16, 256, 0.1, "synth:NaCl", 0.1)
    nat.relax(10000, 1e-3, 0, True, True)
    B = nat.Batch()
    g = nat.grid_info()
    conv = np.where(B.status == nat.ST_CONVERGED)[0]
    assert len(conv) > 100, np.bincount(B.status, minlength=4)
    for s in conv[::9]:
        # the exact (double) inputs of system s at the GPU's final positions
        m = nat.gen_workload_system(1, int(s), 10001, 16, 256, 0.1, g)
        m["apos"] = B.positions(s)
        fo, _, _, Eo = orc.eval(m, grid=g)
        assert fmax(fo) < 1.01e-3, (s, fmax(fo))
        assert abs(B.E[s] - Eo) < 1e-4, (s, B.E[s], Eo)
    nat.set_force_limit(0.0)


def test_persistent_matches_split(nat):
    P("case: persistent vs split")
    nat.set_opt(*FIRE_MULTI)
    nat.set_force_limit(10.0)
    nat.init_workload(1, 48, 20001, 16, 200, 0.1, "synth:NaCl", 0.1)
    nat.relax(10000, 1e-3, 0, True, True)
    B = nat.Batch()
    st0, E0 = np.array(B.status), np.array(B.E)
    nat.relax(10000, 1e-3, 1, True, True)
    B = nat.Batch()
    st1 = np.array(B.status)
    both = (st0 == nat.ST_CONVERGED) & (st1 == nat.ST_CONVERGED)
    assert both.sum() >= 0.6 * len(st0), (np.bincount(st0, minlength=4), np.bincount(st1, minlength=4))
    assert (st0 == st1).mean() >= 0.8
    # same minima for most systems (FP32 rounding order differs between paths)
    assert np.median(np.abs(E0[both] - np.array(B.E)[both])) < 1e-3
    nat.set_force_limit(0.0)


@pytest.mark.parametrize("form", ["plain", "learned"])
def test_schedules_give_identical_results(nat, form):
    """One workgroup relaxes a system from its first step to its last with the
    same code whatever the scheduling: every scheduler mode, and the autotuned
    queue order of a repeated pass, give bit-identical per-system results."""
    import json
    import os
    P("case: schedulers", form)
    here = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "firecore_amd")
    nat.set_opt(*FIRE_MULTI)
    nat.set_force_limit(10.0)
    nat.init_workload(1, 96, 30001, 16, 256, 0.1, "synth:NaCl", 0.1)
    if form == "learned":
        meta = json.load(open(os.path.join(here, "proposal_weights.json")))
        nat.set_proposal_weights(np.load(os.path.join(here, "proposal_weights.npy")))
        nat.set_proposal_every(16)
        nat.set_precond(2, meta["gstar"], meta["grigid"])
    old = nat.set_sched(1)
    runs = []
    try:
        for mode in (1, 1, 0, 2):
            nat.set_sched(mode)
            nat.relax(10000, 1e-3, 0, True, True)
            B = nat.Batch()
            pos = np.concatenate([B.positions(s) for s in range(B.nsys)])
            runs.append((mode, np.array(B.status), np.array(B.nsteps), np.array(B.E), pos))
    finally:
        nat.set_sched(old)
        nat.set_precond(0)
        nat.set_force_limit(0.0)
    assert len(set(int(c["block"]) for c in nat.class_info())) >= 3
    _, st0, ns0, E0, p0 = runs[0]
    assert (st0 == nat.ST_CONVERGED).sum() >= 80, np.bincount(st0, minlength=4)
    for mode, st, ns, E, p in runs[1:]:
        assert (st == st0).all() and (ns == ns0).all(), mode
        assert np.array_equal(E, E0) and np.array_equal(p, p0), mode


def test_nonfinite_system_is_isolated(nat):
    """A system with two coincident atoms goes non-finite at step 0 and is
    retired as failed without affecting the other systems."""
    P("case: NaN isolation")
    bad = nat.gen_molecule(24, 1, 0.1, 2)
    bad["apos"][5] = bad["apos"][6]
    good = [nat.gen_molecule(24, 3, 0.1, 4), nat.gen_molecule(40, 5, 0.1, 6)]
    nat.set_opt(*FIRE_MULTI)
    for path in (0, 1):
        nat.init_systems([good[0], bad, good[1]])
        nat.relax(5000, 1e-3, path, True, True)
        B = nat.Batch()
        assert B.status[1] == nat.ST_NAN and B.nsteps[1] == 0
        assert B.status[0] == nat.ST_CONVERGED and B.status[2] == nat.ST_CONVERGED


def test_step_limit(nat):
    P("case: step limit")
    nat.init_workload(0, 8, 0, 24, 24, 0.1, None, 0.1)
    nat.relax(5, 1e-3, 0, True, True)
    B = nat.Batch()
    assert (B.status == nat.ST_MAXSTEP).all() and (B.nsteps == 5).all()


def test_term_samplers(nat, orc):
    """sample_* entry points evaluate the device's FP32 terms."""
    P("case: samplers")
    from firecore_amd import MMFF_multi as mm
    ang = np.linspace(0.05, np.pi - 0.05, 25)
    for kind, fn in enumerate((mm.sample_evalPiAling, mm.sample_evalAngleCos, mm.sample_evalAngleCosHalf)):
        E, F = fn(ang, K=1.7, ang0=2.0, r1=1.1, r2=1.3)
        Eo, Fo = orc.sample_terms(kind, 1.7, 2.0, 1.1, 1.3, ang)
        # FP32; near theta=0 the cos(theta/2) form divides by sin(theta/2)
        assert np.abs(E - Eo).max() < 2e-6 and np.abs(F - Fo).max() < 5e-5
    rs = np.linspace(1.8, 8.0, 30)
    REQi, REQj = (1.9, 0.06, 0.2, 0.0), (1.5, 0.05, -0.3, 0.0)
    E, F = mm.sampleNonBond(rs, kind=2, REQi=REQi, REQj=REQj)
    Eo, Fo = orc.sample_pair(rs, np.array([3.4, 0.003, -0.06, 0.0]))
    assert np.abs(E - Eo).max() < 1e-5 * np.abs(Eo).max() and np.abs(F - Fo).max() < 1e-5 * np.abs(Fo).max()


def test_reference_style_init_run(nat, orc):
    """init() / getBuffs() / run() as pyBall.MMFF_multi callers use them."""
    P("case: init/run")
    from firecore_amd import MMFF_multi as mm
    mm.init(nSys_=16, xyz_name="synth:natoms=24,seed=0,sigma=0.1", nPBC=(0, 0, 0))
    mm.getBuffs()
    assert mm.gpu_atoms.shape == (16, mm.nvecs, 4) and mm.nvecs == mm.natoms + mm.nnode
    # the reference's own view of gpu_neighs, (nSys, nvecs, 4), stays inside the padded mirror
    ref_view = mm.getIBuff("gpu_neighs", (16, mm.nvecs, 4)).reshape(-1, 4)
    assert np.array_equal(ref_view[:16 * mm.natoms], mm.gpu_neighs.reshape(-1, 4))
    assert (ref_view[16 * mm.natoms:] == -1).all()
    mm.set_opt()
    outE = np.zeros(16)
    outF = np.zeros(16)
    n = mm.run(5000, Fconv=1e-3, outE=outE, outF=outF, iParalel=2)
    assert 0 < n <= 5000 and (outF < 1e-3).all()
    E0 = mm.eval()
    assert abs(E0 - outE[0]) < 1e-4
    # relaxed geometry of replica 0 is consistent with the oracle
    B = nat.Batch()
    fo, _, _, Eo = orc.eval(B.system(0))
    assert fmax(fo) < 1.01e-3 and abs(Eo - outE[0]) < 1e-4


def test_sample_surf_matches_oracle(nat, orc):
    P("case: sampleSurf")
    from firecore_amd import MMFF_multi as mm
    import oracle
    zs = np.linspace(2.0, 8.0, 20)
    E, Fz = mm.sampleSurf("synth:NaCl", zs, kind=12, atyp=0, Q=0.1, K=1.5, pos0=(0.3, -0.4, 0.0))
    g = nat.grid_info()
    plq = oracle.req2plq(np.array([1.9255, np.sqrt(0.00455323095), 0.1, 0.0]), 1.5)
    for z, e, fz in zip(zs, E, Fz):
        eo, fo = orc.grid_sample(g, np.array([0.3, -0.4, z]), plq)
        assert abs(e - eo) < 1e-4 + 1e-5 * abs(eo) and abs(fz - fo[2]) < 1e-4 + 1e-5 * abs(fo[2])


# ------------------------------------------------------------------ large size classes
def _fire_traj_precond(orc, m, fire, nsteps, grid, gstar, grigid):
    """FP64 FIRE trajectory with the constant-gain preconditioner (oracle/precond.py)
    applied to the exact atom forces before the per-vec clamp, following the
    oracle's relax step (MolWorld_sp3::run_no_omp, MolWorld_sp3.h:2291-2474)."""
    from oracle import precond as pc
    dt_max, dt_min, damp_max, finc, fdec, falpha, minLastNeg, Flimit = fire
    N, M = m["natoms"], m["nnode"]
    comp = pc.components(N, m["neighs"])
    gs = np.full(M, gstar)
    gc = np.full(int(comp.max()) + 1, grigid)
    pos = np.array(m["apos"], np.float64)
    vel = np.zeros_like(pos)
    dt, damping, lastNeg = dt_max, damp_max, 0
    for _ in range(nsteps):
        f = orc.eval(m, pos, grid=grid)[0]
        f[:N] = pc.precondition(pos[:N], f[:N], m["neighs"], M, gs, gc, comp)
        fl = np.sqrt((f ** 2).sum(1))
        big = fl > Flimit
        f[big] *= (Flimit / fl[big])[:, None]
        vf, vv, ff = (vel * f).sum(), (vel * vel).sum(), (f * f).sum()
        flen, vlen = np.sqrt(ff), np.sqrt(vv)
        cos_vf = vf / (flen * vlen + 1e-32)
        renorm = vlen / (flen + 1e-32)
        if cos_vf < 0:
            cv, cf = 0.0, 0.0
            dt, lastNeg, damping = max(dt * fdec, dt_min), 0, damp_max
        else:
            cv, cf = 1 - damping, damping
            if lastNeg > minLastNeg:
                dt, damping = min(dt * finc, dt_max), damping * falpha
            lastNeg += 1
        cf *= renorm
        dtf = float(np.float32(dt))
        p = pos[N:]
        f[N:] -= p * (p * f[N:]).sum(1, keepdims=True)
        vel = vel * cv + f * (dtf + renorm * cf)
        vel[N:] -= p * (p * vel[N:]).sum(1, keepdims=True)
        pos = pos + vel * dtf
        pos[N:] /= np.linalg.norm(pos[N:], axis=1, keepdims=True)
    return pos


@pytest.mark.parametrize("form", ["plain", "constant"])
def test_large_class_trajectory_parity(nat, orc, form):
    """Systems of 150, 200 and 256 atoms (the 192- and 256-thread workgroups,
    whose windowed reaction rows wrap around the ring) follow the FP64
    trajectory step by step for 40 FIRE steps on the GridFF, with the per-atom
    window cache, plain and with the constant-gain preconditioner."""
    P("case: large-class trajectory", form)
    fire = tuple(FIRE_REF) + (10.0,)
    nat.set_opt(*FIRE_REF)
    nat.set_force_limit(10.0)
    try:
        # systems of fixed sizes (nmin = nmax = N) placed over the substrate
        nat.init_workload(1, 1, 10001, 16, 16, 0.1, "synth:NaCl", 0.1)
        g = nat.grid_info()
        mols = [nat.gen_workload_system(1, k, 10101, n, n, 0.1, g) for k, n in enumerate((150, 200, 256))]
        nat.init_systems(mols)
        _attach_grid(nat)
        if form == "constant":
            nat.set_precond(1, 2.0, 63.0)
        nat.relax(40, 1e-12, 0, True, True)
        B = nat.Batch()
        assert (B.nsteps == 40).all()
        assert {int(c["block"]) for c in nat.class_info()} == {192, 256}
        for s, m in enumerate(mols):
            if form == "plain":
                _, _, _, traj = orc.relax(m, fire, 1e-12, 40, 0, grid=g, ntraj=40)
                ref = traj[-1]
            else:
                ref = _fire_traj_precond(orc, m, fire, 40, g, 2.0, 63.0)
            d = np.abs(B.positions(s) - ref).max()
            assert d < 2e-4, (s, m["natoms"], d)
    finally:
        nat.set_precond(0)
        nat.set_force_limit(0.0)


def _attach_grid(nat):
    """Attach the synthetic NaCl GridFF to the current batch (init_systems
    batches start without a substrate)."""
    nat.attach_surface("synth:NaCl", 0.1)


def test_grid_z_clamp_edge_on_device(nat, orc):
    """Atoms above the top of the GridFF (iz clamped to nz-5 while tz keeps the
    unclamped index, Bspline.h:759) on the persistent kernel: the first step
    gathers the clamped window, the later ones read it from the window cache;
    the 20-step trajectory follows the FP64 oracle."""
    P("case: z-clamp edge")
    fire = tuple(FIRE_REF) + (10.0,)
    nat.set_opt(*FIRE_REF)
    nat.set_force_limit(10.0)
    try:
        nat.init_workload(1, 1, 10001, 40, 40, 0.1, "synth:NaCl", 0.1)
        g = nat.grid_info()
        top = g["pos0"][2] + g["n"][2] * g["step"][2]
        mols = []
        for k, n in enumerate((24, 40)):
            m = nat.gen_workload_system(1, k, 10201, n, n, 0.1, g)
            a = np.array(m["apos"])
            # centre the molecule 0.3 A below the top plane: its atoms straddle z = nz-5 and beyond
            a[:n, 2] += (top - 0.3) - a[:n, 2].mean()
            m["apos"] = a
            mols.append(m)
        nat.init_systems(mols)
        _attach_grid(nat)
        nat.relax(20, 1e-12, 0, True, True)
        B = nat.Batch()
        for s, m in enumerate(mols):
            N = m["natoms"]
            uz = (np.array(m["apos"])[:N, 2] - g["pos0"][2]) / g["step"][2]
            assert (uz >= g["n"][2] - 4).any(), "no atom in the clamped band"
            _, _, _, traj = orc.relax(m, fire, 1e-12, 20, 0, grid=g, ntraj=20)
            d = np.abs(B.positions(s) - traj[-1]).max()
            assert d < 2e-4, (s, d)
    finally:
        nat.set_force_limit(0.0)


def test_sample_surf_z_edges(nat, orc):
    """Device B-spline samples below iz=3 and above iz=nz-5 (the clamped bands)."""
    P("case: sampleSurf z edges")
    from firecore_amd import MMFF_multi as mm
    import oracle
    g0 = None
    nat.init_workload(1, 1, 10001, 16, 16, 0.1, "synth:NaCl", 0.1)
    g0 = nat.grid_info()
    H = g0["n"][2] * g0["step"][2]
    zs = np.concatenate([np.linspace(-0.4, 0.45, 9), np.linspace(H - 0.7, H + 0.5, 9)])
    E, Fz = mm.sampleSurf("synth:NaCl", zs, kind=12, atyp=0, Q=0.1, K=1.5, pos0=(0.3, -0.4, 0.0))
    g = nat.grid_info()
    plq = oracle.req2plq(np.array([1.9255, np.sqrt(0.00455323095), 0.1, 0.0]), 1.5)
    for z, e, fz in zip(zs, E, Fz):
        eo, fo = orc.grid_sample(g, np.array([0.3, -0.4, z]), plq)
        assert abs(e - eo) < 1e-4 + 2e-6 * abs(eo) and abs(fz - fo[2]) < 1e-4 + 2e-6 * abs(fo[2]), (z, e, eo, fz, fo)


def test_oversize_system_relaxed_by_split_fallback(nat, orc):
    """A system of ~300 atoms fits no workgroup class of the persistent kernel;
    in a mixed batch it is relaxed by the split kernels after the persistent
    launch, and its result (status, Fmax, forces, energy) is checked against
    the FP64 oracle like every other system's."""
    P("case: oversize system")
    nat.set_opt(*FIRE_MULTI)
    nat.set_force_limit(10.0)
    try:
        mols = [nat.gen_molecule(n, s, 0.1, s + 50) for n, s in [(24, 31), (300, 32), (80, 33)]]
        nat.init_systems(mols)
        nat.relax(20000, 1e-3, 0, True, True)
        B = nat.Batch()
        assert (B.status == nat.ST_CONVERGED).all(), B.status
        assert 0 < B.nsteps[1] < 20000
        for s, m in enumerate(mols):
            fo, _, _, Eo = orc.eval(m, B.positions(s))
            assert fmax(fo) < 1.01e-3, (s, fmax(fo))
            assert abs(B.Fmax[s] - fmax(fo)) < 1e-4, (s, B.Fmax[s], fmax(fo))
            assert abs(B.E[s] - Eo) < 1e-4 + 1e-6 * abs(Eo), (s, B.E[s], Eo)
            assert np.abs(B.forces(s) - fo).max() < 1e-4, s
    finally:
        nat.set_force_limit(0.0)


# ------------------------------------------------------------------ screening workload (molecule pairs)
def test_pair_systems_forces(nat, orc):
    """Molecule-pair systems of the screening workload (two molecules of 8-64
    atoms each, random relative orientation, 4-8 A apart, over the GridFF):
    forces and energies, including every intermolecular LJ+Coulomb pair, on
    both kernel paths (split evaluation and a zero-step persistent pass)."""
    P("case: molecule-pair forces")
    nat.set_force_limit(0.0)
    nat.init_workload(2, 16, 20001, 8, 64, 0.1, "synth:NaCl", 0.1)
    g = nat.grid_info()
    nat.eval_forces()
    B = nat.Batch()
    ncomp = [len(set(nat.iview("gpu_comp", B.sizes["ntotA"])[B.info[s, 2]:B.info[s, 2] + B.info[s, 0]]))
             for s in range(B.nsys)]
    assert all(c == 2 for c in ncomp), ncomp
    check_forces(nat, orc, B, grid=g)
    # the intermolecular pairs matter at this separation
    m = nat.gen_workload_system(2, 0, 20001, 8, 64, 0.1, g)
    comp = nat.iview("gpu_comp", B.sizes["ntotA"])[:m["natoms"]]
    far = np.array(m["apos"])
    far[:m["natoms"]][comp == 1] += np.array([0.0, 0.0, 500.0])
    f_near = orc.eval(m, grid=None)[0]
    f_far = orc.eval(m, far, grid=None)[0]
    assert np.abs(f_near - f_far).max() > 1e-3
    nat.relax(0, 1e-12, 0, True, True)
    check_forces(nat, orc, nat.Batch(), grid=g)


# ------------------------------------------------------------------ GridFF generation, pi-vectors
def test_gridff_device_nodes_match_host(nat, nacl_grid):
    """The FP64 device kernel of the GridFF node potentials reproduces the host
    definition (pinned against the reference's evalGridFFPoint_Mors/_Coul on
    the CPU), and the grid that init builds with it (device nodes, host fit)
    is the host-built grid."""
    import ctypes
    P("case: GridFF generation on the device")
    L = nat.lib()
    L.gridff_nacl_nodes.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                                          ctypes.c_void_p]
    L.gridff_nacl_nodes.restype = ctypes.c_int
    nx, ny, nz, dz = 28, 28, 60, 0.1
    for rdamp in (0.0, 1.0):  # bare (reference) and damped (benchmark substrate) Coulomb channel
        Vh = np.zeros((nx, ny, nz, 3))
        Vd = np.zeros((nx, ny, nz, 3))
        assert L.gridff_nacl_nodes(nx, ny, nz, dz, rdamp, 0, Vh.ctypes.data) == 0
        assert L.gridff_nacl_nodes(nx, ny, nz, dz, rdamp, 1, Vd.ctypes.data) == 0
        assert np.abs(Vd - Vh).max() < 1e-12 * (1 + np.abs(Vh).max()), (rdamp, np.abs(Vd - Vh).max())
    nat.init_workload(1, 2, 10001, 16, 32, 0.1, "synth:NaCl", 0.1)
    g = nat.grid_info()
    assert g["n"] == nacl_grid["n"]
    assert np.allclose(g["pos0"], nacl_grid["pos0"]) and np.allclose(g["step"], nacl_grid["step"])
    c0, c1 = np.asarray(g["coefs"]), nacl_grid["coefs"]
    assert np.abs(c0 - c1).max() < 1e-9 * (1 + np.abs(c1).max())


def test_init_systems_derives_missing_pi_vectors(nat):
    """Systems given without pi-vectors (zero rows) get them from the geometry
    (MMFFsp3_loc::initPi + flipPis) before upload."""
    P("case: init_systems pi-vectors")
    mols = [nat.gen_molecule(n, s, 0.1, s + 3) for n, s in [(30, 41), (90, 42)]]
    for m in mols:
        m["apos"] = np.array(m["apos"])
        m["apos"][m["natoms"]:] = 0.0
    nat.init_systems(mols)
    B = nat.Batch()
    for s, m in enumerate(mols):
        N = m["natoms"]
        exp = nat.init_pi_vectors(m["apos"][:N], m["neighs"], m["nnode"], m["Ksp"], m["Kpp"])
        assert np.abs(B.positions(s)[N:] - exp).max() < 1e-6, s
