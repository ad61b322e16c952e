"""FP64 CPU oracle of the multi-system MMFF relaxation path (test infrastructure).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
package, and only to check results or as the timed CPU baseline.  The product
library (firecore_amd/libMMFFmulti_lib.so) never loads it.

Two back-ends with one call signature:
  * ``liboracle.so``      -- this repository's FP64 restatement (oracle.cpp)
  * ``_ref/libfcref.so``  -- the reference's own CPU classes compiled from
    /root/reference by oracle/Makefile (absent where the reference is absent)

Parity pins of the restatement (see tests/test_oracle_*.py): bit-level
agreement with _ref on random molecules (forces, recoils, energies, FIRE
trajectories, back-neighbours), the golden fixtures under tests/golden/
generated from _ref, and closed-form checks (finite differences of the energy,
Newton's third law, B-spline partition of unity and linear reproduction).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libfcref.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def _p(a, t=_dp):
    return None if a is None else a.ctypes.data_as(t)


def build(with_ref=None):
    """Compile liboracle.so (and _ref when the reference tree is present)."""
    import subprocess
    targets = ["all"]
    if with_ref is None:
        with_ref = os.path.isdir("/root/reference/cpp/common")
    if with_ref:
        targets.append("ref")
    subprocess.run(["make", "-s", "-C", HERE] + targets, check=True)


class Backend:
    """ctypes view of liboracle.so (prefix 'orc_') or _ref/libfcref.so ('ref_')."""

    def __init__(self, path, prefix):
        self.lib = ctypes.CDLL(path)
        self.prefix = prefix
        L = self.lib
        f = getattr(L, prefix + "eval")
        f.restype = ctypes.c_double
        f = getattr(L, prefix + "relax")
        f.restype = ctypes.c_int
        f = getattr(L, prefix + "grid_sample")
        f.restype = ctypes.c_double

    def _fn(self, name):
        return getattr(self.lib, self.prefix + name)

    @staticmethod
    def _sys_args(m, sw):
        sw = np.ascontiguousarray(np.asarray(sw, dtype=np.int32))
        return [ctypes.c_int(m["natoms"]), ctypes.c_int(m["nnode"]), _p(m["REQs"]), _p(m["neighs"], _ip),
                _p(m["apars"]), _p(m["bLs"]), _p(m["bKs"]), _p(m["Ksp"]), _p(m["Kpp"]), _p(sw, _ip)], sw

    @staticmethod
    def _grid_args(grid):
        if grid is None:
            return [ctypes.c_int(0)] * 3 + [None, None, None, ctypes.c_double(1.5)]
        return [ctypes.c_int(grid["n"][0]), ctypes.c_int(grid["n"][1]), ctypes.c_int(grid["n"][2]),
                _p(grid["pos0"]), _p(grid["step"]), _p(grid["coefs"]), ctypes.c_double(grid["alpha"])]

    def eval(self, m, apos=None, sw=(1, 1, 1, 1, 1), grid=None):
        """Assembled forces (nvecs,3), recoils (nnode,4,3) x2 and total energy."""
        N, M = m["natoms"], m["nnode"]
        apos = np.ascontiguousarray(m["apos"] if apos is None else apos, dtype=np.float64)
        f = np.zeros((N + M, 3))
        fng = np.zeros((M, 4, 3))
        fngpi = np.zeros((M, 4, 3))
        a, sw_keep = self._sys_args(m, sw)
        args = a[:2] + [_p(apos)] + a[2:] + self._grid_args(grid) + [_p(f), _p(fng), _p(fngpi)]
        E = self._fn("eval")(*args)
        return f, fng, fngpi, E

    def relax(self, m, fire, Fconv, nmax, conv_mode=0, sw=(1, 1, 1, 1, 1), grid=None, ntraj=0):
        """fire = (dt_max, dt_min, damp_max, finc, fdec, falpha, minLastNeg[, Flimit]);
        Flimit (per-vec force clamp) is honoured by the oracle only."""
        N, M = m["natoms"], m["nnode"]
        apos = np.array(m["apos"], dtype=np.float64, copy=True)
        fire = np.zeros(8) + np.r_[np.asarray(fire, dtype=np.float64), np.zeros(8 - len(fire))]
        E = ctypes.c_double(0)
        traj = np.zeros((max(ntraj, 1), N + M, 3))
        a, sw_keep = self._sys_args(m, sw)
        args = a[:2] + [_p(apos)] + a[2:] + self._grid_args(grid) + [
            _p(fire), ctypes.c_double(Fconv), ctypes.c_int(nmax), ctypes.c_int(conv_mode), ctypes.byref(E),
            _p(traj) if ntraj > 0 else None, ctypes.c_int(ntraj)]
        n = self._fn("relax")(*args)
        return n, apos, E.value, traj[:ntraj]

    def explore(self, m, fire, Fconv, nmax, T, gamma, n_explore, n_relax=0, seed=0, sysidx=0, kmax=16,
                sw=(1, 1, 1, 1, 1), grid=None, ntraj=0):
        """Exploring run of one system (relax / Langevin phases, oracle only):
        returns (minima found, final positions, minima energies, minima
        positions, trajectory of the first ntraj steps)."""
        N, M = m["natoms"], m["nnode"]
        apos = np.array(m["apos"], dtype=np.float64, copy=True)
        fire = np.zeros(8) + np.r_[np.asarray(fire, dtype=np.float64), np.zeros(8 - len(fire))]
        minE = np.zeros(kmax)
        minpos = np.zeros((kmax, N + M, 3))
        traj = np.zeros((max(ntraj, 1), N + M, 3))
        a, sw_keep = self._sys_args(m, sw)
        f = self.lib.orc_explore
        f.restype = ctypes.c_int
        args = a[:2] + [_p(apos)] + a[2:] + self._grid_args(grid) + [
            _p(fire), ctypes.c_double(Fconv), ctypes.c_int(nmax), ctypes.c_double(T), ctypes.c_double(gamma),
            ctypes.c_int(n_explore), ctypes.c_int(n_relax), ctypes.c_uint(seed & 0xFFFFFFFF), ctypes.c_int(sysidx),
            ctypes.c_int(kmax), _p(minE), _p(minpos), _p(traj) if ntraj > 0 else None, ctypes.c_int(ntraj)]
        n = f(*args)
        k = max(0, min(n, kmax))
        return n, apos, minE[:k], minpos[:k], traj[:ntraj]

    def drive_xi(self, seed, sysidx, vec, axis, step):
        f = self.lib.orc_drive_xi
        f.restype = ctypes.c_double
        f.argtypes = [ctypes.c_uint] * 5
        return f(seed, sysidx, vec, axis, step)

    def set_surface(self, apos=None, REQs=None, lvec=None, nPBC=(1, 1, 0), alpha=1.5):
        """Explicit substrate atoms (instead of a grid) for the following calls;
        apos None removes them.  Oracle: images -nPBC..nPBC of lvec; reference
        build: ref_set_surface (its own image shifts)."""
        if self.prefix == "ref_":
            f = self.lib.ref_set_surface
            f.restype = None
            if apos is None:
                f(ctypes.c_int(0), None, None, None, None, ctypes.c_double(alpha))
                return
            a = np.ascontiguousarray(apos, np.float64)
            q = np.ascontiguousarray(REQs, np.float64)
            lv = np.ascontiguousarray(lvec, np.float64).reshape(9)
            npb = np.ascontiguousarray(nPBC, np.int32)
            f(ctypes.c_int(len(a)), _p(a), _p(q), _p(lv), _p(npb, _ip), ctypes.c_double(alpha))
            return
        f = self.lib.orc_set_surface
        f.restype = None
        if apos is None:
            f(ctypes.c_int(0), None, None, ctypes.c_int(0), None, ctypes.c_double(alpha))
            return
        a = np.ascontiguousarray(apos, np.float64)
        q = np.ascontiguousarray(REQs, np.float64)
        L = np.asarray(lvec, np.float64).reshape(3, 3)
        nx, ny, nz = (int(v) for v in nPBC)
        sh = np.ascontiguousarray([L[0] * ix + L[1] * iy + L[2] * iz for iz in range(-nz, nz + 1)
                                   for iy in range(-ny, ny + 1) for ix in range(-nx, nx + 1)], np.float64)
        f(ctypes.c_int(len(a)), _p(a), _p(q), ctypes.c_int(len(sh)), _p(sh), ctypes.c_double(alpha))

    def set_pbc(self, natoms=0, shifts=None, neighCell=None):
        """Periodic molecule (image shifts (npbc,3), neighbour images (natoms,4))
        for the following eval/relax/explore calls on systems of `natoms` atoms
        (0 clears).  Oracle only; the reference build takes set_pbc_ref."""
        f = self.lib.orc_set_pbc
        f.restype = None
        if natoms <= 0 or shifts is None:
            f(ctypes.c_int(0), ctypes.c_int(0), None, None)
            return
        sh = np.ascontiguousarray(shifts, np.float64)
        nc = np.ascontiguousarray(neighCell, np.int32)
        f(ctypes.c_int(natoms), ctypes.c_int(len(sh)), _p(sh), _p(nc, _ip))

    def set_pbc_ref(self, lvec=None, nPBC=None):
        """Reference build: periodic molecule (lattice rows, image range) for the
        following calls, its own shifts and neighbour images (None: vacuum)."""
        f = self.lib.ref_set_pbc
        f.restype = None
        if lvec is None:
            f(None, None)
            return
        lv = np.ascontiguousarray(lvec, np.float64).reshape(9)
        npb = np.ascontiguousarray(nPBC, np.int32)
        f(_p(lv), _p(npb, _ip))

    def neigh_cells_ref(self, m, lvec, nPBC):
        """Reference build: (shifts, neighCell) of molecule m."""
        f = self.lib.ref_neigh_cells
        f.restype = ctypes.c_int
        N, M = m["natoms"], m["nnode"]
        apos = np.ascontiguousarray(m["apos"], np.float64)
        ng = np.ascontiguousarray(m["neighs"], np.int32)
        lv = np.ascontiguousarray(lvec, np.float64).reshape(9)
        npb = np.ascontiguousarray(nPBC, np.int32)
        nsh = int(np.prod(2 * npb + 1))
        nc = np.zeros((N, 4), np.int32)
        sh = np.zeros((nsh, 3))
        f(ctypes.c_int(N), ctypes.c_int(M), _p(apos), _p(ng, _ip), _p(lv), _p(npb, _ip), _p(nc, _ip), _p(sh))
        return sh, nc

    def db_reset(self):
        """Reference build: a fresh MolecularDatabase with its default descriptors."""
        self.lib.ref_db_reset.restype = None
        self.lib.ref_db_reset()

    def db_add(self, pos, types):
        """Reference build: MolecularDatabase::addIfNewDescriptor of a
        structure (-1 = new member, else the index it duplicates)."""
        f = self.lib.ref_db_add
        f.restype = ctypes.c_int
        p = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
        t = np.ascontiguousarray(types, np.int32)
        return f(ctypes.c_int(len(p)), _p(p), _p(t, _ip))

    def db_count(self):
        self.lib.ref_db_count.restype = ctypes.c_int
        return self.lib.ref_db_count()

    def db_descriptor(self, pos):
        f = self.lib.ref_db_descriptor
        f.restype = ctypes.c_int
        p = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
        out = np.zeros(8)
        n = f(ctypes.c_int(len(p)), _p(p), _p(out))
        return out[:n]

    def db_distance(self, x, tx, y, ty):
        f = self.lib.ref_db_distance
        f.restype = ctypes.c_double
        x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
        y = np.ascontiguousarray(y, np.float64).reshape(-1, 3)
        return f(ctypes.c_int(len(x)), _p(x), _p(np.ascontiguousarray(tx, np.int32), _ip), _p(y),
                 _p(np.ascontiguousarray(ty, np.int32), _ip))

    def build_xyz_ref(self, res_dir, xyz_noext, nPBC=(1, 1, 0), maxn=4096):
        """Reference build: the reference's own MolWorld_sp3::init of an .xyz
        molecule (run in res_dir, where its common_resources/ live); returns
        (molecule dict with atypes, neighCell, lvec 3x3 or None)."""
        f = self.lib.ref_build_xyz
        f.restype = ctypes.c_int
        apos = np.zeros((2 * maxn, 3))
        REQs = np.zeros((maxn, 4))
        neighs = np.zeros((maxn, 4), np.int32)
        ncell = np.zeros((maxn, 4), np.int32)
        pars = [np.zeros((maxn, 4)) for _ in range(5)]
        atypes = np.zeros(maxn, np.int32)
        lv = np.zeros(9)
        per = ctypes.c_int(0)
        nn = ctypes.c_int(0)
        npb = np.ascontiguousarray(nPBC, np.int32)
        n = f(res_dir.encode(), xyz_noext.encode(), _p(npb, _ip), ctypes.c_int(maxn), ctypes.byref(nn), _p(apos),
              _p(REQs), _p(neighs, _ip), _p(ncell, _ip), *[_p(p) for p in pars], _p(atypes, _ip), _p(lv),
              ctypes.byref(per))
        if n < 0:
            raise RuntimeError("reference build of %s failed" % xyz_noext)
        m = nn.value
        d = dict(natoms=n, nnode=m, apos=apos[:n + m].copy(), REQs=REQs[:n].copy(), neighs=neighs[:n].copy(),
                 neighCell=ncell[:n].copy(), atypes=atypes[:n].copy())
        for k, p in zip(("apars", "bLs", "bKs", "Ksp", "Kpp"), pars):
            d[k] = p[:m].copy()
        return d, (lv.reshape(3, 3) if per.value else None)

    def set_restraints(self, natoms=0, constr=None, constrK=None, bbox=None):
        """Positional restraints / walls of the GPU update kernel for the
        following eval/relax calls on systems of `natoms` atoms (0 clears).
        Only the FP64 restatement has them (the reference CPU path has none)."""
        f = self._fn("set_restraints")
        if natoms <= 0:
            f(ctypes.c_int(0), None, None, None)
            return
        c = np.ascontiguousarray(constr, np.float64)
        k = np.ascontiguousarray(constrK, np.float64)
        b = np.ascontiguousarray(np.zeros(9) if bbox is None else bbox, np.float64).reshape(9)
        f(ctypes.c_int(natoms), _p(c), _p(k), _p(b))

    def make_back_neighs(self, natoms, nnode, neighs):
        neighs = np.ascontiguousarray(neighs, dtype=np.int32)
        bk = np.zeros((natoms, 4), np.int32)
        ng = np.zeros((natoms, 4), np.int32)
        r = self._fn("make_back_neighs")(ctypes.c_int(natoms), ctypes.c_int(nnode), _p(neighs, _ip), _p(bk, _ip),
                                        _p(ng, _ip))
        return r, bk, ng

    def grid_sample(self, grid, p, plq):
        p = np.ascontiguousarray(p, dtype=np.float64)
        plq = np.ascontiguousarray(plq, dtype=np.float64)
        f = np.zeros(3)
        E = self._fn("grid_sample")(ctypes.c_int(grid["n"][0]), ctypes.c_int(grid["n"][1]),
                                    ctypes.c_int(grid["n"][2]), _p(grid["pos0"]), _p(grid["step"]),
                                    _p(grid["coefs"]), _p(p), _p(plq), _p(f))
        return E, f

    def sample_terms(self, kind, k, ang0, r1, r2, angles):
        angles = np.ascontiguousarray(angles, dtype=np.float64)
        n = len(angles)
        E = np.zeros(n)
        F = np.zeros(n)
        self._fn("sample_terms")(ctypes.c_int(kind), ctypes.c_double(k), ctypes.c_double(ang0),
                                 ctypes.c_double(r1), ctypes.c_double(r2), ctypes.c_int(n), _p(angles), _p(E), _p(F))
        return E, F

    def sample_pair(self, rs, REQij):
        rs = np.ascontiguousarray(rs, dtype=np.float64)
        REQij = np.ascontiguousarray(REQij, dtype=np.float64)
        E = np.zeros(len(rs))
        F = np.zeros(len(rs))
        self._fn("sample_pair")(ctypes.c_int(len(rs)), _p(rs), _p(REQij), _p(E), _p(F))
        return E, F


def ref_init_pi(apos, neighs, nnode, Ksp, Kpp, flip=True):
    """The reference's own MMFFsp3_loc::initPi (+ flipPis) through oracle/_ref,
    or None where the reference build is absent."""
    if not os.path.exists(REF_SO):
        return None
    lib = ctypes.CDLL(REF_SO)
    apos = np.ascontiguousarray(apos, np.float64)
    neighs = np.ascontiguousarray(neighs, np.int32)
    out = np.zeros((nnode, 3))
    lib.ref_init_pi(ctypes.c_int(len(neighs)), ctypes.c_int(nnode), _p(apos), _p(neighs, _ip),
                    _p(np.ascontiguousarray(Ksp, np.float64)), _p(np.ascontiguousarray(Kpp, np.float64)),
                    ctypes.c_int(int(flip)), _p(out))
    return out


def ref_gridff_points(apos, REQs, lvec, alpha, nm, nc, pts):
    """The reference's own GridFF node potentials (Pauli, London, bare Coulomb)
    through oracle/_ref at points pts (n,3); None where the build is absent.
    nm, nc: (x, y) image ranges of the Morse and Coulomb sums."""
    if not os.path.exists(REF_SO):
        return None
    lib = ctypes.CDLL(REF_SO)
    apos = np.ascontiguousarray(apos, np.float64)
    REQs = np.ascontiguousarray(REQs, np.float64)
    lvec = np.ascontiguousarray(lvec, np.float64).reshape(9)
    pts = np.ascontiguousarray(pts, np.float64)
    out = np.zeros((len(pts), 3))
    lib.ref_gridff_points(ctypes.c_int(len(apos)), _p(apos), _p(REQs), _p(lvec), ctypes.c_double(alpha),
                          ctypes.c_int(nm[0]), ctypes.c_int(nm[1]), ctypes.c_int(nc[0]), ctypes.c_int(nc[1]),
                          ctypes.c_int(len(pts)), _p(pts), _p(out))
    return out


def oracle():
    return Backend(ORACLE_SO, "orc_")


def reference():
    """The reference build, or None when it was not built here."""
    if not os.path.exists(REF_SO):
        return None
    return Backend(REF_SO, "ref_")


def req2plq(REQ, alpha):
    lib = ctypes.CDLL(ORACLE_SO)
    out = np.zeros(4)
    REQ = np.ascontiguousarray(REQ, dtype=np.float64)
    lib.orc_req2plq(_p(REQ), ctypes.c_double(alpha), _p(out))
    return out


def relax_batch(systems, fire, Fconv, nmax, conv_mode=0, sw=(1, 1, 1, 1, 1), grid=None, nthreads=1,
                precision="f64"):
    """Relax a list of systems on `nthreads` host threads: FP64 (the oracle),
    or precision="f32" for the same restatement with every array and
    operation in single precision (the FP32 CPU baseline, not a parity
    reference).  Returns (steps, E, Fmax, positions list)."""
    lib = ctypes.CDLL(ORACLE_SO)
    ns = len(systems)
    natoms = np.array([m["natoms"] for m in systems], np.int32)
    nnode = np.array([m["nnode"] for m in systems], np.int32)
    cat = lambda k, dt=np.float64: np.ascontiguousarray(np.concatenate([np.asarray(m[k], dt).reshape(-1) for m in systems]))
    apos = cat("apos")
    REQs, neighs = cat("REQs"), cat("neighs", np.int32)
    apars, bLs, bKs, Ksp, Kpp = cat("apars"), cat("bLs"), cat("bKs"), cat("Ksp"), cat("Kpp")
    sw = np.ascontiguousarray(np.asarray(sw, np.int32))
    fire = np.zeros(8) + np.r_[np.asarray(fire, dtype=np.float64), np.zeros(8 - len(fire))]
    steps = np.zeros(ns, np.int32)
    E = np.zeros(ns)
    Fm = np.zeros(ns)
    g = Backend._grid_args(grid)
    fn = lib.orc_relax_batch_f32 if precision == "f32" else lib.orc_relax_batch
    fn(ctypes.c_int(ns), _p(natoms, _ip), _p(nnode, _ip), _p(apos), _p(REQs), _p(neighs, _ip),
       _p(apars), _p(bLs), _p(bKs), _p(Ksp), _p(Kpp), _p(sw, _ip), *g, _p(fire),
       ctypes.c_double(Fconv), ctypes.c_int(nmax), ctypes.c_int(conv_mode), ctypes.c_int(nthreads),
       _p(steps, _ip), _p(E), _p(Fm))
    out = []
    off = 0
    for m in systems:
        V = m["natoms"] + m["nnode"]
        out.append(apos[off * 3:(off + V) * 3].reshape(V, 3))
        off += V
    return steps, E, Fm, out
