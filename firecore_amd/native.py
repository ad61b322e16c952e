"""Loader and typed bindings of libMMFFmulti_lib.so (the HIP library).

The library is built in-tree by ``firecore_amd.build``; loading fails loudly
when it is missing -- there is no Python or CPU fallback for any compute.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libMMFFmulti_lib.so")

c_int, c_double, c_bool, c_char_p, c_void_p, c_longlong = (ctypes.c_int, ctypes.c_double, ctypes.c_bool,
                                                           ctypes.c_char_p, ctypes.c_void_p, ctypes.c_longlong)
c_double_p = ctypes.POINTER(ctypes.c_double)
c_float_p = ctypes.POINTER(ctypes.c_float)
c_int_p = ctypes.POINTER(ctypes.c_int)

_lib = None

# status codes written per system
ST_RUNNING, ST_CONVERGED, ST_MAXSTEP, ST_NAN = 0, 1, 2, 3
# system-info record (int32 x 8): natoms, nnode, a0, n0, v0, mask0, corr0, ncorr
SYSINFO_FIELDS = ("natoms", "nnode", "a0", "n0", "v0", "mask0", "corr0", "ncorr")


# ints per system of the "sys_info" buffer (SysInfo: natoms, nnode, a0, n0, v0,
# mask0, corr0, ncorr, sh0, npbc)
SYSINFO_INTS = 10


def _p(a, t=c_double_p):
    return None if a is None else a.ctypes.data_as(t)


def lib():
    """The loaded library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libMMFFmulti_lib.so is not built (run `python -m firecore_amd.build`); "
                           "this package has no CPU fallback")
    L = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_LOCAL)
    L.init_workload.argtypes = [c_int, c_int, c_longlong, c_int, c_int, c_double, c_char_p, c_double]
    L.init_workload.restype = c_int
    L.init_systems.argtypes = [c_int, c_int_p, c_int_p, c_double_p, c_double_p, c_int_p, c_double_p, c_double_p,
                               c_double_p, c_double_p, c_double_p, c_int_p]
    L.init_systems.restype = c_int
    L.relax.argtypes = [c_int, c_double, c_int, c_bool, c_bool]
    L.relax_noalloc.argtypes = [c_int, c_double, c_bool, c_bool]
    L.relax_noalloc.restype = c_int
    L.relax.restype = c_int
    L.get_last_ms.restype = c_double
    L.eval_forces.restype = c_int
    L.set_conv_mode.argtypes = [c_int]
    L.get_batch_sizes.argtypes = [c_int_p]
    L.get_grid_info.argtypes = [c_double_p]
    L.get_grid_info.restype = c_int
    L.get_grid_coefs.restype = c_double_p
    L.gen_molecule_arrays.argtypes = [c_int, c_longlong, c_double, c_longlong, c_int_p] + [c_void_p] * 9
    L.gen_molecule_arrays.restype = c_int
    L.make_back_neighs.argtypes = [c_int, c_int, c_int_p, c_int_p]
    L.make_back_neighs.restype = c_int
    L.make_exclusions.argtypes = [c_int, c_int, c_int_p, c_int, ctypes.POINTER(ctypes.c_uint), c_int_p, c_int_p]
    L.make_exclusions.restype = c_int
    L.make_nacl_gridff.argtypes = [c_int, c_int, c_int, c_double, c_double, c_int, c_double_p, c_double_p]
    L.make_nacl_gridff.restype = c_int
    L.getfBuff.argtypes = [c_char_p]
    L.getfBuff.restype = c_float_p
    L.getIBuff.argtypes = [c_char_p]
    L.getIBuff.restype = c_int_p
    L.getBuff.argtypes = [c_char_p]
    L.getBuff.restype = c_double_p
    L.set_opt.argtypes = [c_double, c_double, c_double, c_double, c_double, c_double, c_int, c_double, c_double]
    L.set_opt.restype = None
    L.setSwitches_multi.argtypes = [c_int] * 9
    L.setSwitches_multi.restype = None
    L.attach_surface.argtypes = [c_char_p, c_double]
    L.attach_surface.restype = c_int
    _lib = L
    return L


# ------------------------------------------------------------ host helpers
def gen_molecule(natoms, seed, sigma=0.0, pseed=0):
    """Synthetic molecule (host generator; no device work) as a dict of arrays
    in the MMFFsp3_loc layout."""
    L = lib()
    n = natoms
    nn = c_int(0)
    apos = np.zeros((2 * n, 3))
    REQs = np.zeros((n, 4))
    neighs = np.zeros((n, 4), np.int32)
    pars = [np.zeros((n, 4)) for _ in range(5)]
    elem = np.zeros(n, np.int32)
    r = L.gen_molecule_arrays(n, seed, sigma, pseed, ctypes.byref(nn), apos.ctypes.data, REQs.ctypes.data,
                              neighs.ctypes.data, *[p.ctypes.data for p in pars], elem.ctypes.data)
    if r < 0:
        raise ValueError("cannot generate a %d-atom molecule" % natoms)
    m = nn.value
    d = dict(natoms=n, nnode=m, apos=np.ascontiguousarray(apos[:n + m]), REQs=REQs, neighs=neighs, elem=elem)
    for k, p in zip(("apars", "bLs", "bKs", "Ksp", "Kpp"), pars):
        d[k] = np.ascontiguousarray(p[:m])
    return d


def gen_workload_system(kind, i, seed0, nmin, nmax, sigma=0.1, grid=None):
    """System i of init_workload(kind, ...) generated on the host (same arrays
    the device batch receives, in double precision)."""
    L = lib()
    if not hasattr(L, "_gw"):
        L.gen_workload_arrays.argtypes = [c_int, c_int, c_longlong, c_int, c_int, c_double, c_double_p, c_double_p,
                                          c_int_p, c_int_p] + [c_void_p] * 9
        L.gen_workload_arrays.restype = c_int
        L._gw = True
    nmx = nmax * 2 + 8
    N, M = c_int(0), c_int(0)
    apos = np.zeros((2 * nmx, 3))
    REQs = np.zeros((nmx, 4))
    neighs = np.zeros((nmx, 4), np.int32)
    pars = [np.zeros((nmx, 4)) for _ in range(5)]
    elem = np.zeros(nmx, np.int32)
    pos0 = cell = None
    if grid is not None:
        pos0 = np.ascontiguousarray(grid["pos0"], np.float64)
        cell = np.array([grid["n"][0] * grid["step"][0], grid["n"][1] * grid["step"][1]])
    r = L.gen_workload_arrays(kind, i, seed0, nmin, nmax, sigma, _p(pos0), _p(cell), ctypes.byref(N),
                              ctypes.byref(M), apos.ctypes.data, REQs.ctypes.data, neighs.ctypes.data,
                              *[p.ctypes.data for p in pars], elem.ctypes.data)
    if r < 0:
        raise ValueError("cannot generate workload system")
    n, m = N.value, M.value
    d = dict(natoms=n, nnode=m, apos=np.ascontiguousarray(apos[:n + m]), REQs=np.ascontiguousarray(REQs[:n]),
             neighs=np.ascontiguousarray(neighs[:n]), elem=np.ascontiguousarray(elem[:n]))
    for k, p in zip(("apars", "bLs", "bKs", "Ksp", "Kpp"), pars):
        d[k] = np.ascontiguousarray(p[:m])
    return d


def make_back_neighs(natoms, nnode, neighs):
    neighs = np.ascontiguousarray(neighs, np.int32)
    bk = np.zeros((natoms, 4), np.int32)
    r = lib().make_back_neighs(natoms, nnode, _p(neighs, c_int_p), _p(bk, c_int_p))
    return r, bk


def init_pi_vectors(apos, neighs, nnode, Ksp, Kpp, flip=True):
    """Initial pi-vectors (nnode,3) of a molecule from its atom positions."""
    apos = np.ascontiguousarray(apos, np.float64)
    neighs = np.ascontiguousarray(neighs, np.int32)
    natoms = len(neighs)
    Ksp = np.ascontiguousarray(Ksp, np.float64)
    Kpp = np.ascontiguousarray(Kpp, np.float64)
    out = np.zeros((nnode, 3))
    L = lib()
    L.init_pi_vectors.argtypes = [c_int, c_int, c_double_p, c_int_p, c_double_p, c_double_p, c_int, c_double_p]
    L.init_pi_vectors.restype = c_int
    if L.init_pi_vectors(natoms, nnode, _p(apos), _p(neighs, c_int_p), _p(Ksp), _p(Kpp), int(flip), _p(out)) != 0:
        raise ValueError("init_pi_vectors: bad input")
    return out


def make_exclusions(natoms, nnode, neighs, subtract13=True):
    neighs = np.ascontiguousarray(neighs, np.int32)
    W = (natoms + 31) // 32
    mask = np.zeros((natoms, W), np.uint32)
    w = c_int(0)
    nc = c_int(0)
    lib().make_exclusions(natoms, nnode, _p(neighs, c_int_p), int(subtract13),
                          mask.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)), ctypes.byref(w), ctypes.byref(nc))
    bits = np.unpackbits(mask.view(np.uint8), bitorder="little").reshape(natoms, W * 32)[:, :natoms]
    return bits.astype(bool), nc.value


def build_xyz(xyz, tables_dir, maxn=4096):
    """Typed molecule of an .xyz file (the builder of init(xyz_name)) with the
    type tables of `tables_dir` (ElementTypes.dat, AtomTypes.dat,
    BondTypes.dat); returns (molecule dict, lvec 3x3 or None)."""
    import os
    L = lib()
    L.build_xyz_arrays.argtypes = [c_char_p] * 4 + [c_int, c_int_p] + [c_void_p] * 11
    L.build_xyz_arrays.restype = c_int
    apos = np.zeros((2 * maxn, 3))
    REQs = np.zeros((maxn, 4))
    neighs = np.zeros((maxn, 4), np.int32)
    pars = [np.zeros((maxn, 4)) for _ in range(5)]
    elem = np.zeros(maxn, np.int32)
    lv = np.zeros(9)
    per = c_int(0)
    nn = c_int(0)
    t = lambda f: os.path.join(tables_dir, f).encode()
    n = L.build_xyz_arrays(xyz.encode(), t("ElementTypes.dat"), t("AtomTypes.dat"), t("BondTypes.dat"), maxn,
                           ctypes.byref(nn), apos.ctypes.data, REQs.ctypes.data, neighs.ctypes.data,
                           *[p.ctypes.data for p in pars], elem.ctypes.data, lv.ctypes.data, ctypes.addressof(per))
    if n < 0:
        raise ValueError("cannot build %s" % xyz)
    m = nn.value
    d = dict(natoms=n, nnode=m, apos=np.ascontiguousarray(apos[:n + m]), REQs=np.ascontiguousarray(REQs[:n]),
             neighs=np.ascontiguousarray(neighs[:n]), elem=np.ascontiguousarray(elem[:n]))
    for k, p in zip(("apars", "bLs", "bKs", "Ksp", "Kpp"), pars):
        d[k] = np.ascontiguousarray(p[:m])
    return d, (lv.reshape(3, 3) if per.value else None)


def make_neigh_cells(apos, neighs, lvec, nPBC):
    """(shifts (npbc,3), neighCell (natoms,4)) of a periodic molecule."""
    L = lib()
    L.make_neigh_cells.argtypes = [c_int, c_double_p, c_int_p, c_double_p, c_int_p, c_double_p, c_int_p]
    L.make_neigh_cells.restype = c_int
    apos = np.ascontiguousarray(apos, np.float64)
    neighs = np.ascontiguousarray(neighs, np.int32)
    n = len(neighs)
    lv = np.ascontiguousarray(lvec, np.float64).reshape(9)
    npb = np.ascontiguousarray(nPBC, np.int32)
    nsh = int(np.prod(2 * npb + 1))
    sh = np.zeros((nsh, 3))
    nc = np.zeros((n, 4), np.int32)
    L.make_neigh_cells(n, _p(apos), _p(neighs, c_int_p), _p(lv), _p(npb, c_int_p), _p(sh), _p(nc, c_int_p))
    return sh, nc


def attach_surface_atoms(surf="synth:NaCl", nPBC=(1, 1, 0)):
    """The substrate as explicit atoms (no GridFF); surf None removes them."""
    L = lib()
    L.attach_surface_atoms.argtypes = [c_char_p, c_int_p]
    L.attach_surface_atoms.restype = c_int
    npb = np.ascontiguousarray(nPBC, np.int32)
    r = L.attach_surface_atoms(None if surf is None else surf.encode(), _p(npb, c_int_p))
    if r < 0:
        raise RuntimeError("attach_surface_atoms failed")


def surface_atoms(surf=None, maxn=100000):
    """(apos, REQs, lvec 3x3, nPBC) of the substrate atoms of `surf` read on the
    host (nPBC None), or of the ones in use (None if none)."""
    L = lib()
    L.get_surface_atoms.argtypes = [c_char_p, c_double_p, c_double_p, c_double_p, c_int_p, c_int]
    L.get_surface_atoms.restype = c_int
    name = None if surf is None else surf.encode()
    n = L.get_surface_atoms(name, None, None, None, None, 0)
    if n < 0:
        raise RuntimeError("cannot read surface %s" % surf)
    if n == 0:
        return None
    a, q, lv, npb = np.zeros((n, 3)), np.zeros((n, 4)), np.zeros(9), np.zeros(3, np.int32)
    L.get_surface_atoms(name, _p(a), _p(q), _p(lv), _p(npb, c_int_p), n)
    return a, q, lv.reshape(3, 3), (None if surf is not None else tuple(int(v) for v in npb))


def set_lattice(isys, lvec=None, nPBC=None):
    """Make system isys (all if < 0) periodic with lattice rows lvec (3x3) and
    image range nPBC, or a vacuum molecule again (lvec None)."""
    L = lib()
    L.set_lattice.argtypes = [c_int, c_double_p, c_int_p]
    L.set_lattice.restype = c_int
    lv = None if lvec is None else np.ascontiguousarray(lvec, np.float64).reshape(9)
    npb = None if nPBC is None else np.ascontiguousarray(nPBC, np.int32)
    r = L.set_lattice(isys, None if lv is None else _p(lv), None if npb is None else _p(npb, c_int_p))
    if r < 0:
        raise RuntimeError("set_lattice failed")


def make_nacl_grid(nx=56, ny=56, nz=200, dz=0.1, nthreads=8, Rdamp=1.0):
    """Synthetic NaCl(001) GridFF coefficients (host direct sums + fit); Rdamp =
    Coulomb damping of the node sums (1 A: the benchmark substrate; 0: bare)."""
    coefs = np.zeros((nx, ny, nz, 3))
    geom = np.zeros(7)
    lib().make_nacl_gridff(nx, ny, nz, dz, Rdamp, nthreads, _p(coefs), _p(geom))
    return dict(n=(nx, ny, nz), pos0=geom[0:3].copy(), step=geom[3:6].copy(), alpha=float(geom[6]), coefs=coefs)


# ------------------------------------------------------------ batch API
def init_workload(kind, nsys, seed0, nmin, nmax, sigma=0.1, surf=None, grid_step=0.1):
    n = lib().init_workload(kind, nsys, seed0, nmin, nmax, sigma, surf.encode() if surf else None, grid_step)
    if n < 0:
        raise RuntimeError("init_workload failed")
    return n


def init_systems(systems):
    """Upload a list of molecule dicts (gen_molecule layout) as one batch."""
    cat = lambda k, dt=np.float64: np.ascontiguousarray(np.concatenate([np.asarray(m[k], dt).reshape(-1) for m in systems]))
    natoms = np.array([m["natoms"] for m in systems], np.int32)
    nnode = np.array([m["nnode"] for m in systems], np.int32)
    arrs = [cat("apos"), cat("REQs"), cat("neighs", np.int32), cat("apars"), cat("bLs"), cat("bKs"), cat("Ksp"),
            cat("Kpp"), cat("elem", np.int32)]
    n = lib().init_systems(len(systems), _p(natoms, c_int_p), _p(nnode, c_int_p), _p(arrs[0]), _p(arrs[1]),
                           _p(arrs[2], c_int_p), _p(arrs[3]), _p(arrs[4]), _p(arrs[5]), _p(arrs[6]), _p(arrs[7]),
                           _p(arrs[8], c_int_p))
    if n < 0:
        raise RuntimeError("init_systems failed")
    return n


def attach_surface(surf="synth:NaCl", grid_step=0.1):
    """Load (or create) the GridFF of `surf` and attach it to the current batch."""
    if lib().attach_surface(surf.encode(), grid_step) != 0:
        raise RuntimeError("attach_surface failed")


def batch_sizes():
    out = np.zeros(9, np.int32)
    lib().get_batch_sizes(_p(out, c_int_p))
    keys = ("nsys", "ntotA", "ntotM", "ntotV", "maxN", "maxM", "gnx", "gny", "gnz")
    return dict(zip(keys, (int(x) for x in out)))


def grid_info():
    g = np.zeros(7)
    n = lib().get_grid_info(_p(g))
    s = batch_sizes()
    if n <= 0:
        return None
    ptr = lib().get_grid_coefs()
    coefs = np.ctypeslib.as_array(ptr, shape=(s["gnx"], s["gny"], s["gnz"], 3))
    return dict(n=(s["gnx"], s["gny"], s["gnz"]), pos0=g[0:3].copy(), step=g[3:6].copy(), alpha=float(g[6]),
                coefs=coefs)


def fview(name, n):
    ptr = lib().getfBuff(name.encode())
    if not ptr:
        raise KeyError(name)
    return np.ctypeslib.as_array(ptr, shape=(n,))


def iview(name, n):
    ptr = lib().getIBuff(name.encode())
    if not ptr:
        raise KeyError(name)
    return np.ctypeslib.as_array(ptr, shape=(n,))


def dview(name, n):
    ptr = lib().getBuff(name.encode())
    if not ptr:
        raise KeyError(name)
    return np.ctypeslib.as_array(ptr, shape=(n,))


class Batch:
    """Numpy views of the host mirrors of the current CSR-packed batch."""

    def __init__(self):
        s = batch_sizes()
        self.sizes = s
        ns, A, M, V = s["nsys"], s["ntotA"], s["ntotM"], s["ntotV"]
        self.nsys = ns
        self.info = iview("sys_info", ns * SYSINFO_INTS).reshape(ns, SYSINFO_INTS)
        self.atoms = fview("gpu_atoms", V * 4).reshape(V, 4)
        self.aforces = fview("gpu_aforces", V * 4).reshape(V, 4)
        self.avel = fview("gpu_avel", V * 4).reshape(V, 4)
        self.REQs = fview("gpu_REQs", A * 4).reshape(A, 4)
        self.PLQs = fview("gpu_PLQs", A * 4).reshape(A, 4)
        self.neighs = iview("gpu_neighs", A * 4).reshape(A, 4)
        self.bkNeighs = iview("gpu_bkNeighs", V * 4).reshape(V, 4)
        self.elem = iview("gpu_elem", A)
        self.MMpars = fview("gpu_MMpars", M * 4).reshape(M, 4)
        self.BLs = fview("gpu_BLs", M * 4).reshape(M, 4)
        self.BKs = fview("gpu_BKs", M * 4).reshape(M, 4)
        self.Ksp = fview("gpu_Ksp", M * 4).reshape(M, 4)
        self.Kpp = fview("gpu_Kpp", M * 4).reshape(M, 4)
        self.constr = fview("gpu_constr", A * 4).reshape(A, 4)
        self.constrK = fview("gpu_constrK", A * 4).reshape(A, 4)
        self.bboxes = fview("gpu_bboxes", ns * 12).reshape(ns, 3, 4)
        self.nsteps = iview("sys_nsteps", ns)
        self.status = iview("sys_status", ns)
        self.Fmax = fview("sys_Fmax", ns)
        self.E = dview("sys_E", ns)

    def system(self, s, double=True):
        """System s as a molecule dict (gen_molecule layout) from the mirrors."""
        N, M, a0, n0, v0 = (int(x) for x in self.info[s, :5])
        d = dict(natoms=N, nnode=M,
                 apos=self.atoms[v0:v0 + N + M, :3].astype(np.float64),
                 REQs=self.REQs[a0:a0 + N].astype(np.float64),
                 neighs=np.ascontiguousarray(self.neighs[a0:a0 + N]),
                 elem=np.ascontiguousarray(self.elem[a0:a0 + N]),
                 apars=self.MMpars[n0:n0 + M].astype(np.float64),
                 bLs=self.BLs[n0:n0 + M].astype(np.float64),
                 bKs=self.BKs[n0:n0 + M].astype(np.float64),
                 Ksp=self.Ksp[n0:n0 + M].astype(np.float64),
                 Kpp=self.Kpp[n0:n0 + M].astype(np.float64))
        return d

    def forces(self, s):
        N, M, a0, n0, v0 = (int(x) for x in self.info[s, :5])
        return self.aforces[v0:v0 + N + M, :3].astype(np.float64)

    def positions(self, s):
        N, M, a0, n0, v0 = (int(x) for x in self.info[s, :5])
        return self.atoms[v0:v0 + N + M, :3].astype(np.float64)


def upload(bParams=True, bForces=False, bVel=False, blvec=True):
    """Copy the host mirrors (positions, restraints, walls, parameters) to the device."""
    L = lib()
    L.upload.argtypes = [c_bool] * 4
    L.upload.restype = None
    L.upload(bParams, bForces, bVel, blvec)


def relax(nstep_max, Fconv, path=0, restore_start=False, download=True):
    n = lib().relax(nstep_max, Fconv, path, restore_start, download)
    if n < 0:
        raise RuntimeError("relax failed")
    return n


def relax_noalloc(nstep_max, Fconv, restore_start=False, download=True):
    """relax() on the persistent kernel with no device allocation (static scratch)."""
    n = lib().relax_noalloc(nstep_max, Fconv, restore_start, download)
    if n < 0:
        raise RuntimeError("relax_noalloc failed")
    return n


def explore(nstep_max, Fconv, T, gamma, n_explore, n_relax=0, seed=0, kmax=16, positions=True):
    """Exploring run of the whole batch (split kernels); returns the minima found."""
    L = lib()
    L.explore.argtypes = [c_int, c_double, c_double, c_double, c_int, c_int, c_longlong, c_int, c_bool]
    L.explore.restype = c_int
    n = L.explore(nstep_max, Fconv, T, gamma, n_explore, n_relax, seed, kmax, positions)
    if n < 0:
        raise RuntimeError("explore failed")
    return n


def minima(isys, maxn=16):
    """(found, energies, positions) of system isys from the last explore()."""
    L = lib()
    L.get_minima.argtypes = [c_int, c_double_p, c_double_p, c_int]
    L.get_minima.restype = c_int
    s = batch_sizes()
    info = iview("sys_info", s["nsys"] * SYSINFO_INTS).reshape(s["nsys"], SYSINFO_INTS)
    V = int(info[isys, 0] + info[isys, 1])
    E = np.zeros(maxn)
    P = np.zeros((maxn, V, 3))
    n = L.get_minima(isys, _p(E), _p(P), maxn)
    if n < 0:
        raise RuntimeError("get_minima failed")
    k = min(n, maxn)
    return n, E[:k], P[:k]


def database():
    """Members of the minima database: list of dicts (natoms, E, isys,
    occurrence, pos (natoms, 3))."""
    L = lib()
    L.get_database_member.argtypes = [c_int, c_double_p, c_int_p, c_int_p, c_double_p, c_int]
    L.get_database_member.restype = c_int
    n = L.get_database_member(-1, None, None, None, None, 0)
    out = []
    for i in range(n):
        na = L.get_database_member(i, None, None, None, None, 0)
        E = np.zeros(1)
        s = np.zeros(1, np.int32)
        occ = np.zeros(1, np.int32)
        P = np.zeros((na, 3))
        L.get_database_member(i, _p(E), _p(s, c_int_p), _p(occ, c_int_p), _p(P), na)
        out.append(dict(natoms=na, E=float(E[0]), isys=int(s[0]), occurrence=int(occ[0]), pos=P))
    return out


def database_add(pos, types, E=0.0):
    """Offer a structure to the minima database: index of the member it
    duplicates, or -1 when it was added."""
    L = lib()
    L.database_add.argtypes = [c_int, c_double_p, c_int_p, c_double]
    L.database_add.restype = c_int
    pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
    types = np.ascontiguousarray(types, np.int32)
    r = L.database_add(len(pos), _p(pos), _p(types, c_int_p), float(E))
    if r < -1:
        raise ValueError("database_add: invalid structure")
    return r


def clear_database():
    lib().clear_database()


def database_descriptor(pos):
    L = lib()
    L.database_descriptor.argtypes = [c_int, c_double_p, c_double_p]
    L.database_descriptor.restype = c_int
    pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
    d = np.zeros(4)
    n = L.database_descriptor(len(pos), _p(pos), _p(d))
    return d[:n]


def database_distance(x, tx, y, ty):
    L = lib()
    L.database_distance.argtypes = [c_int, c_double_p, c_int_p, c_double_p, c_int_p]
    L.database_distance.restype = ctypes.c_double
    x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
    y = np.ascontiguousarray(y, np.float64).reshape(-1, 3)
    tx = np.ascontiguousarray(tx, np.int32)
    ty = np.ascontiguousarray(ty, np.int32)
    return L.database_distance(len(x), _p(x), _p(tx, c_int_p), _p(y), _p(ty, c_int_p))


def set_save_to_database(on=True):
    """setSwitches_multi(..., bSaveToDatabase): explore() offers its converged
    minima to the database while on."""
    L = lib()
    L.setSwitches_multi.argtypes = [c_int] * 9
    L.setSwitches_multi(0, 0, 0, 0, 0, 0, 0, 0, 1 if on else -1)


def last_ms():
    return lib().get_last_ms()


def set_cost_model(feedback, c0=1000.0, c1=20.0):
    L = lib()
    L.set_cost_model.argtypes = [c_int, c_double, c_double]
    L.set_cost_model.restype = c_int
    if L.set_cost_model(int(feedback), c0, c1) != 0:
        raise ValueError("set_cost_model: invalid coefficients")


def set_class_mask(mask):
    L = lib()
    L.set_class_mask.argtypes = [c_int]
    L.set_class_mask.restype = c_int
    r = L.set_class_mask(int(mask))
    if r < 0:
        raise ValueError("set_class_mask: mask outside 0..15")
    return r


def device_status():
    """0, or the HIP error code of a device failure not reported by an earlier call."""
    L = lib()
    L.device_status.argtypes = []
    L.device_status.restype = c_int
    return L.device_status()


def set_record_times(on):
    L = lib()
    L.set_record_times.argtypes = [c_int]
    L.set_record_times.restype = c_int
    return L.set_record_times(int(on))


def system_times():
    """(nsys, 2) start/end of each system in the last recorded relaxation (us)."""
    L = lib()
    L.get_system_times.argtypes = [c_double_p]
    L.get_system_times.restype = c_int
    out = np.zeros((batch_sizes()["nsys"], 2))
    if L.get_system_times(_p(out)) != 0:
        raise RuntimeError("no system times recorded")
    return out


def class_info():
    """Persistent-kernel size classes of the current batch with the device
    time (ms) of each class launch in the last relax."""
    L = lib()
    info = np.zeros(6 * 8, np.int32)
    ms = np.zeros(8)
    L.get_class_info.restype = c_int
    n = L.get_class_info(_p(info, c_int_p), _p(ms), 8)
    out = []
    for c in range(n):
        b, ns, nmax, mmax, ng, lds = (int(x) for x in info[6 * c:6 * c + 6])
        order = np.zeros(ns, np.int32)
        L.get_class_order(c, _p(order, c_int_p), ns)
        out.append(dict(block=b, nsys=ns, nmax=nmax, mmax=mmax, ngroups=ng, lds=lds, ms=float(ms[c]), order=order))
    return out


def eval_forces():
    if lib().eval_forces() != 0:
        raise RuntimeError("eval_forces failed")


def set_conv_mode(mode):
    lib().set_conv_mode(mode)


def set_force_limit(Flimit):
    L = lib()
    L.set_force_limit.argtypes = [c_double]
    L.set_force_limit(Flimit)


def eval_precond():
    """Exact forces (Batch.forces) and preconditioned forces (returned, (ntotV,4))."""
    L = lib()
    L.eval_precond.restype = c_int
    if L.eval_precond() != 0:
        raise RuntimeError("eval_precond failed")
    nt = batch_sizes()["ntotV"]
    return fview("gpu_pcforces", nt * 4).reshape(nt, 4).copy()


def set_sched(mode):
    L = lib()
    L.set_sched.argtypes = [c_int]
    L.set_sched.restype = c_int
    r = L.set_sched(mode)
    if r < 0:
        raise ValueError("set_sched: invalid mode")
    return r


def set_precond(mode, gstar=3.0, grigid=63.0):
    L = lib()
    L.set_precond.argtypes = [c_int, c_double, c_double]
    L.set_precond.restype = c_int
    if L.set_precond(mode, gstar, grigid) != 0:
        raise ValueError("set_precond: invalid arguments")


def set_opt(dt_max=0.05, dt_min=0.005, damp_max=0.1, finc=1.1, fdec=0.5, falpha=0.8, minLastNeg=3,
            cvf_min=-0.1, cvf_max=0.1):
    lib().set_opt(dt_max, dt_min, damp_max, finc, fdec, falpha, minLastNeg, cvf_min, cvf_max)


# ---- proposal network -------------------------------------------------------
def proposal_nparam():
    L = lib()
    L.get_proposal_nparam.restype = c_int
    return L.get_proposal_nparam()


def set_proposal_weights(theta):
    L = lib()
    th = np.ascontiguousarray(theta, np.float32)
    L.set_proposal_weights.argtypes = [ctypes.c_void_p, c_int]
    L.set_proposal_weights.restype = c_int
    if L.set_proposal_weights(th.ctypes.data, th.size) != 0:
        raise ValueError("set_proposal_weights failed")


def get_proposal_weights():
    L = lib()
    n = proposal_nparam()
    th = np.zeros(n, np.float32)
    L.get_proposal_weights.argtypes = [ctypes.c_void_p, c_int]
    L.get_proposal_weights.restype = c_int
    if L.get_proposal_weights(th.ctypes.data, n) != 0:
        raise RuntimeError("get_proposal_weights failed")
    return th


def set_proposal_stall(windows):
    """FIRE fallback of the learned form after `windows` proposal points in a
    row without a new lowest max|F| (0 = never)."""
    L = lib()
    L.set_proposal_stall.argtypes = [c_int]
    L.set_proposal_stall.restype = c_int
    if L.set_proposal_stall(int(windows)) != 0:
        raise ValueError("set_proposal_stall: windows must be >= 0")


def proposal_fallback():
    """Per system: True where the last learned relaxation fell back to plain FIRE."""
    L = lib()
    L.get_proposal_fallback.argtypes = [c_void_p]
    L.get_proposal_fallback.restype = c_int
    out = np.zeros(batch_sizes()["nsys"], np.int32)
    if L.get_proposal_fallback(out.ctypes.data) != 0:
        raise RuntimeError("get_proposal_fallback failed")
    return out.astype(bool)


def set_proposal_every(every):
    L = lib()
    L.set_proposal_every.argtypes = [c_int]
    L.set_proposal_every.restype = c_int
    if L.set_proposal_every(every) != 0:
        raise ValueError("set_proposal_every: invalid period")


def proposal_forward():
    """(outputs (ntotA,2), features (ntotA,16)) at the current positions."""
    L = lib()
    nA = batch_sizes()["ntotA"]
    out = np.zeros((nA, 2), np.float32)
    feats = np.zeros((nA, 16), np.float32)
    L.proposal_forward.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    L.proposal_forward.restype = c_int
    if L.proposal_forward(out.ctypes.data, feats.ctypes.data) != 0:
        raise RuntimeError("proposal_forward failed")
    return out, feats


def proposal_loss_grad(eta, dmax):
    L = lib()
    s = batch_sizes()
    dout = np.zeros((s["ntotA"], 2), np.float32)
    loss = np.zeros(s["nsys"], np.float32)
    L.proposal_loss_grad.argtypes = [c_double, c_double, ctypes.c_void_p, ctypes.c_void_p]
    L.proposal_loss_grad.restype = c_int
    if L.proposal_loss_grad(eta, dmax, dout.ctypes.data, loss.ctypes.data) != 0:
        raise RuntimeError("proposal_loss_grad failed")
    return dout, loss


def proposal_backward(dout):
    L = lib()
    d = np.ascontiguousarray(dout, np.float32)
    g = np.zeros(proposal_nparam(), np.float32)
    L.proposal_backward.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    L.proposal_backward.restype = c_int
    if L.proposal_backward(d.ctypes.data, g.ctypes.data) != 0:
        raise RuntimeError("proposal_backward failed")
    return g


def record_snapshots(every, n):
    L = lib()
    L.record_snapshots.argtypes = [c_int, c_int]
    L.record_snapshots.restype = c_int
    if L.record_snapshots(every, n) != 0:
        raise RuntimeError("record_snapshots failed")


def train_proposal(iters, lr, eta, dmax, seed=0):
    L = lib()
    hist = np.zeros(max(iters, 1), np.float32)
    L.train_proposal.argtypes = [c_int, c_double, c_double, c_double, ctypes.c_longlong, ctypes.c_void_p]
    L.train_proposal.restype = c_int
    if L.train_proposal(iters, lr, eta, dmax, seed, hist.ctypes.data) != 0:
        raise RuntimeError("train_proposal failed")
    return hist[:iters]


def last_chunks():
    L = lib()
    L.get_last_chunks.restype = c_int
    return L.get_last_chunks()


def set_graphs(on):
    L = lib()
    L.set_graphs.argtypes = [c_int]
    L.set_graphs.restype = c_int
    return L.set_graphs(1 if on else 0)


def set_learned_inkernel(on):
    L = lib()
    L.set_learned_inkernel.argtypes = [c_int]
    L.set_learned_inkernel.restype = c_int
    return L.set_learned_inkernel(int(on))
