"""The C-ABI library loads and exports every symbol include/mmff_multi.h
declares, and every symbol the reference's pyBall/MMFF_multi.py binds."""
import ctypes
import os
import re

from firecore_amd import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# names bound by pyBall/MMFF_multi.py at import (argtypes/restype set)
REFERENCE_BOUND = """sample_DistConstr sample_evalPiAling sample_evalAngleCos sample_evalAngleCosHalf sampleNonBond
sampleSurf sampleSurf_vecs printBuffNames getIBuff getBuff getfBuff init_buffers setVerbosity init insertSMILES
setSwitches_multi checkInvariants set_opt setOptLog setTrjName saveXYZ getTypeName eval run MDloop pack_system
unpack_system upload_sys download_sys upload download change_lvec optimizeLattice_1d optimizeLattice_2d_multi
upload_pop addDistConstrain addAngConstrain shift_atoms_ax shift_atoms rotate_atoms_ax scanTranslation_ax
scanTranslation scanRotation_ax scanRotation scan""".split()


def header_functions():
    txt = open(os.path.join(ROOT, "include", "mmff_multi.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(", txt, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "extern")))


def test_library_loads_and_exports_header_symbols():
    L = native.lib()
    missing = [n for n in header_functions() if not hasattr(L, n)]
    assert not missing, missing
    assert len(header_functions()) >= 60


def test_reference_bound_symbols_exported():
    L = ctypes.CDLL(native.LIB_PATH)
    missing = [n for n in REFERENCE_BOUND if not hasattr(L, n)]
    assert not missing, missing
    hdr = set(header_functions())
    assert set(REFERENCE_BOUND) <= hdr


def test_python_module_mirror_imports():
    from firecore_amd import MMFF_multi  # noqa: F401  (binds every symbol at import, as the reference module)
    assert MMFF_multi.lib is native.lib()


def test_host_setters_validate_their_arguments():
    """Mode setters check their argument without touching a device."""
    old = native.set_sched(2)
    try:
        assert native.set_sched(1) == 2
        for bad in (-1, 3):
            try:
                native.set_sched(bad)
            except ValueError:
                pass
            else:
                raise AssertionError("set_sched accepted %d" % bad)
    finally:
        native.set_sched(old)


def test_persistent_lds_budget():
    """LDS per workgroup of the persistent kernel (host-side layout function):
    the windowed reaction rows and half exclusion rows keep three 256-atom
    workgroups and four 192-atom workgroups within a CU's 160 KB, and the
    learned form's padded layout always holds the first activation buffer."""
    L = native.lib()
    L.fc_lds_bytes.restype = ctypes.c_int
    L.fc_hslot_elems.restype = ctypes.c_int
    cu = 160 * 1024
    for nmax, mmax, block, per_cu in [(256, 128, 256, 3), (192, 100, 192, 4), (128, 66, 128, 6)]:
        wmax = ((nmax >> 1) >> 5) + 1
        base = L.fc_lds_bytes(nmax, mmax, wmax, block, 0, 0)
        learned = L.fc_lds_bytes(nmax, mmax, wmax, block, 1, 0)
        assert learned >= base
        assert per_cu * learned <= cu, (nmax, learned)
        # first activation buffer (nmax rows of 68 bf16) fits between facc and the gains
        facc = 16 * (nmax + mmax) + 16 * nmax * 2
        assert learned - facc >= nmax * 68 * 2
    for nmax in (16, 64, 100, 256):
        assert L.fc_hslot_elems(nmax) == ((nmax * 68 * 2 + 15) & ~15) // 2


# ---------------------------------------------------------------- argtypes/restype pin
# C type of a header parameter -> the ctypes type a binding passes for it
C2CT = {"int": "c_int", "double": "c_double", "bool": "c_bool", "char*": "c_char_p", "int*": "c_int_p",
        "double*": "c_double_p", "float*": "c_float_p", "void*": "c_void_p", "void": "None", "long long": "c_longlong"}
# Deliberate, documented departures of the reference's Python binding from its
# own C signatures (the C side is authoritative here):
#   insertSMILES: binds (c_char_p, c_bool), C takes (char*)      libMMFF.h:25
#   setTrjName:   binds 2 arguments, C takes 3                   libMMFF.h:53
#   rotate_atoms_ax: rebound with rotate_atoms' argument list    MMFF_multi.py:515-534
#   addDistConstrain: C has a trailing bool bOldIndex            libMMFF.h:41
#   scanTranslation_ax, scanTranslation: C has (const char* trjName, bool bAddjustCaps)
#                 where the binding has one bool                 libMMFF.h:134, 138
#   scanRotation_ax, scanRotation: C ends with const char* trjName, the binding
#                 with a bool                                     libMMFF.h:141, 147
KNOWN_PY_C_MISMATCH = {"insertSMILES", "setTrjName", "rotate_atoms_ax", "addDistConstrain", "scanTranslation_ax",
                       "scanTranslation", "scanRotation_ax", "scanRotation"}


def _ctype(decl):
    d = decl.replace("const ", "").strip()
    d = re.sub(r"\s+", " ", d)
    stars = d.count("*")
    base = d.replace("*", "").strip()
    # drop the parameter name (last word) where there is one
    parts = base.split(" ")
    if len(parts) > 1 and parts[-1] not in ("int", "double", "bool", "char", "float", "void", "long"):
        parts = parts[:-1]
    base = " ".join(parts)
    return C2CT.get(base + "*" * stars, base + "*" * stars)


def header_prototypes():
    txt = open(os.path.join(ROOT, "include", "mmff_multi.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    out = {}
    for m in re.finditer(r"^([A-Za-z_][\w \*]*?)\b([A-Za-z_]\w*)\s*\(([^)]*)\)\s*;", txt, flags=re.M):
        ret, name, args = m.group(1), m.group(2), m.group(3).strip()
        a = [] if args in ("", "void") else [_ctype(x) for x in args.split(",")]
        out[name] = (a, _ctype(ret))
    return out


def reference_bindings():
    out = {}
    for line in open(os.path.join(ROOT, "tests", "golden", "abi_reference.txt")):
        if line.startswith("#") or not line.strip():
            continue
        name, args, ret = (x.strip() for x in line.split("|"))
        out[name] = ([a for a in args.split(",") if a], ret)
    return out


def test_header_matches_reference_ctypes_bindings():
    """Every binding of pyBall/MMFF_multi.py (name -> argtypes -> restype,
    fixture tests/golden/abi_reference.txt) matches include/mmff_multi.h,
    except the documented Python/C departures of the reference."""
    hdr = header_prototypes()
    ref = reference_bindings()
    assert len(ref) == 45
    bad = []
    for name, (args, ret) in ref.items():
        assert name in hdr, name
        if name in KNOWN_PY_C_MISMATCH:
            continue
        if hdr[name] != (args, ret):
            bad.append((name, hdr[name], (args, ret)))
    assert not bad, bad
    # the departures are real (so the exception list stays minimal)
    for name in KNOWN_PY_C_MISMATCH:
        assert hdr[name] != ref[name], name


def test_mirror_module_binds_header_signatures():
    """firecore_amd/MMFF_multi.py binds the reference names with the header's
    C signatures (including those where the reference's binding departs)."""
    from firecore_amd import MMFF_multi as mm
    import ctypes as ct
    hdr = header_prototypes()
    names = {ct.c_int: "c_int", ct.c_double: "c_double", ct.c_bool: "c_bool", ct.c_char_p: "c_char_p",
             ct.c_void_p: "c_void_p", ct.POINTER(ct.c_double): "c_double_p", ct.POINTER(ct.c_int): "c_int_p",
             ct.POINTER(ct.c_float): "c_float_p", None: "None"}
    checked = 0
    for name in reference_bindings():
        f = getattr(mm.lib, name)
        if f.argtypes is None:
            continue
        got = [names.get(t, getattr(t, "__name__", str(t))) for t in f.argtypes]
        # numpy ndpointer argtypes stand for double*/int* pointers
        got = ["c_double_p" if g.startswith("ndpointer_<f8") else "c_int_p" if g.startswith("ndpointer_<i4") else g
               for g in got]
        assert got == hdr[name][0], (name, got, hdr[name][0])
        checked += 1
    assert checked >= 40
