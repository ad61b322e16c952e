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
