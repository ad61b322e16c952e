import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(autouse=True)
def _device_clean_after_gpu_test(request):
    """After every device test, wait for the device and fail that test if it
    left an unreported device error (asynchronous faults surface at a later
    call otherwise, and would be charged to whichever test makes it)."""
    yield
    if request.node.get_closest_marker("gpu") is not None:
        from firecore_amd import native
        assert native.device_status() == 0, "device error after " + request.node.nodeid


@pytest.fixture(scope="session")
def orc():
    import oracle
    if not os.path.exists(oracle.ORACLE_SO):
        oracle.build(with_ref=False)
    return oracle.oracle()


@pytest.fixture(scope="session")
def ref():
    import oracle
    r = oracle.reference()
    if r is None:
        pytest.skip("reference build oracle/_ref is not present on this machine")
    return r


@pytest.fixture(scope="session")
def nacl_grid():
    from firecore_amd import native
    return native.make_nacl_grid(56, 56, 200, 0.1, 8)


# Fire parameters of the reference CPU path after MolWorld_sp3 initialisation
# (initOpt(dt_default=0.05), DynamicOpt.h:47-95):
# dt_max, dt_min, damp_max, finc, fdec, falpha, minLastNeg
FIRE_REF = (0.05, 0.005, 0.1, 1.1, 0.5, 0.8, 3)
# pyBall/MMFF_multi.py set_opt() defaults
FIRE_MULTI = (0.1, 0.02, 0.2, 1.1, 0.5, 0.8, 5)
