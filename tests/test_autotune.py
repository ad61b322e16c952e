"""Plan selection of the launch-plan autotuner (firecore_amd/autotune.py), on
recorded calibration times (no device needed)."""
from firecore_amd import autotune


def _masks(**kw):
    return {int(k[1:], 16): v for k, v in kw.items()}


def test_default_kept_within_margin():
    t = _masks(xf=350.23, xe=349.57, xd=353.41, xc=360.38, xb=348.01, xa=350.32, x9=356.3, x8=356.77)
    assert autotune._pick(t, 0xF, rank=lambda m: bin(m).count("1")) == 0xF  # best alternative only 0.6 % faster


def test_near_ties_go_to_fewer_classes():
    # two calibrations of one form: 0xb and 0xa within 0.3 % of each other, both far below the default
    rank = lambda m: bin(m).count("1")  # noqa: E731
    a = _masks(xf=403.35, xe=408.65, xd=384.59, xc=390.48, xb=374.52, xa=373.4, x9=384.31, x8=386.58)
    b = _masks(xf=402.42, xe=410.84, xd=384.48, xc=390.48, xb=372.11, xa=372.37, x9=386.16, x8=387.99)
    assert autotune._pick(a, 0xF, rank=rank) == autotune._pick(b, 0xF, rank=rank) == 0xA
    # without a rank the faster one wins, and the two calibrations disagree
    assert autotune._pick(a, 0xF) == 0xA and autotune._pick(b, 0xF) == 0xB


def test_clear_winner_taken():
    rank = lambda m: bin(m).count("1")  # noqa: E731
    t = _masks(xf=400.0, xb=360.0, xa=380.0, x8=390.0)
    assert autotune._pick(t, 0xF, rank=rank) == 0xB
    assert autotune._pick({"sched1": 404.0, "sched2": 405.2, "sched0": 424.3}, "sched1") == "sched1"
    assert autotune._pick({"sched1": 393.8, "sched2": 390.5, "sched0": 377.1}, "sched1") == "sched0"
