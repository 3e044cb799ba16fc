"""The tile-dim autotuner checked against measurement on the GPU.

bgb_autotune_measure packs each workload at T = 4, 8, 16, 32 and times every
kernel configuration (run length 8 / 16; packed x gathered or staged in LDS
for the bit op). The profile-driven model (bgb_autotune_profile, from the
tile statistics of printMatrixStats_T / estimatevec, utility.cu:808-975)
must pick a configuration whose measured time is within 10% of the fastest
one, on the benchmark workloads: R-MAT scale 20 and 22 (edge factor 16) and
the 4096 x 4096 grid, for bin_bin_full and bin_full_full. The reference
fixes T at compile time (Makefile:12-46).
"""
import sys

import pytest

pytestmark = pytest.mark.gpu

_GRAPHS = {}


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def workload(name):
    from bgb import graph
    if name not in _GRAPHS:
        _GRAPHS[name] = {"rmat20": lambda: graph.rmat(20, 16, 0.57, 0.19, 0.19, seed=1),
                         "rmat22": lambda: graph.rmat(22, 16, 0.57, 0.19, 0.19, seed=1),
                         "grid4096": lambda: graph.grid2d(4096, 4096)}[name]()
    return _GRAPHS[name]


@pytest.mark.timeout(600)
@pytest.mark.parametrize("op", ("bin_bin", "bin_full_full"))
@pytest.mark.parametrize("name", ("rmat20", "rmat22", "grid4096"))
def test_model_pick_within_10pct_of_measured_best(cuda, name, op):
    from bgb import b2sr as B
    from bgb import graph
    csr = workload(name)
    meas = B.autotune_measure(csr, op, reps=10)
    times = meas["times_us"]
    best_key = min(times, key=times.get)
    best = times[best_key]
    cfg = graph.autotune_profile(csr, op)
    key = (cfg["T"], cfg["run_length"], "lds" if (op == "bin_bin" and cfg["x_staging"] == 2) else "memory")
    got = times[key]
    log(f"[tune] {name} {op}: model {key} {got:.1f} us, measured best {best_key} {best:.1f} us "
        f"(ratio {got / best:.3f}); predicted {cfg['cost_us']}")
    assert got <= 1.10 * best, (key, got, best_key, best)
    # the measured tuner returns its fastest configuration
    pick = (meas["T"], meas["run_length"],
            "lds" if (op == "bin_bin" and meas["T"] <= 16 and meas["x_staging"] == 2) else "memory")
    assert times[pick] == best


def test_measured_tuner_restores_knobs_and_rejects_bad_arguments(cuda):
    import ctypes
    import bgb
    from bgb import b2sr as B
    from bgb import graph
    csr = graph.rmat(12, 8, seed=3)
    bgb.call("bgb_set_run_length", 8, 16)
    try:
        B.autotune_measure(csr, "bin_bin", reps=2)
        assert bgb.lib().bgb_get_run_length(8) == 16
        assert all(bgb.lib().bgb_get_x_staging(T) == 0 for T in (4, 8, 16, 32))
    finally:
        bgb.call("bgb_set_run_length", 8, 8)
    t = ctypes.c_int(0)
    assert bgb.lib().bgb_autotune_measure(5, 10, 10, 0, 1, None, 15, 2, ctypes.byref(t), None, None, None,
                                          None) == 1
