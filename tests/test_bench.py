"""bench.py's host-side pieces, without a GPU: the one-GPU refusal, the
algorithmic bytes of a launch, the lookup of the committed counter
summaries (including kernels whose launch name hides template arguments)
and the CPU baseline's run of the reference build."""
import os
import subprocess
import sys

import pytest

import bench

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def test_more_than_one_gpu_is_refused_before_any_gpu_work():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one MI355X" in (r.stdout + r.stderr)


def test_kernel_bytes_is_bytes_per_cell_times_fields_times_cells():
    ni, nj, nk = 256, 128, 20
    per_cell, fields = bench.KERNEL_BYTES["k_hv_apply"]
    assert bench.kernel_bytes("k_hv_apply<T>", ni, nj, nk, 1) == per_cell * fields * ni * (nj + 1) * nk
    # template arguments do not change the bytes; unknown kernels have none
    assert bench.kernel_bytes("k_hv_apply<float>", ni, nj, nk, 1) == bench.kernel_bytes("k_hv_apply<T>", ni, nj, nk, 1)
    assert bench.kernel_bytes("k_no_such_kernel", ni, nj, nk, 1) is None


def test_measured_traffic_lookup():
    t, src = bench.measured_traffic("jupiter_256x128x20", "fp64", "k_hv_apply<T>")
    assert t and t > 0 and src.endswith("traffic_jupiter_256x128x20_fp64.json")
    # the zonal hyperviscosity's profile name carries its register layout
    t, _ = bench.measured_traffic("hot_jupiter_1024x512x64", "fp64", "k_zonal_hv<T>")
    assert t and t > 0
    # the mixed mode's FP32 instantiation is found by its base name, and so
    # is the fused del^4 row kernel, whose profile name carries its template
    t, src = bench.measured_traffic("jupiter_256x128x20", "mixed", "k_hsu_mer<T>")
    assert t and t > 0 and src.endswith("traffic_jupiter_256x128x20_mixed.json")
    t, _ = bench.measured_traffic("jupiter_ensemble_64x128x64x16", "mixed", "k_hv4_rows")
    assert t and t > 0
    # launch name k_set_p2etc<1>, profile name k_set_p2etc<1, true> (the
    # mixed mode's math library)
    t1, _ = bench.measured_traffic("jupiter_256x128x20", "mixed", "k_set_p2etc<1>")
    t0, _ = bench.measured_traffic("jupiter_256x128x20", "mixed", "k_set_p2etc<0>")
    assert t1 and t0 and t1 != t0
    assert bench.measured_traffic("no_such_workload", "fp64", "k_hv_apply<T>") == (None, None)


def test_filter_band_rows():
    # rows with |lat| >= 45 deg: half of the h-rows on these globes
    assert bench.filtered_rows(128) == 64
    assert bench.filtered_rows(64) == 32


def test_reference_cpu_leg_times_the_reference_build():
    """bench.py's CPU baseline runs the reference's own timestep() where
    oracle/_ref/epic_ref was built, on the same seeded case."""
    if not os.path.exists(bench.REF_EXE):
        pytest.skip("oracle/_ref/epic_ref not built")
    import epic_amd as ea
    v, n, el = bench.cpu_baseline_reference(ea.HELD_SUAREZ, 32, 16, 8, 600, 0.5)
    assert n >= 2 and el > 0
    assert abs(v - 32 * 17 * 8 * n / el) < 1e-6 * v
