"""The measurement arithmetic of bench.py on the CPU: kernel-name reduction,
the HIP-event roofline with its bracket correction, and the reduction of
rocprofv3 counter passes (marker split, calibration-copy scaling, bytes and
rates per kernel), on synthetic rocprofv3 CSV files whose answers are known.
Planted mistakes (a wrong scale, a marker out of place) must change or
reject the result."""

import csv
import importlib.util
import json
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", ROOT / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_kernel_base_name(bench):
    f = bench.kernel_base_name
    assert f("void seemps::(anonymous namespace)::zgemm_kernel<0, 1, "
             "seemps::(anonymous namespace)::PrecF64, 32, 32, true>(long, long)") == "zgemm_kernel"
    assert f("lz_orth1(long, int, HIP_vector_type<double, 2u> const*)") == "lz_orth1"
    assert f("__amd_rocclr_copyBuffer") == "__amd_rocclr_copyBuffer"


def test_roofline_bracket_correction(bench, monkeypatch):
    """achieved = flops per launch / (event average - (bracket - empty
    kernel's own duration)); the dominant kernel is the one with most time
    among those that carry flops."""
    monkeypatch.setattr(bench, "read_empty_dispatch_us", lambda: 2.0)
    records = {
        "zgemm_kernel": dict(launches=100, ms=2.0, flops=100 * 4.0e8),   # 20 us per launch
        "polar_svd": dict(launches=2, ms=1.5, flops=2 * 1.0e9),
        "phase_sweep": dict(launches=1, ms=50.0, flops=1.0),            # never dominant
        "lz_finish": dict(launches=500, ms=9.0, flops=0.0),             # no flops
    }
    r = bench.roofline(records, bracket_us=6.0)
    assert r["kernel"] == "zgemm_kernel"
    assert r["avg_launch_us_events"] == 20.0
    assert r["avg_launch_us"] == 16.0            # 20 - (6 - 2)
    assert r["achieved"] == pytest.approx(4.0e8 / 16e-6 / 1e12, rel=1e-4)
    assert r["frac"] == pytest.approx(r["achieved"] / bench.PEAK_FP64_TFLOPS, rel=1e-3)
    # without the empty kernel's duration on record nothing is subtracted
    monkeypatch.setattr(bench, "read_empty_dispatch_us", lambda: None)
    assert bench.roofline(records, bracket_us=6.0)["avg_launch_us"] == 20.0
    # the correction never removes more than half of the event average
    monkeypatch.setattr(bench, "read_empty_dispatch_us", lambda: 0.0)
    assert bench.roofline(records, bracket_us=500.0)["avg_launch_us"] == 10.0


def _write_pass(directory, dispatches, counter):
    """A counter_collection.csv of rocprofv3: one row per dispatch and
    counter, as the tool writes it (values in KiB for the size counters)."""
    directory.mkdir(parents=True, exist_ok=True)
    path = directory / "host" / "pmc_counter_collection.csv"
    path.parent.mkdir(parents=True, exist_ok=True)
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value",
                    "Start_Timestamp", "End_Timestamp"])
        t = 1000
        for i, (kernel, value, ns) in enumerate(dispatches):
            w.writerow([i + 1, kernel, counter, value, t, t + ns])
            t += ns + 100
    return path


MARK = "void seemps::(anonymous namespace)::timing_empty_kernel()"
COPY = "void seemps::(anonymous namespace)::zscale2d_kernel(long)"


def test_pmc_small_summary(bench, tmp_path):
    """Counters read half the fetched bytes and all written bytes of the
    calibration copy: every kernel's bytes are scaled by the same factors.
    Only the dispatches between the two markers count."""
    copy_kib = bench.PMC_COPY_ELEMS * 16 / 1024
    # kernel, FETCH KiB (half of the true bytes), WRITE KiB, duration ns
    work = [("zgemm_kernel", 50.0, 30.0, 2000), ("zgemm_kernel", 150.0, 10.0, 6000),
            ("svd_jacobi", 1000.0, 4000.0, 1_000_000)]
    before = [("zgemm_kernel", 9e9, 9e9, 10)]   # outside the markers: ignored
    fetch = before + [(MARK, 0, 0, 10)] + [(k, f, 0, d) for k, f, _, d in work] + \
        [(MARK, 0, 0, 10), (COPY, copy_kib / 2, 0, 50_000)]
    write = before + [(MARK, 0, 0, 10)] + [(k, w, 0, d) for k, _, w, d in work] + \
        [(MARK, 0, 0, 10), (COPY, copy_kib, 0, 50_000)]
    _write_pass(tmp_path / "f", [(k, v, d) for k, v, _, d in fetch], "FETCH_SIZE")
    _write_pass(tmp_path / "w", [(k, v, d) for k, v, _, d in write], "WRITE_SIZE")
    out = bench.pmc_small_summary(str(tmp_path / "f"), str(tmp_path / "w"))
    assert out["fetch_scale"] == 2.0 and out["write_scale"] == 1.0
    rows = {r["kernel"]: r for r in out["kernels"]}
    z = rows["zgemm_kernel"]
    assert z["dispatches"] == 2
    # (2 x 50 + 30) + (2 x 150 + 10) KiB over 8000 ns
    zb = ((2 * 50 + 30) + (2 * 150 + 10)) * 1024
    assert z["bytes_per_dispatch"] == round(zb / 2)
    assert z["us_per_dispatch"] == 4.0
    assert z["gb_per_s"] == round(zb / 8000, 1)
    s = rows["svd_jacobi"]
    assert s["gb_per_s"] == round((2 * 1000 + 4000) * 1024 / 1_000_000, 1)
    assert list(rows) == ["svd_jacobi", "zgemm_kernel"]      # by total time
    a = out["all_kernels"]
    assert a["dispatches"] == 3 and a["kernel_ns"] == 1_008_000
    assert a["bytes"] == round(zb + (2 * 1000 + 4000) * 1024)
    assert json.dumps(out)  # serialisable as the bench prints it


def test_pmc_small_summary_rejects_missing_marker(bench, tmp_path):
    copy_kib = bench.PMC_COPY_ELEMS * 16 / 1024
    rows = [(MARK, 0, 10), ("zgemm_kernel", 5.0, 100), (COPY, copy_kib, 1000)]
    _write_pass(tmp_path / "f", [(k, v / 2 if k == COPY else v, d) for k, v, d in rows],
                "FETCH_SIZE")
    _write_pass(tmp_path / "w", rows, "WRITE_SIZE")
    with pytest.raises(SystemExit, match="markers"):
        bench.pmc_small_summary(str(tmp_path / "f"), str(tmp_path / "w"))


def test_pmc_summary_splits_forms(bench, tmp_path):
    """--pmc-workload: the chosen form lies between the first two of the last
    three markers, the plain form between the last two; bytes per dispatch
    are averaged over the kernel's dispatches in each form."""
    copy_kib = bench.PMC_COPY_ELEMS * 16 / 1024
    seq = [(MARK, 0, 0), ("zgemm_kernel", 10.0, 4.0), ("zgemm_kernel", 30.0, 8.0),
           (MARK, 0, 0), ("zgemm_kernel", 100.0, 50.0), ("svd_jacobi", 7.0, 9.0),
           (MARK, 0, 0), (COPY, copy_kib / 2, copy_kib)]
    _write_pass(tmp_path / "f", [(k, f, 100) for k, f, _ in seq], "FETCH_SIZE")
    _write_pass(tmp_path / "w", [(k, w, 100) for k, _, w in seq], "WRITE_SIZE")
    out = bench.pmc_summary(str(tmp_path / "f"), str(tmp_path / "w"))
    ch, pl = out["forms"]["chosen"]["zgemm_kernel"], out["forms"]["plain"]["zgemm_kernel"]
    assert ch["dispatches"] == 2 and pl["dispatches"] == 1
    assert ch["fetch_bytes"] == round(2 * (10 + 30) / 2 * 1024)
    assert ch["write_bytes"] == round((4 + 8) / 2 * 1024)
    assert pl["traffic_bytes"] == round((2 * 100 + 50) * 1024)
    assert "svd_jacobi" not in out["forms"]["chosen"]
    assert out["forms"]["plain"]["svd_jacobi"]["fetch_bytes"] == round(2 * 7 * 1024)


def test_committed_profiles_parse(bench):
    """The committed summaries hold what bench.py reads from them: the empty
    kernel's duration and the dominant kernel's bytes and MFMA work per
    dispatch of the chi = 256 update."""
    assert 0.5 < bench.read_empty_dispatch_us() < 5.0
    traffic, flops = bench.read_traffic("zgemm_kernel", "chosen")
    assert traffic is not None and 1e6 < traffic < 1e9
    assert flops is not None and flops > 0
    plain_bytes, _ = bench.read_traffic("svd_jacobi", "plain")
    assert plain_bytes > traffic
