"""The bench's per-kernel HBM traffic reduction (rocprofv3 FETCH_SIZE / WRITE_SIZE CSVs).

Synthetic counter files: a 512 MiB calibration copy, tail launches of two shapes (with and
without the deferred x update) and update launches one of which exited early.
"""
import csv
import importlib.util
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def _bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", ROOT / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _write(path, counter, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Kernel_Name", "Grid_Size", "Counter_Name",
                                          "Counter_Value"])
        w.writeheader()
        for name, grid, kib in rows:
            w.writerow({"Kernel_Name": name, "Grid_Size": grid, "Counter_Name": counter,
                        "Counter_Value": kib})


def test_traffic_filters_early_exits_per_launch_shape(tmp_path):
    bench = _bench()
    copy_kib = bench.COPY_BYTES / 1024
    tail = "void gdp::k_stail<float, true>(...)"
    upd = "void gdp::k_update<float, 16, 64, true>(...)"
    # FETCH reads half the copy's bytes (scale 2), WRITE all of them (scale 1)
    fetch_rows = [("__amd_rocclr_copyBuffer", 1 << 20, copy_kib / 2)]
    write_rows = [("__amd_rocclr_copyBuffer", 1 << 20, copy_kib)]
    # five tail launches with the x update (100 KiB fetched), one without (1 KiB)
    fetch_rows += [(tail, 131072, 100.0)] * 5 + [(tail, 1536, 1.0)]
    write_rows += [(tail, 131072, 50.0)] * 5 + [(tail, 1536, 0.5)]
    # three full update launches and one that exited early (left out of the mean)
    fetch_rows += [(upd, 786432, 80.0)] * 3 + [(upd, 786432, 0.2)]
    write_rows += [(upd, 786432, 40.0)] * 3 + [(upd, 786432, 0.1)]
    fcsv, wcsv = tmp_path / "fetch.csv", tmp_path / "write.csv"
    _write(fcsv, "FETCH_SIZE", fetch_rows)
    _write(wcsv, "WRITE_SIZE", write_rows)

    out = bench.reduce_traffic(str(fcsv), str(wcsv))
    assert out["calibration"]["FETCH_SIZE"]["scale"] == pytest.approx(2.0)
    assert out["calibration"]["WRITE_SIZE"]["scale"] == pytest.approx(1.0)
    k = out["kernels"]
    # both tail shapes count, as the tail's algorithmic bytes average over them
    assert k["tail"]["fetch_bytes"] == round((5 * 100.0 + 1.0) / 6 * 1024 * 2)
    assert k["tail"]["write_bytes"] == round((5 * 50.0 + 0.5) / 6 * 1024)
    # the early-exited update is dropped
    upd_k = k["update (x deferred)"]
    assert upd_k["fetch_bytes"] == round(80.0 * 1024 * 2)
    assert upd_k["hbm_bytes_per_launch"] == round(80.0 * 1024 * 2 + 40.0 * 1024)


def test_traffic_keeps_the_two_update_variants_apart(tmp_path):
    """k_update with its x half deferred moves about half the bytes of the full update
    on the same grid: the two are kinds of their own (as the plan profiler names them),
    each filtered against its own maximum, so the deferred launches are not taken for
    early exits."""
    bench = _bench()
    copy_kib = bench.COPY_BYTES / 1024
    full = "void gdp::k_update<float, 16, 64, false>(...)"
    deferred = "void gdp::k_update<float, 16, 64, true>(...)"
    fetch_rows = [("__amd_rocclr_copyBuffer", 1 << 20, copy_kib / 2)]
    write_rows = [("__amd_rocclr_copyBuffer", 1 << 20, copy_kib)]
    # one full update, four deferred ones (0.45 of the full bytes: below half of the
    # full update's count) and one deferred launch that exited early
    fetch_rows += [(full, 786432, 100.0)] + [(deferred, 786432, 45.0)] * 4 + \
        [(deferred, 786432, 0.1)]
    write_rows += [(full, 786432, 50.0)] + [(deferred, 786432, 22.5)] * 4 + \
        [(deferred, 786432, 0.05)]
    fcsv, wcsv = tmp_path / "fetch.csv", tmp_path / "write.csv"
    _write(fcsv, "FETCH_SIZE", fetch_rows)
    _write(wcsv, "WRITE_SIZE", write_rows)
    k = bench.reduce_traffic(str(fcsv), str(wcsv))["kernels"]
    assert k["update"]["fetch_bytes"] == round(100.0 * 1024 * 2)
    assert k["update"]["write_bytes"] == round(50.0 * 1024)
    assert k["update (x deferred)"]["fetch_bytes"] == round(45.0 * 1024 * 2)
    assert k["update (x deferred)"]["write_bytes"] == round(22.5 * 1024)


def test_streamed_and_sweep_order_down_kernels_rank_as_one_family():
    """A plan may run its finest down level with the streamed kernel and coarser ones with
    the sweep-order kernel: levels are ranked by area across both."""
    bench = _bench()
    names = {
        ("void gdp::k_sdown_stream<float, 2, 32>(...)", 4096): None,  # area 4096 * 32
        ("void gdp::k_sdown<float, 2, 4, false>(...)", 8192): None,   # area 8192 * 4
        ("void gdp::k_sdown_blk<float, 2, 8>(...)", 4096): None,      # area 4096 * 8 / 4
        ("void gdp::k_sdown<float, 2, 1, false>(...)", 4096): None,   # area 4096 * 1
    }
    grids = {"k_sdown<": {bench._level_area(n, g) for n, g in names}}
    kinds = [bench.kernel_kind(n, g, grids) for n, g in names]
    assert kinds == ["down L0", "down L1", "down L2", "down L3"]


def test_strip_and_block_tiled_up_kernels_rank_as_one_family():
    """The block-tiled up kernel's rows argument counts four strips: a level it runs ranks
    among the strip kernel's levels by area, and its dots flag is read like theirs."""
    bench = _bench()
    names = [
        ("void gdp::k_sup_blk<float, 2, 16, true>(...)", 16384),   # area 16384 * 16 / 4
        ("void gdp::k_sup<float, 2, 4, false, false>(...)", 4096),  # area 4096 * 4
        ("void gdp::k_sup_blk<float, 2, 32, false>(...)", 1024),   # area 1024 * 32 / 4
    ]
    grids = {"k_sup<": {bench._level_area(n, g) for n, g in names}}
    kinds = [bench.kernel_kind(n, g, grids) for n, g in names]
    assert kinds == ["up L0+dots", "up L1", "up L2"]


def test_fused_update_down_kernel_is_the_finest_level_with_the_update():
    """The streamed kernel carrying the PCG update (its fourth template argument) ranks
    with the plain one of the same level and gets the plan profiler's kind name."""
    bench = _bench()
    names = [
        ("void gdp::k_sdown_stream<float, 2, 16, true>(...)", 8192),   # area 8192 * 16
        ("void gdp::k_sdown_stream<float, 2, 16, false>(...)", 8192),  # the same level
        ("void gdp::k_sdown_blk<float, 2, 8, true>(...)", 32768),     # area 32768 * 8 / 4
        ("void gdp::k_sdown_blk<float, 2, 8, false>(...)", 4096),     # area 4096 * 8 / 4
    ]
    grids = {"k_sdown<": {bench._level_area(n, g) for n, g in names}}
    kinds = [bench.kernel_kind(n, g, grids) for n, g in names]
    assert kinds == ["down L0+update", "down L0", "down L1+update", "down L2"]
