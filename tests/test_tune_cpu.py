"""The autotuner's host side, without a GPU: the table rows the sweep tool
derives from a tuning file, the compiled-in table, and the choice API of the
C ABI (w90gpu_tune_get / w90gpu_tune_set touch no device)."""
import os
import re

import pytest

from wannier90_amd import tune

CLASS_KERNELS = {1: {"fused32", "fused64"}, 2: {"fused64", "direct64", "fused48"}, 3: {"gemm64", "gemm32"}}


def test_nn_ranges_split_at_midpoints():
    assert tune.nn_ranges([6, 8, 12]) == {6: (1, 7), 8: (8, 10), 12: (11, 12)}
    assert tune.nn_ranges([8]) == {8: (1, 12)}


def test_table_rows_merge_runs_of_non_default_choices():
    ch = {(30, 6): ("fused64", {}), (31, 6): ("fused32", {}), (32, 6): ("fused64", {}),
          (33, 6): ("direct64", {}), (70, 6): ("gemm32", {}), (71, 6): ("gemm32", {}),
          (73, 6): ("gemm32", {}), (72, 8): ("gemm32", {}), (100, 8): ("gemm64", {})}
    rows = tune.table_rows(ch)
    assert rows == [(30, 30, 1, 7, "fused64"), (32, 32, 1, 7, "fused64"), (33, 33, 1, 7, "direct64"),
                    (70, 71, 1, 7, "gemm32"), (73, 73, 1, 7, "gemm32"), (72, 72, 8, 12, "gemm32")]


def test_a_run_does_not_cross_a_class_boundary():
    ch = {(32, 12): ("fused64", {}), (33, 12): ("fused64", {})}
    # fused64 is the 33-64 default: only the 32-band row is kept
    assert tune.table_rows(ch) == [(32, 32, 1, 12, "fused64")]


def test_read_choices_parses_the_engine_format(tmp_path):
    f = tmp_path / "t.txt"
    f.write_text("# comment\n70 6 gemm32 gemm64=812.50 gemm32=640.25 items=2048\n40 8 fused64 fused64=10.0 "
                 "direct64=11.0 items=2048\n")
    ch = tune.read_choices(str(f))
    assert ch[(70, 6)] == ("gemm32", {"gemm64": 812.5, "gemm32": 640.25})
    assert ch[(40, 8)][0] == "fused64"


def _table_rows_in_header():
    text = open(tune.HEADER).read()
    return [(int(a), int(b), int(c), int(d), k.lower())
            for a, b, c, d, k in re.findall(r"\{(\d+), (\d+), (\d+), (\d+), RK_(\w+)\}", text)]


def test_compiled_table_rows_are_valid():
    for lo, hi, nlo, nhi, k in _table_rows_in_header():
        cls = tune.wann_class(lo)
        assert cls >= 1 and tune.wann_class(hi) == cls and lo <= hi
        assert 1 <= nlo <= nhi <= 12
        assert k in CLASS_KERNELS[cls] and k != tune.DEFAULTS[cls]


def test_choice_api_follows_table_and_fixed_choices():
    from wannier90_amd import capi

    if not os.path.exists(capi.LIB_PATH):
        pytest.skip("libw90gpu.so not built")
    # defaults where the table has no row, and the table's rows
    assert capi.tune_get(8, 6) == "w16"
    covered = {(nw, nn): k for lo, hi, nlo, nhi, k in _table_rows_in_header()
               for nw in range(lo, hi + 1) for nn in range(nlo, nhi + 1)}
    for nw in (20, 40, 80, 128):
        for nn in (6, 8, 12):
            want = covered.get((nw, nn), tune.DEFAULTS[tune.wann_class(nw)])
            assert capi.tune_get(nw, nn) == want
    capi.tune_set(48, 6, "fused48")
    assert capi.tune_get(48, 6) == "fused48"
    capi.tune_set(48, 6, None)
    # a fixed choice, then back
    before = capi.tune_get(80, 6)
    other = ({"gemm64", "gemm32"} - {before}).pop()
    capi.tune_set(80, 6, other)
    assert capi.tune_get(80, 6) == other
    capi.tune_set(80, 6, None)
    assert capi.tune_get(80, 6) == before
    # a kernel outside the class, a 48 tile for more than 48 bands, an
    # unknown kernel, an invalid key
    for bad in ((80, 6, "fused32"), (20, 6, "gemm32"), (49, 6, "fused48"), (40, 6, "nope"), (0, 6, "w16"),
                (20, 13, "fused32")):
        with pytest.raises(capi.W90Error):
            capi.tune_set(*bad)
