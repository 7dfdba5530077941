"""The C ABI library: loads without a GPU, exports every declared entry point,
matches the ctypes layouts, and its host-side numeric helpers are exact."""
import ctypes as C
import os
import re
import subprocess
from fractions import Fraction

import numpy as np
import pytest

from firepony_amd import abi, engine
from oracle import pyoracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "firepony_amd.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:fp_status|void|const char \*)\s*\*?\s*(fp_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(engine.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 18
    for n in names:
        assert hasattr(lib, n), n
    assert set(engine.EXPORTED) <= set(names)


def test_create_without_gpu_fails_cleanly():
    # on a machine without a HIP device the engine must refuse, not fall back
    if os.environ.get("HIP_VISIBLE_DEVICES") is None and os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is visible here")
    with pytest.raises(engine.FireponyError):
        from firepony_amd import synth
        g, rs, n = synth.small_workload(num_reads=10)
        engine.Engine(g)


LAYOUT_PROBE = r"""
#include <stdio.h>
#include <stddef.h>
#include "firepony_amd.h"
#define S(t) printf(#t " %zu\n", sizeof(t));
#define O(t, f) printf(#t "." #f " %zu\n", offsetof(t, f));
int main(void) {
    S(fp_options) S(fp_header_view) S(fp_reference_view) S(fp_variants_view) S(fp_batch_view)
    S(fp_table_out) S(fp_stats)
    O(fp_batch_view, num_cigar_ops) O(fp_batch_view, read_group) O(fp_batch_view, num_qualities)
    O(fp_stats, baq_cells) O(fp_stats, prep_kernel_ms) O(fp_table_out, empirical_quality)
    O(fp_variants_view, num_features) O(fp_options, sort_digit_bits)
    return 0;
}
"""


def test_ctypes_layout_matches_header(tmp_path):
    src = tmp_path / "probe.c"
    src.write_text(LAYOUT_PROBE)
    exe = tmp_path / "probe"
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    out = dict(line.rsplit(" ", 1) for line in subprocess.check_output([str(exe)]).decode().split("\n") if line)
    py = {"fp_options": abi.FpOptions, "fp_header_view": abi.FpHeaderView, "fp_reference_view": abi.FpReferenceView,
          "fp_variants_view": abi.FpVariantsView, "fp_batch_view": abi.FpBatchView, "fp_table_out": abi.FpTableOut,
          "fp_stats": abi.FpStats}
    for k, v in out.items():
        if "." in k:
            t, f = k.split(".")
            assert getattr(py[t], f).offset == int(v), k
        else:
            assert C.sizeof(py[k]) == int(v), k


def test_fixed_point_is_exact():
    lib = C.CDLL(engine.LIB_PATH)
    lib.fp_selftest_fixed.argtypes = [abi.f64p, C.c_uint64, abi.u64p, abi.u64p, abi.f64p]
    vals = [0.0, 1.0, 0.5, 1 / 3, 2 / 7, 1 / 250, 3 / 65535, 1 / 65535, float(np.float32(1 / 3)),
            float(np.float32(5 / 251))]
    rng = np.random.default_rng(0)
    for n in rng.integers(1, 600, 200):
        vals.append(int(rng.integers(0, n + 1)) / int(n))
    v = np.array(vals, np.float64)
    lo = np.zeros(len(v), np.uint64)
    hi = np.zeros(len(v), np.uint64)
    back = np.zeros(len(v), np.float64)
    lib.fp_selftest_fixed(v.ctypes.data_as(abi.f64p), len(v), lo.ctypes.data_as(abi.u64p),
                          hi.ctypes.data_as(abi.u64p), back.ctypes.data_as(abi.f64p))
    for x, l, h, b in zip(vals, lo, hi, back):
        fixed = (int(h) << 64) | int(l)
        assert Fraction(fixed, 2 ** 80) == Fraction(x), x       # exact representation
        assert b == x


def test_fixed_to_double_rounds_to_nearest_even():
    lib = C.CDLL(engine.LIB_PATH)
    lib.fp_selftest_fixed_to_double.restype = C.c_double
    lib.fp_selftest_fixed_to_double.argtypes = [C.c_uint64, C.c_uint64]
    rng = np.random.default_rng(1)
    cases = [(1 << 53) + 1, (1 << 53) + 3, (1 << 54) + 2, (1 << 54) + 6, (1 << 100) + (1 << 47),
             (1 << 100) + (3 << 47), (1 << 127) + 12345, 2 ** 128 - 1]
    cases += [int(rng.integers(0, 2 ** 63)) << int(rng.integers(0, 64)) | int(rng.integers(0, 2 ** 63))
              for _ in range(300)]
    for x in cases:
        x &= 2 ** 128 - 1
        got = lib.fp_selftest_fixed_to_double(x & (2 ** 64 - 1), x >> 64)
        # Python's int -> float conversion rounds to nearest even; 2^-80 scaling is exact
        assert got == float(Fraction(x, 2 ** 80)), x


def test_map_thresholds_reproduce_host_formula():
    lib = C.CDLL(engine.LIB_PATH)
    lib.fp_selftest_map_thresholds.argtypes = [abi.f64p]
    thr = np.zeros(128, np.float64)
    lib.fp_selftest_map_thresholds(thr.ctypes.data_as(abi.f64p))
    t = thr[5:102]
    assert np.all(np.diff(t) < 0)

    def via_thresholds(y):
        if y == 0.0 or not y <= thr[5]:
            return 4
        v = max(q for q in range(5, 102) if y <= thr[q])
        return 99 if v >= 101 else v

    rng = np.random.default_rng(2)
    maxes = list(1.0 - 10.0 ** rng.uniform(-15, 0, 3000)) + [1.0, 0.0, 0.5]
    for q in range(5, 102):     # both sides of every threshold
        for y in (thr[q], np.nextafter(thr[q], 1.0), np.nextafter(thr[q], 0.0)):
            maxes.append(1.0 - float(y))
    for m in maxes:
        y = 1.0 - m   # what the kernel evaluates
        assert via_thresholds(y) == pyoracle.lib().fpo_baq_map_quality(m), m
