"""ctypes wrapper of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product never does.
"""
import ctypes as C
import os

import numpy as np

from firepony_amd import abi

_LIB = None
HERE = os.path.dirname(os.path.abspath(__file__))

ACCUM_EXACT = 0
ACCUM_REF_FLOAT = 1


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "_build", "libfp_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing; run `bash oracle/build.sh`")
        L = C.CDLL(path)
        L.fpo_create.restype = C.c_void_p
        L.fpo_create.argtypes = [C.c_int, C.c_int, C.c_int]
        L.fpo_setup.argtypes = [C.c_void_p, C.POINTER(abi.FpHeaderView),
                                C.POINTER(abi.FpReferenceView), C.POINTER(abi.FpVariantsView)]
        L.fpo_process_batch.argtypes = [C.c_void_p, C.POINTER(abi.FpBatchView)]
        L.fpo_postprocess.argtypes = [C.c_void_p]
        L.fpo_gather.argtypes = [C.c_void_p, C.c_void_p]
        L.fpo_set_read_group_names.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(C.c_char_p)]
        L.fpo_get_table.argtypes = [C.c_void_p, C.c_int, C.POINTER(abi.FpTableOut)]
        L.fpo_write_report.argtypes = [C.c_void_p, C.c_char_p]
        L.fpo_stats.argtypes = [C.c_void_p, C.POINTER(abi.FpStats)]
        L.fpo_debug_size.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_uint64)]
        L.fpo_debug_fetch.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_uint64]
        L.fpo_debug_baq_real.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_uint64]
        L.fpo_destroy.argtypes = [C.c_void_p]
        L.fpo_round_n.restype = C.c_double
        L.fpo_round_n.argtypes = [C.c_double, C.c_int]
        L.fpo_baq_map_quality.restype = C.c_int
        L.fpo_baq_map_quality.argtypes = [C.c_double]
        L.fpo_empirical_quality.restype = C.c_double
        L.fpo_empirical_quality.argtypes = [C.c_uint64, C.c_double, C.c_double, C.c_int,
                                            C.POINTER(C.c_double)]
        L.fpo_hmm_glocal.argtypes = [abi.u8p, C.c_int, abi.u8p, abi.u8p, C.c_int, C.c_int,
                                     abi.u8p, abi.u32p]
        L.fpo_fractional_errors.argtypes = [abi.u8p, abi.u8p, C.c_int, abi.f64p]
        _LIB = L
    return _LIB


def read_table(getter, handle, table_id):
    """Fetch a table through a get_table-style entry point into numpy arrays."""
    out = abi.FpTableOut()
    rc = getter(handle, table_id, C.byref(out))
    if rc != 0:
        raise RuntimeError(f"get_table failed: {rc}")
    n = int(out.count)
    t = {
        "keys": np.zeros(n, np.uint32),
        "observations": np.zeros(n, np.uint64),
        "mismatches": np.zeros(n, np.float64),
        "mismatches_fixed_lo": np.zeros(n, np.uint64),
        "mismatches_fixed_hi": np.zeros(n, np.uint64),
        "expected_errors": np.zeros(n, np.float64),
        "estimated_quality": np.zeros(n, np.float64),
        "empirical_quality": np.zeros(n, np.float64),
    }
    out = abi.FpTableOut()
    out.count = n
    for k, v in t.items():
        ctype = {np.uint32: abi.u32p, np.uint64: abi.u64p, np.float64: abi.f64p}[v.dtype.type]
        setattr(out, k, v.ctypes.data_as(ctype))
    rc = getter(handle, table_id, C.byref(out))
    if rc != 0:
        raise RuntimeError(f"get_table failed: {rc}")
    return t


class Oracle:
    def __init__(self, genome, frac_error_f64=False, accumulate=ACCUM_EXACT,
                 disable_output_rounding=False):
        self._h = lib().fpo_create(int(frac_error_f64), accumulate, int(disable_output_rounding))
        self.genome = genome
        rc = lib().fpo_setup(self._h, C.byref(genome.header), C.byref(genome.reference),
                             C.byref(genome.variants))
        if rc:
            raise RuntimeError(f"fpo_setup: {rc}")

    def process(self, batch):
        rc = lib().fpo_process_batch(self._h, C.byref(batch.view))
        if rc:
            raise RuntimeError(f"fpo_process_batch: {rc}")

    def postprocess(self):
        lib().fpo_postprocess(self._h)

    def gather(self, other):
        """Add other's observation tables and read counters into this oracle (exact mode)."""
        rc = lib().fpo_gather(self._h, other._h)
        if rc:
            raise RuntimeError(f"fpo_gather: {rc}")

    def set_read_group_names(self, names):
        arr = (C.c_char_p * max(1, len(names)))(*[n.encode() for n in names])
        rc = lib().fpo_set_read_group_names(self._h, len(names), arr)
        if rc:
            raise RuntimeError(f"fpo_set_read_group_names: {rc}")

    def table(self, table_id):
        return read_table(lib().fpo_get_table, self._h, table_id)

    def report(self, path):
        rc = lib().fpo_write_report(self._h, path.encode())
        if rc:
            raise RuntimeError(f"fpo_write_report: {rc}")
        with open(path) as f:
            return f.read()

    def stats(self):
        s = abi.FpStats()
        lib().fpo_stats(self._h, C.byref(s))
        return s

    def debug(self, item, dtype):
        n = C.c_uint64()
        lib().fpo_debug_size(self._h, item, C.byref(n))
        buf = np.zeros(int(n.value), np.uint8)
        if n.value:
            lib().fpo_debug_fetch(self._h, item, buf.ctypes.data_as(C.c_void_p), n.value)
        return buf.view(dtype)

    def baq_real(self, nq):
        out = np.zeros(nq, np.float64)
        lib().fpo_debug_baq_real(self._h, out.ctypes.data_as(C.POINTER(C.c_double)), nq)
        return out

    def __del__(self):
        if getattr(self, "_h", None):
            lib().fpo_destroy(self._h)
            self._h = None
