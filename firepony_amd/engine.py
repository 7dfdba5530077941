"""ctypes binding of the firepony_amd C ABI (include/firepony_amd.h).

Mirrors the reference's firepony_pipeline (device/pipeline.h:44-66):
Engine(...) ~ firepony_pipeline::create + setup, process() ~ the worker loop
body, gather() ~ gather_intermediates, postprocess(), statistics().
The native library is required: every call goes to the HIP engine, and a
missing or unloadable library raises instead of falling back.
"""
import ctypes as C
import os

import numpy as np

from . import abi

_LIB = None
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "libfirepony_amd.so")

EXPORTED = [
    "fp_options_init", "fp_engine_create", "fp_engine_setup", "fp_engine_process_batch",
    "fp_engine_process_batch_device", "fp_engine_synchronize", "fp_engine_gather",
    "fp_engine_postprocess", "fp_engine_write_report", "fp_engine_get_table", "fp_engine_stats",
    "fp_engine_get_name", "fp_engine_get_total_memory", "fp_engine_debug_size",
    "fp_engine_debug_fetch", "fp_engine_reset", "fp_engine_destroy", "fp_last_error",
    "fp_status_string", "fp_job_open", "fp_job_submit", "fp_job_submit_device", "fp_job_close",
    "fp_job_write_report", "fp_job_get_table", "fp_job_stats", "fp_job_release",
    "fp_engine_set_read_group_names", "fp_engine_process_batch_device_noalloc",
]


class FireponyError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"firepony_amd status {status}: {msg}")
        self.status = status


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise FireponyError(-1, f"{LIB_PATH} is missing; run `python -m firepony_amd.build`")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.fp_options_init.argtypes = [C.POINTER(abi.FpOptions)]
        L.fp_options_init.restype = None
        L.fp_engine_create.argtypes = [C.c_int, C.POINTER(abi.FpOptions), C.POINTER(vp)]
        L.fp_engine_setup.argtypes = [vp, C.POINTER(abi.FpHeaderView), C.POINTER(abi.FpReferenceView),
                                      C.POINTER(abi.FpVariantsView)]
        L.fp_engine_process_batch.argtypes = [vp, C.POINTER(abi.FpBatchView)]
        L.fp_engine_process_batch_device.argtypes = [vp, C.POINTER(abi.FpBatchView)]
        L.fp_engine_process_batch_device_noalloc.argtypes = [vp, C.POINTER(abi.FpBatchView)]
        L.fp_engine_synchronize.argtypes = [vp]
        L.fp_engine_gather.argtypes = [vp, vp]
        L.fp_engine_postprocess.argtypes = [vp]
        L.fp_engine_write_report.argtypes = [vp, C.c_char_p]
        L.fp_engine_get_table.argtypes = [vp, C.c_int, C.POINTER(abi.FpTableOut)]
        L.fp_engine_stats.argtypes = [vp, C.POINTER(abi.FpStats)]
        L.fp_engine_get_name.argtypes = [vp, C.c_char_p, C.c_size_t]
        L.fp_engine_get_total_memory.argtypes = [vp, C.POINTER(C.c_uint64)]
        L.fp_engine_debug_size.argtypes = [vp, C.c_int, C.POINTER(C.c_uint64)]
        L.fp_engine_debug_fetch.argtypes = [vp, C.c_int, vp, C.c_uint64]
        L.fp_engine_reset.argtypes = [vp]
        L.fp_engine_destroy.argtypes = [vp]
        L.fp_engine_destroy.restype = None
        L.fp_last_error.restype = C.c_char_p
        L.fp_status_string.restype = C.c_char_p
        L.fp_status_string.argtypes = [C.c_int]
        L.fp_job_open.argtypes = [vp, C.c_uint32, C.POINTER(C.c_char_p), C.POINTER(vp)]
        L.fp_job_submit.argtypes = [vp, C.POINTER(abi.FpBatchView)]
        L.fp_job_submit_device.argtypes = [vp, C.POINTER(abi.FpBatchView)]
        L.fp_job_close.argtypes = [vp]
        L.fp_job_write_report.argtypes = [vp, C.c_char_p]
        L.fp_job_get_table.argtypes = [vp, C.c_int, C.POINTER(abi.FpTableOut)]
        L.fp_job_stats.argtypes = [vp, C.POINTER(abi.FpStats)]
        L.fp_job_release.argtypes = [vp]
        L.fp_job_release.restype = None
        L.fp_engine_set_read_group_names.argtypes = [vp, C.c_uint32, C.POINTER(C.c_char_p)]
        for name in EXPORTED:
            f = getattr(L, name)
            if f.restype is C.c_int and name not in ("fp_options_init",):
                f.restype = C.c_int
        _LIB = L
    return _LIB


def _check(rc):
    if rc != 0:
        raise FireponyError(rc, lib().fp_last_error().decode())


def default_options(**kw):
    o = abi.FpOptions()
    lib().fp_options_init(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class Engine:
    """One engine on one HIP device (firepony_pipeline on a compute device)."""

    def __init__(self, genome, device=0, **options):
        self.options = default_options(**options)
        h = C.c_void_p()
        _check(lib().fp_engine_create(device, C.byref(self.options), C.byref(h)))
        self._h = h
        self.genome = genome
        _check(lib().fp_engine_setup(self._h, C.byref(genome.header), C.byref(genome.reference),
                                     C.byref(genome.variants)))

    def process(self, batch):
        """Host batch: copy to the device and run the per-batch pipeline."""
        _check(lib().fp_engine_process_batch(self._h, C.byref(batch.view)))

    def process_device(self, view):
        """Batch whose arrays already live in this engine's device memory."""
        _check(lib().fp_engine_process_batch_device(self._h, C.byref(view)))

    def process_device_noalloc(self, view):
        """Device batch through the entry point that allocates no device memory (plain engines)."""
        _check(lib().fp_engine_process_batch_device_noalloc(self._h, C.byref(view)))

    def synchronize(self):
        _check(lib().fp_engine_synchronize(self._h))

    def gather(self, other):
        _check(lib().fp_engine_gather(self._h, other._h))

    def postprocess(self):
        _check(lib().fp_engine_postprocess(self._h))

    def reset(self):
        _check(lib().fp_engine_reset(self._h))

    def set_read_group_names(self, names):
        """Names by read-group id, as assigned while the reads were loaded."""
        arr = (C.c_char_p * max(1, len(names)))(*[n.encode() for n in names])
        _check(lib().fp_engine_set_read_group_names(self._h, len(names), arr))

    def table(self, table_id):
        out = abi.FpTableOut()
        _check(lib().fp_engine_get_table(self._h, table_id, C.byref(out)))
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
        _check(lib().fp_engine_get_table(self._h, table_id, C.byref(out)))
        return t

    def report(self, path):
        _check(lib().fp_engine_write_report(self._h, path.encode()))
        with open(path) as f:
            return f.read()

    def open_job(self, read_group_names):
        """Serving engines: admit a job with its own read groups.  Returns None when the
        job's tables do not fit the read-group slots now (retry after a release)."""
        names = (C.c_char_p * len(read_group_names))(*[n.encode() for n in read_group_names])
        h = C.c_void_p()
        rc = lib().fp_job_open(self._h, len(read_group_names), names, C.byref(h))
        if rc == abi.FP_ERR_BUDGET:
            return None
        _check(rc)
        return Job(self, h)

    def stats(self):
        s = abi.FpStats()
        _check(lib().fp_engine_stats(self._h, C.byref(s)))
        return s

    def name(self):
        buf = C.create_string_buffer(256)
        _check(lib().fp_engine_get_name(self._h, buf, 256))
        return buf.value.decode()

    def debug(self, item, dtype):
        n = C.c_uint64()
        _check(lib().fp_engine_debug_size(self._h, item, C.byref(n)))
        buf = np.zeros(int(n.value), np.uint8)
        if n.value:
            _check(lib().fp_engine_debug_fetch(self._h, item, buf.ctypes.data_as(C.c_void_p), n.value))
        return buf.view(dtype)

    def close(self):
        if getattr(self, "_h", None):
            lib().fp_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def _read_table(getter, table_id):
    out = abi.FpTableOut()
    _check(getter(table_id, C.byref(out)))
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
    _check(getter(table_id, C.byref(out)))
    return t


class Job:
    """A recalibration job on a serving engine (its tables live in the engine's read-group slots)."""

    def __init__(self, engine, handle):
        self.engine = engine
        self._h = handle

    def submit(self, batch):
        _check(lib().fp_job_submit(self._h, C.byref(batch.view)))

    def submit_device(self, view):
        _check(lib().fp_job_submit_device(self._h, C.byref(view)))

    def close(self):
        _check(lib().fp_job_close(self._h))

    def table(self, table_id):
        return _read_table(lambda t, o: lib().fp_job_get_table(self._h, t, o), table_id)

    def report(self, path):
        _check(lib().fp_job_write_report(self._h, path.encode()))
        with open(path) as f:
            return f.read()

    def stats(self):
        s = abi.FpStats()
        _check(lib().fp_job_stats(self._h, C.byref(s)))
        return s

    def release(self):
        if self._h:
            lib().fp_job_release(self._h)
            self._h = None
