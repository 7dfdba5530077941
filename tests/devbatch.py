"""Device-resident copies of generated batches, made with the HIP runtime the
engine library itself uses (one runtime per process)."""
import ctypes as C

import numpy as np

from firepony_amd import abi

_HIP = None


def hip():
    global _HIP
    if _HIP is None:
        _HIP = C.CDLL("libamdhip64.so")
        _HIP.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _HIP.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _HIP.hipFree.argtypes = [C.c_void_p]
        _HIP.hipHostMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t, C.c_uint]
        _HIP.hipHostFree.argtypes = [C.c_void_p]
    return _HIP


FIELDS = [("chromosome", abi.u16p), ("alignment_start", abi.u32p), ("alignment_stop", abi.u32p),
          ("mate_chromosome", abi.u32p), ("mate_alignment_start", abi.u32p),
          ("inferred_insert_size", abi.i32p), ("cigars", abi.u32p), ("cigar_start", abi.u32p),
          ("cigar_len", abi.u32p), ("reads", abi.u32p), ("read_start", abi.u32p), ("read_len", abi.u32p),
          ("qualities", abi.u8p), ("qual_start", abi.u32p), ("qual_len", abi.u32p), ("flags", abi.u16p),
          ("mapq", abi.u8p), ("read_group", abi.u32p)]


class DeviceCopy:
    """fp_batch_view whose arrays live in device memory; frees them on close."""

    def __init__(self, batch):
        arrs = batch.arrays()
        self.ptrs = []
        v = abi.FpBatchView()
        hv = batch.view
        v.num_reads, v.num_cigar_ops = hv.num_reads, hv.num_cigar_ops
        v.num_read_bases, v.num_qualities = hv.num_read_bases, hv.num_qualities
        for name, ptype in FIELDS:
            a = np.ascontiguousarray(arrs[name])
            p = C.c_void_p()
            nbytes = max(a.nbytes, 4)
            if hip().hipMalloc(C.byref(p), nbytes) != 0:
                raise RuntimeError("hipMalloc failed")
            self.ptrs.append(p)
            if a.nbytes and hip().hipMemcpy(p, a.ctypes.data, a.nbytes, 1) != 0:
                raise RuntimeError("hipMemcpy failed")
            setattr(v, name, C.cast(p, ptype))
        self.view = v

    def close(self):
        for p in self.ptrs:
            hip().hipFree(p)
        self.ptrs = []


class PinnedCopy:
    """fp_batch_view whose arrays live in pinned host memory (hipHostMalloc); frees them on close.
    The engine reads such a batch over the host link in one copy launch."""

    def __init__(self, batch):
        arrs = batch.arrays()
        self.ptrs = []
        v = abi.FpBatchView()
        hv = batch.view
        v.num_reads, v.num_cigar_ops = hv.num_reads, hv.num_cigar_ops
        v.num_read_bases, v.num_qualities = hv.num_read_bases, hv.num_qualities
        for name, ptype in FIELDS:
            a = np.ascontiguousarray(arrs[name])
            p = C.c_void_p()
            if hip().hipHostMalloc(C.byref(p), max(a.nbytes, 4), 0) != 0:
                raise RuntimeError("hipHostMalloc failed")
            self.ptrs.append(p)
            C.memmove(p, a.ctypes.data, a.nbytes)
            setattr(v, name, C.cast(p, ptype))
        self.view = v

    def close(self):
        for p in self.ptrs:
            hip().hipHostFree(p)
        self.ptrs = []
