"""ctypes binding of the input loaders (include/firepony_amd_io.h) and the
file-to-report flow built on them.

The reference's main program (firepony.cu:225-423) loads the reference
(loader/reference.cu), the known sites (loader/variants.cu), then streams the
alignment file in batches (io_thread.cu:97-147, loader/alignments.cu) through
its pipelines and writes the GATK report.  `recalibrate` does the same with
one engine: a loader thread reads batch k+1 while the engine runs batch k.
"""
import ctypes as C
import os
from concurrent.futures import ThreadPoolExecutor

from . import abi

_LIB = None
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "libfirepony_amd_io.so")

EXPORTED = [
    "fp_reference_load", "fp_reference_get_view", "fp_reference_lookup", "fp_reference_name",
    "fp_reference_free", "fp_variants_load", "fp_variants_get_view", "fp_variants_counts",
    "fp_variants_free", "fp_alignments_open", "fp_alignments_next_batch", "fp_alignments_header",
    "fp_alignments_counts", "fp_alignments_close", "fp_io_last_error",
]


class LoaderError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"firepony_amd loader status {status}: {msg}")
        self.status = status


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise LoaderError(-1, f"{LIB_PATH} is missing; run `python -m firepony_amd.build`")
        L = C.CDLL(LIB_PATH)
        vp, u64p = C.c_void_p, C.POINTER(C.c_uint64)
        L.fp_reference_load.argtypes = [C.c_char_p, C.POINTER(vp)]
        L.fp_reference_get_view.argtypes = [vp, C.POINTER(abi.FpReferenceView)]
        L.fp_reference_lookup.argtypes = [vp, C.c_char_p]
        L.fp_reference_lookup.restype = C.c_uint32
        L.fp_reference_name.argtypes = [vp, C.c_uint32, C.c_char_p, C.c_size_t]
        L.fp_reference_free.argtypes = [vp]
        L.fp_reference_free.restype = None
        L.fp_variants_load.argtypes = [C.c_char_p, vp, C.POINTER(vp)]
        L.fp_variants_get_view.argtypes = [vp, C.POINTER(abi.FpVariantsView)]
        L.fp_variants_counts.argtypes = [vp, u64p, u64p]
        L.fp_variants_free.argtypes = [vp]
        L.fp_variants_free.restype = None
        L.fp_alignments_open.argtypes = [C.c_char_p, vp, C.POINTER(vp)]
        L.fp_alignments_next_batch.argtypes = [vp, C.c_uint32, C.POINTER(abi.FpBatchView)]
        L.fp_alignments_header.argtypes = [vp, C.POINTER(abi.FpHeaderView), C.POINTER(C.c_uint32)]
        L.fp_alignments_counts.argtypes = [vp, u64p, u64p]
        L.fp_alignments_close.argtypes = [vp]
        L.fp_alignments_close.restype = None
        L.fp_io_last_error.restype = C.c_char_p
        _LIB = L
    return _LIB


def _check(rc):
    if rc != 0:
        raise LoaderError(rc, lib().fp_io_last_error().decode(errors="replace"))


class Reference:
    """A FASTA reference (reference_file_handle, loader/reference.cu)."""

    def __init__(self, path):
        h = C.c_void_p()
        _check(lib().fp_reference_load(os.fsencode(path), C.byref(h)))
        self._h = h
        self.view = abi.FpReferenceView()
        _check(lib().fp_reference_get_view(self._h, C.byref(self.view)))

    def lookup(self, name):
        i = lib().fp_reference_lookup(self._h, name.encode())
        return None if i == 0xFFFFFFFF else int(i)

    def name(self, i):
        buf = C.create_string_buffer(4096)
        _check(lib().fp_reference_name(self._h, i, buf, len(buf)))
        return buf.value.decode()

    def __len__(self):
        return int(self.view.num_sequences)

    def close(self):
        if getattr(self, "_h", None):
            lib().fp_reference_free(self._h)
            self._h = None

    def __del__(self):
        self.close()


class KnownSites:
    """A VCF of known sites (load_vcf, loader/variants.cu)."""

    def __init__(self, path, reference):
        h = C.c_void_p()
        _check(lib().fp_variants_load(os.fsencode(path), reference._h, C.byref(h)))
        self._h = h
        self.reference = reference
        self.view = abi.FpVariantsView()
        _check(lib().fp_variants_get_view(self._h, C.byref(self.view)))

    def counts(self):
        r, s = C.c_uint64(), C.c_uint64()
        _check(lib().fp_variants_counts(self._h, C.byref(r), C.byref(s)))
        return int(r.value), int(s.value)

    def close(self):
        if getattr(self, "_h", None):
            lib().fp_variants_free(self._h)
            self._h = None

    def __del__(self):
        self.close()


class LoadedBatch:
    """One batch read by an Alignments handle: `view` is the boundary view, valid until the
    handle loads its second batch after this one."""

    def __init__(self, view):
        self.view = view

    @property
    def num_reads(self):
        return int(self.view.num_reads)

    def arrays(self):
        return abi.view_arrays(self.view)


class Alignments:
    """A SAM or BAM file read in batches (alignment_file, loader/alignments.cu)."""

    def __init__(self, path, reference):
        h = C.c_void_p()
        _check(lib().fp_alignments_open(os.fsencode(path), reference._h, C.byref(h)))
        self._h = h
        self.reference = reference

    def next_batch(self, batch_size):
        v = abi.FpBatchView()
        _check(lib().fp_alignments_next_batch(self._h, batch_size, C.byref(v)))
        return LoadedBatch(v) if v.num_reads else None

    def header(self):
        """(header view, read-group names by id so far, upper bound of read-group ids)"""
        h, m = abi.FpHeaderView(), C.c_uint32()
        _check(lib().fp_alignments_header(self._h, C.byref(h), C.byref(m)))
        names = [h.read_group_names[i].decode() for i in range(h.num_read_groups)]
        return h, names, int(m.value)

    def counts(self):
        r, u = C.c_uint64(), C.c_uint64()
        _check(lib().fp_alignments_counts(self._h, C.byref(r), C.byref(u)))
        return int(r.value), int(u.value)

    def close(self):
        if getattr(self, "_h", None):
            lib().fp_alignments_close(self._h)
            self._h = None

    def __del__(self):
        self.close()


class FileInputs:
    """The reference, known sites and alignment header of one run, in the form the engine's
    setup takes (`.header`, `.reference`, `.variants`).  The header's read-group count is the
    file's upper bound of read-group ids; their names are known once the reads are loaded."""

    def __init__(self, alignments_path, reference_path, known_sites_path):
        self.ref = Reference(reference_path)
        self.sites = KnownSites(known_sites_path, self.ref)
        self.alignments = Alignments(alignments_path, self.ref)
        h, _, max_rg = self.alignments.header()
        self.max_read_groups = max(1, max_rg)
        self._placeholder = (C.c_char_p * self.max_read_groups)(*([b""] * self.max_read_groups))
        self.header = abi.FpHeaderView()
        self.header.num_chromosomes = h.num_chromosomes
        self.header.chromosome_lengths = h.chromosome_lengths
        self.header.num_read_groups = self.max_read_groups
        self.header.read_group_names = self._placeholder
        self.reference = self.ref.view
        self.variants = self.sites.view

    def batches(self, batch_size=20000):
        """Batches of the alignment file; the next one loads on a worker thread while the
        caller uses the current one (io_thread.cu:97-147)."""
        with ThreadPoolExecutor(max_workers=1) as ex:
            fut = ex.submit(self.alignments.next_batch, batch_size)
            while True:
                b = fut.result()
                if b is None:
                    return
                fut = ex.submit(self.alignments.next_batch, batch_size)
                yield b

    def read_group_names(self):
        return self.alignments.header()[1]


def recalibrate(alignments_path, reference_path, known_sites_path, report_path, batch_size=20000,
                device=0, **engine_options):
    """BQSR+BAQ tables and GATK report of one alignment file (firepony.cu:225-423); returns the
    engine's statistics.  engine_options are fp_options fields."""
    from .engine import Engine
    inp = FileInputs(alignments_path, reference_path, known_sites_path)
    e = Engine(inp, device=device, **engine_options)
    try:
        for b in inp.batches(batch_size):
            e.process(b)
        e.synchronize()
        e.set_read_group_names(inp.read_group_names())
        e.postprocess()
        e.report(report_path)
        return e.stats()
    finally:
        e.close()
