"""Python handle on the seeded synthetic workload generator (csrc/synth.cpp)."""
import ctypes as C
import os

import numpy as np

from . import abi

_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "libfpsynth.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing; run `python -m firepony_amd.build`")
        L = C.CDLL(path)
        L.fpsyn_genome_create.restype = C.c_void_p
        L.fpsyn_genome_create.argtypes = [C.POINTER(abi.FpsynGenomeSpec)]
        L.fpsyn_genome_destroy.argtypes = [C.c_void_p]
        L.fpsyn_genome_reference.argtypes = [C.c_void_p, C.POINTER(abi.FpReferenceView)]
        L.fpsyn_genome_variants.argtypes = [C.c_void_p, C.POINTER(abi.FpVariantsView)]
        L.fpsyn_genome_header.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(abi.FpHeaderView)]
        L.fpsyn_batch_create.restype = C.c_void_p
        L.fpsyn_batch_create.argtypes = [C.c_void_p, C.POINTER(abi.FpsynReadSpec), C.c_uint64,
                                         C.c_uint32, C.c_uint32]
        L.fpsyn_batch_view.argtypes = [C.c_void_p, C.POINTER(abi.FpBatchView)]
        L.fpsyn_batch_destroy.argtypes = [C.c_void_p]
        L.fpsyn_named_spec.restype = C.c_int
        L.fpsyn_named_spec.argtypes = [C.c_int, C.POINTER(abi.FpsynGenomeSpec),
                                       C.POINTER(abi.FpsynReadSpec), C.POINTER(C.c_uint64)]
        _LIB = L
    return _LIB


def named_spec(config):
    gs, rs, n = abi.FpsynGenomeSpec(), abi.FpsynReadSpec(), C.c_uint64()
    if lib().fpsyn_named_spec(config, C.byref(gs), C.byref(rs), C.byref(n)) != 0:
        raise ValueError(f"unknown workload {config}")
    return gs, rs, int(n.value)


class Genome:
    """Reference contigs + known sites; owns the native storage."""

    def __init__(self, spec, num_read_groups):
        self._h = lib().fpsyn_genome_create(C.byref(spec))
        if not self._h:
            raise ValueError("bad genome spec")
        self.spec = spec
        self.reference = abi.FpReferenceView()
        self.variants = abi.FpVariantsView()
        self.header = abi.FpHeaderView()
        lib().fpsyn_genome_reference(self._h, C.byref(self.reference))
        lib().fpsyn_genome_variants(self._h, C.byref(self.variants))
        lib().fpsyn_genome_header(self._h, num_read_groups, C.byref(self.header))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().fpsyn_genome_destroy(self._h)
            self._h = None


class Batch:
    """Reads [first, first + n) of a workload, in the boundary layout."""

    def __init__(self, genome, read_spec, first, n, threads=8):
        self.genome = genome
        self._h = lib().fpsyn_batch_create(genome._h, C.byref(read_spec), first, n, threads)
        if not self._h:
            raise ValueError("bad read spec")
        self.view = abi.FpBatchView()
        lib().fpsyn_batch_view(self._h, C.byref(self.view))

    def arrays(self):
        """numpy views of every array of the batch (no copies)."""
        return abi.view_arrays(self.view)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().fpsyn_batch_destroy(self._h)
            self._h = None


def small_workload(seed=1, num_reads=2000, edge=0.0, lengths=(100,), read_groups=2,
                   contigs=(200000,), sites=400, clip=0.1, indel=0.1, paired=0,
                   n_runs=0.0, multi_sites=0.0, mismatch=0.01, q2_tail=0.05):
    """A small custom workload for parity tests; returns (genome, read_spec, n)."""
    gs = abi.FpsynGenomeSpec()
    gs.seed = seed
    gs.num_contigs = len(contigs)
    for i, c in enumerate(contigs):
        gs.contig_length[i] = c
    gs.num_sites = sites
    gs.n_run_fraction = n_runs
    gs.multi_base_site_frac = multi_sites
    rs = abi.FpsynReadSpec()
    rs.seed = seed + 1000
    rs.num_lengths = len(lengths)
    for i, L in enumerate(lengths):
        rs.lengths[i] = L
    rs.num_read_groups = read_groups
    for i in range(read_groups):
        rs.rg_mismatch[i] = mismatch * (1 + i)
    rs.indel_read_frac = indel
    rs.indel_min, rs.indel_max = 1, 10
    rs.clip_read_frac = clip
    rs.clip_min, rs.clip_max = 0, 20
    rs.q2_tail_frac = q2_tail
    rs.q2_tail_max = 20
    rs.paired = paired
    rs.insert_mean, rs.insert_sd = 300.0, 60.0
    rs.edge_case_frac = edge
    return Genome(gs, read_groups), rs, num_reads
