"""Writes the golden fixtures of this directory.

The reference ships no golden vectors and cannot be built or run here, so the
fixtures are the oracle's outputs on seeded synthetic inputs, frozen: each
fixture holds a digest of its inputs (so a change of the generator is told
apart from a change of the pipeline), the per-read and per-base intermediates
of one batch, the three observation tables and the GATK report.

    python tests/golden/make_golden.py        # from the repository root

Cases:
  config0_first1000  the first 1,000 reads of the CPU-runnable workload
                     (1 Mbp contig, 100 bp reads, 1 % mismatches, 2 % indels)
  edge_menu_600      600 reads of the edge-case menu (paired reads, adapter
                     clipping, N runs, multi-base known sites, clips, indels,
                     three read groups, two contigs)
"""
import gzip
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from firepony_amd import abi, synth  # noqa: E402

PER_READ = [(abi.FP_DBG_ACTIVE_READS, np.uint32, "active_reads"), (abi.FP_DBG_READ_WINDOW, np.uint32, "read_window"),
            (abi.FP_DBG_NOINS_WINDOW, np.uint32, "noins_window"), (abi.FP_DBG_REF_WINDOW, np.uint32, "ref_window"),
            (abi.FP_DBG_NUM_ERRORS, np.uint16, "num_errors"), (abi.FP_DBG_HMM_WINDOW, np.uint32, "hmm_window"),
            (abi.FP_DBG_BANDWIDTH, np.uint16, "bandwidth"), (abi.FP_DBG_HQ_WINDOW, np.uint32, "hq_window")]
PER_BASE = [(abi.FP_DBG_ERROR_FLAGS, np.uint8, "error_flags"), (abi.FP_DBG_ACTIVE_MASK, np.uint8, "active_mask")]
TABLES = [(abi.FP_TABLE_QUALITY, "quality"), (abi.FP_TABLE_CYCLE, "cycle"), (abi.FP_TABLE_CONTEXT, "context")]


def cases():
    """(name, genome, read spec, first, count, options) of every fixture."""
    gs, rs, _ = synth.named_spec(0)
    g0 = synth.Genome(gs, rs.num_read_groups)
    yield "config0_first1000", g0, rs, 0, 1000, {}
    g1, rs1, _ = synth.small_workload(seed=7, num_reads=600, edge=0.35, lengths=(36, 100, 151), read_groups=3,
                                      contigs=(60000, 40000), sites=300, clip=0.3, indel=0.2, paired=1,
                                      n_runs=0.01, multi_sites=0.3)
    yield "edge_menu_600", g1, rs1, 0, 600, {"frac_error_f64": 1}


def input_digest(batch):
    h = hashlib.sha256()
    arrs = batch.arrays()
    for k in sorted(arrs):
        h.update(k.encode())
        h.update(np.ascontiguousarray(arrs[k]).tobytes())
    return h.hexdigest()


def collect(runner, batch, tmp_report):
    """The fixture arrays of one pipeline (oracle or engine) after one batch and postprocess."""
    out = {}
    for item, dt, name in PER_READ + PER_BASE:
        out[name] = runner.debug(item, dt)
    out["baq"] = runner.debug(abi.FP_DBG_BAQ, np.uint8)
    out["frac_errors"] = runner.debug(abi.FP_DBG_FRAC_ERRORS, np.float64)
    for t, name in TABLES:
        tb = runner.table(t)
        for k in ("keys", "observations", "mismatches_fixed_lo", "mismatches_fixed_hi"):
            out[f"{name}_{k}"] = np.asarray(tb[k])
    runner.postprocess()
    out["report"] = runner.report(tmp_report)
    return out


def main():
    from oracle.pyoracle import Oracle
    for name, g, rs, first, count, opts in cases():
        b = synth.Batch(g, rs, first, count)
        o = Oracle(g, frac_error_f64=bool(opts.get("frac_error_f64", 0)))
        o.process(b)
        st = o.stats()
        res = collect(o, b, os.path.join(HERE, f".{name}.tmp"))
        os.remove(os.path.join(HERE, f".{name}.tmp"))
        report = res.pop("report")
        arrays = {k: v for k, v in res.items()}
        arrays["stats"] = np.array([st.total_reads, st.filtered_reads, st.baq_reads], np.uint64)
        arrays["input_sha256"] = np.array(input_digest(b))
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **arrays)
        with gzip.open(os.path.join(HERE, f"{name}_report.txt.gz"), "wt") as f:
            f.write(report)
        print(f"{name}: {count} reads, {int(st.baq_reads)} BAQ reads, report {len(report)} bytes")


if __name__ == "__main__":
    main()
