"""The five benchmark workloads (synthetic, seeded) and their scoring setups.

C1  1,000 DNA from 20 ancestors (10% substitutions, 2% indels), lengths
    U{200..600}; EDNAFULL, open 16 / extend 4; PercentIdentity; Double distances.
C2  10,000 protein, background frequencies from a seeded Dirichlet(1),
    lengths round(lognormal(ln 300, 0.5)) clipped to [50, 2000]; BLOSUM62,
    open 10 / extend 1; PercentSimilarity; Int16 distances.
C3  50,000 DNA from 200 ancestors (15% / 3%), lengths 60% N(450,50) and
    40% N(1450,100) clipped to [100, 2000]; EDNAFULL 16/4; Int16.
C4  100,000 DNA mutated as C3 with C2's log-normal lengths; the 25,000-row
    stripe of one GPU; EDNAFULL 16/4; Int16; SWMS block files.
C5  serving: panel of 20,000 DNA (lengths U{300..1500}) + queries with
    lengths U{50..2000}; EDNAFULL 16/4.
"""
import numpy as np

from . import (ALIGN_NUCLEIC, ALIGN_PROTEIN, TDIST_DOUBLE, TDIST_INT16, load_scoring_matrix, synth_dna,
               synth_protein)

LEN_UNIFORM, LEN_BIMODAL, LEN_LOGNORMAL = 0, 1, 2

CONFIGS = {
    "c1": dict(kind="dna", seed=1, n=1000, n_anc=20, sub=0.10, indel=0.02, len_kind=LEN_UNIFORM,
               len_params=[200, 600], matrix="EDNAFULL", gap_open=16, gap_ext=4,
               alignment_type=ALIGN_NUCLEIC, tdistance=TDIST_DOUBLE),
    "c2": dict(kind="protein", seed=2, n=10000, len_kind=LEN_LOGNORMAL, len_params=[300, 0.5, 50, 2000],
               matrix="BLOSUM62", gap_open=10, gap_ext=1, alignment_type=ALIGN_PROTEIN, tdistance=TDIST_INT16),
    "c3": dict(kind="dna", seed=3, n=50000, n_anc=200, sub=0.15, indel=0.03, len_kind=LEN_BIMODAL,
               len_params=[0.6, 450, 50, 1450, 100, 100, 2000], matrix="EDNAFULL", gap_open=16, gap_ext=4,
               alignment_type=ALIGN_NUCLEIC, tdistance=TDIST_INT16),
    "c4": dict(kind="dna", seed=4, n=100000, n_anc=200, sub=0.15, indel=0.03, len_kind=LEN_LOGNORMAL,
               len_params=[300, 0.5, 50, 2000], matrix="EDNAFULL", gap_open=16, gap_ext=4,
               alignment_type=ALIGN_NUCLEIC, tdistance=TDIST_INT16),
    "c5_panel": dict(kind="dna", seed=5, n=20000, n_anc=200, sub=0.15, indel=0.03, len_kind=LEN_UNIFORM,
                     len_params=[300, 1500], matrix="EDNAFULL", gap_open=16, gap_ext=4,
                     alignment_type=ALIGN_NUCLEIC, tdistance=TDIST_INT16),
    "c5_queries": dict(kind="dna", seed=55, n=5000, n_anc=200, sub=0.15, indel=0.03, len_kind=LEN_UNIFORM,
                       len_params=[50, 2000], matrix="EDNAFULL", gap_open=16, gap_ext=4,
                       alignment_type=ALIGN_NUCLEIC, tdistance=TDIST_INT16),
}


def sequences(name, n=None):
    c = CONFIGS[name]
    n = c["n"] if n is None else n
    if c["kind"] == "dna":
        return synth_dna(c["seed"], n, c["n_anc"], c["sub"], c["indel"], c["len_kind"], c["len_params"])
    return synth_protein(c["seed"], n, c["len_kind"], c["len_params"])


def scoring(name):
    c = CONFIGS[name]
    return load_scoring_matrix(c["matrix"]), c["gap_open"], c["gap_ext"]


def block_cells(lengths, r0, r1, c0, c1, diag_lower=False):
    L = np.asarray(lengths, dtype=np.float64)
    rows, cols = L[r0:r1 + 1], L[c0:c1 + 1]
    if not diag_lower:
        return float(rows.sum() * cols.sum())
    tot = 0.0
    cum = np.concatenate([[0.0], np.cumsum(cols)])
    for k, li in enumerate(rows):
        tot += li * cum[k]
    return tot
