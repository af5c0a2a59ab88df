"""MI355X-native SALSA pairwise-distance path (all-pairs Smith-Waterman-Gotoh).

Python bindings (ctypes) over the two native libraries built in-tree:

* ``libswg.so``        -- the C ABI of ``include/swg.h`` (HIP kernels for gfx950)
* ``libsalsa_host.so`` -- the C++ host mirror of ``include/salsa_host.h``
  (block decomposition, matrix/block-file writers, scoring-matrix loader,
  synthetic workloads)

There is no CPU fallback: every compute entry point goes through the HIP
library and fails loudly when it is missing or when no GPU is present.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MATRIX_DIR = os.path.join(HERE, "matrices")

# return codes (include/swg.h)
SWG_OK = 0
SWG_E_ARGUMENT = -1
SWG_E_SYMBOL = -2
SWG_E_NONINTEGRAL = -3
SWG_E_DEVICE = -4
SWG_E_UNSUPPORTED = -5
SWG_E_JC_UNDEFINED = -6
SWG_E_CAPACITY = -7

ALIGN_PROTEIN, ALIGN_NUCLEIC = 0, 1
DIST_PERCENT_IDENTITY, DIST_KIMURA2, DIST_JUKES_CANTOR, DIST_MINMAX_NORM = 0, 1, 2, 3
MODE_SW_LOCAL, MODE_NW_GLOBAL = 0, 1
TDIST_INT16, TDIST_DOUBLE = 0, 1
DIAG_LOWER_MIRROR, DIAG_FULL = 0, 1
PROGRAM_SWMPI, PROGRAM_SWMS = 0, 1
(SCHED_PERSISTENT, SCHED_PLAIN, SCHED_PERSISTENT_I32, SCHED_PERSISTENT_I16, SCHED_PERSISTENT_LPP16,
 SCHED_PERSISTENT_LPS16) = 0, 1, 2, 3, 4, 5
(OPT_STREAM_CHUNK_BYTES, OPT_KERNEL_TIMING, OPT_SCORE_STRIP_ROWS, OPT_TUNE_MIN_CELLS, OPT_REVERSE_STRIP_ROWS,
 OPT_REVERSE_ROW_CAP, OPT_LPS_LANES, OPT_SCORE_MAX3) = 1, 2, 3, 4, 5, 6, 7, 8
KERNEL_NAMES = {1: "one_pass", 2: "score16", 3: "sort", 4: "rerun_track", 5: "rerun_capture", 6: "lpp16",
                7: "global", 8: "reverse", 9: "rect64", 10: "distance", 11: "lps16"}
MATRIX_SIZE = 127

ABI_SYMBOLS = (
    "swg_create", "swg_set_sequences", "swg_compute_block", "swg_compute_block_device",
    "swg_synchronize", "swg_set_schedule", "swg_last_kernel_stats", "swg_last_error", "swg_destroy",
    "swg_serve_create", "swg_serve_submit", "swg_serve_wait", "swg_serve_times", "swg_serve_release",
    "swg_serve_stats", "swg_serve_destroy", "swg_set_option", "swg_last_kernel_times",
    "swg_compute_block_device_noalloc", "swg_last_block_config",
)
NOALLOC_SCRATCH_BYTES = 64 << 20  # SWG_NOALLOC_SCRATCH_BYTES
HOST_SYMBOLS = (
    "salsa_range_partition", "salsa_block_is_computed", "salsa_compute_rank", "salsa_compute_full_matrix",
    "salsa_write_matrix", "salsa_write_swms_block", "salsa_load_scoring_matrix",
    "salsa_synth_dna", "salsa_synth_protein", "salsa_swms_compute_rank", "salsa_read_restart_file",
    "salsa_read_matrix_rows", "salsa_read_partial_stripes", "salsa_distance_value", "salsa_write_index_file",
)


class SwgError(RuntimeError):
    """Error returned by libswg (code + swg_last_error message)."""

    def __init__(self, code, msg):
        super().__init__("swg error %d: %s" % (code, msg))
        self.code = code


class SmithWatermanGotohException(SwgError):
    """Raised where the reference raises SmithWatermanGotohException
    (Salsa.CoreTPL/Bio/Algorithms/SmithWaterman/SmithWatermanGotohException.cs:30-47)."""


class JukesCantorUndefined(SwgError):
    """Raised where the reference throws Exception("Jukes and Cantor Distance
    Undefined - Log(Zero)") (Salsa.CoreTPL/Bio/Algorithms/Alignment.cs:340-345)."""


class swg_kernel_time(ctypes.Structure):
    _fields_ = [("kernel", ctypes.c_int32), ("reserved", ctypes.c_int32), ("ms", ctypes.c_double),
                ("cells", ctypes.c_double), ("cells_run", ctypes.c_double)]


class swg_block_config(ctypes.Structure):
    _fields_ = [("variant", ctypes.c_int32), ("strip_rows", ctypes.c_int32), ("rerun_strip_rows", ctypes.c_int32),
                ("pairs_per_wave", ctypes.c_int32), ("tuning", ctypes.c_int32), ("trials", ctypes.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class swg_params(ctypes.Structure):
    _fields_ = [
        ("mode", ctypes.c_int32), ("alignment_type", ctypes.c_int32), ("distance_fn", ctypes.c_int32),
        ("tdistance", ctypes.c_int32), ("diag_mode", ctypes.c_int32), ("schedule", ctypes.c_int32),
        ("gap_open", ctypes.c_float), ("gap_ext", ctypes.c_float),
        ("scores", ctypes.POINTER(ctypes.c_float)), ("device", ctypes.c_int32), ("program", ctypes.c_int32),
    ]


_libs = {}


def _load(name):
    if name not in _libs:
        path = os.path.join(HERE, name)
        if not os.path.exists(path):
            raise RuntimeError("%s is not built (run __graft_entry__.build())" % path)
        _libs[name] = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
    return _libs[name]


def lib():
    """The HIP library (C ABI of include/swg.h)."""
    L = _load("libswg.so")
    if not getattr(L, "_typed", False):
        vp, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
        L.swg_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(swg_params)]
        L.swg_set_sequences.argtypes = [vp, vp, vp, i32]
        L.swg_compute_block.argtypes = [vp, i32, i32, i32, i32, i32, vp, i64, vp, ctypes.POINTER(i64)]
        L.swg_compute_block_device.argtypes = [vp, i32, i32, i32, i32, i32, vp, i64, vp, ctypes.POINTER(i64)]
        L.swg_compute_block_device_noalloc.argtypes = [vp, i32, i32, i32, i32, i32, vp, i64, vp,
                                                       ctypes.POINTER(i64)]
        L.swg_synchronize.argtypes = [vp]
        L.swg_set_schedule.argtypes = [vp, i32, i32]
        L.swg_last_kernel_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                            ctypes.POINTER(i32), ctypes.POINTER(i32)]
        L.swg_last_kernel_times.argtypes = [vp, ctypes.POINTER(swg_kernel_time), i32, ctypes.POINTER(i32)]
        L.swg_last_block_config.argtypes = [vp, ctypes.POINTER(swg_block_config)]
        L.swg_last_error.restype = ctypes.c_char_p
        L.swg_destroy.argtypes = [vp]
        L.swg_destroy.restype = None
        L.swg_serve_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(swg_params), vp, vp, i32, i32, i32, i32,
                                       i64, i32]
        L.swg_serve_submit.argtypes = [vp, vp, i32, vp, vp, ctypes.POINTER(i64)]
        L.swg_serve_wait.argtypes = [vp, i64, i32, ctypes.POINTER(i32)]
        L.swg_serve_times.argtypes = [vp, i64, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(ctypes.c_double)]
        L.swg_serve_release.argtypes = [vp, i64]
        L.swg_serve_stats.argtypes = [vp, ctypes.POINTER(i64), ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(ctypes.c_double)]
        L.swg_set_option.argtypes = [vp, i32, i64]
        L.swg_serve_destroy.argtypes = [vp]
        L.swg_serve_destroy.restype = None
        L._typed = True
    return L


def host():
    """The C++ host-mirror library (include/salsa_host.h)."""
    lib()
    H = _load("libsalsa_host.so")
    if not getattr(H, "_typed", False):
        vp, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
        H.salsa_range_partition.argtypes = [i32, i32, i32, vp, vp]
        H.salsa_block_is_computed.argtypes = [i32, i32]
        H.salsa_compute_rank.argtypes = [vp, i32, i32, i32, vp, i32, ctypes.POINTER(i64)]
        H.salsa_compute_full_matrix.argtypes = [vp, i32, i32, vp, i32, ctypes.POINTER(i64)]
        H.salsa_write_matrix.argtypes = [ctypes.c_char_p, vp, i64]
        H.salsa_write_index_file.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p), vp, i32, i32]
        H.salsa_write_swms_block.argtypes = [ctypes.c_char_p, ctypes.c_char_p, i32, ctypes.c_char_p, ctypes.c_char_p,
                                             i32, i32, i32, i32, i32, i32, ctypes.POINTER(ctypes.c_char_p), vp, i64, i32]
        H.salsa_load_scoring_matrix.argtypes = [ctypes.c_char_p, vp]
        H.salsa_synth_dna.argtypes = [ctypes.c_uint64, i32, i32, ctypes.c_double, ctypes.c_double, i32, vp, vp, i64, vp]
        H.salsa_synth_dna.restype = i64
        H.salsa_synth_protein.argtypes = [ctypes.c_uint64, i32, i32, vp, vp, i64, vp]
        H.salsa_synth_protein.restype = i64
        H.salsa_swms_compute_rank.argtypes = [vp, i32, i32, i32, i32, i32, ctypes.c_char_p, ctypes.c_char_p,
                                              ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
                                              ctypes.POINTER(ctypes.c_char_p), i32, i32, ctypes.POINTER(i32),
                                              ctypes.POINTER(i64)]
        H.salsa_read_restart_file.argtypes = [ctypes.c_char_p, i32, ctypes.POINTER(i32), ctypes.POINTER(i32),
                                              ctypes.POINTER(i32)]
        H.salsa_read_matrix_rows.argtypes = [ctypes.c_char_p, i32, i64, i32, i32, vp]
        H.salsa_read_partial_stripes.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, i32, i32, vp, vp,
                                                 i32, i32, vp]
        H.salsa_distance_value.argtypes = [i32, i32, ctypes.c_double, ctypes.c_double, i32]
        H.salsa_distance_value.restype = ctypes.c_double
        H._typed = True
    return H


def _check(rc, exc=SwgError):
    if rc != SWG_OK:
        msg = lib().swg_last_error()
        if rc == SWG_E_JC_UNDEFINED and exc is SwgError:
            exc = JukesCantorUndefined
        raise exc(rc, msg.decode() if msg else "")


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def _out_ld(out, dtype, rows, cols):
    """Row pitch (elements) of a caller-supplied 2-D output array: it must
    hold TDistance values with unit column stride and at least rows x cols
    elements (a row or column slice of a larger matrix is fine)."""
    if not isinstance(out, np.ndarray) or out.ndim != 2:
        raise ValueError("out must be a 2-D numpy array")
    if out.dtype != np.dtype(dtype):
        raise ValueError("out has dtype %s, the context's TDistance is %s" % (out.dtype, np.dtype(dtype)))
    if out.shape[0] < rows or out.shape[1] < cols:
        raise ValueError("out of shape %s is smaller than the block (%d x %d)" % (out.shape, rows, cols))
    if rows > 0 and cols > 0:
        if out.strides[1] != out.itemsize or out.strides[0] % out.itemsize or out.strides[0] < cols * out.itemsize:
            raise ValueError("out must have unit column stride and a row pitch of at least the block width")
        if not out.flags.writeable:
            raise ValueError("out is read-only")
    return out.strides[0] // out.itemsize if out.shape[0] > 0 else cols


# ---------------------------------------------------------------- scoring
def load_scoring_matrix(name_or_path):
    """ScoringMatrix.Load (Salsa.CoreTPL/Bio/Algorithms/ScoringMatrix.cs:145-196):
    127x127 float32 table, row = s1 symbol, unlisted symbols score 0."""
    path = name_or_path if os.path.sep in name_or_path else os.path.join(MATRIX_DIR, name_or_path)
    out = np.zeros((MATRIX_SIZE, MATRIX_SIZE), dtype=np.float32)
    rc = host().salsa_load_scoring_matrix(path.encode(), _ptr(out))
    if rc != SWG_OK:
        raise SwgError(rc, "cannot load scoring matrix %s" % path)
    return out


# ---------------------------------------------------------------- sequences
class SequenceSet:
    """Concatenated upper-case residues + offsets (the reference's
    IList<Sequence> _sequences, Salsa.SmithWatermanMPI/Program.cs:65)."""

    def __init__(self, residues, offsets, labels=None):
        self.residues = np.ascontiguousarray(residues, dtype=np.uint8)
        self.offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        self.labels = labels

    @classmethod
    def from_strings(cls, seqs, labels=None):
        bs = [s.encode("ascii") if isinstance(s, str) else bytes(s) for s in seqs]
        offs = np.zeros(len(bs) + 1, dtype=np.int64)
        offs[1:] = np.cumsum([len(b) for b in bs])
        res = np.frombuffer(b"".join(bs), dtype=np.uint8) if bs and offs[-1] else np.zeros(0, np.uint8)
        return cls(res.copy(), offs, labels)

    def __len__(self):
        return len(self.offsets) - 1

    @property
    def lengths(self):
        return np.diff(self.offsets).astype(np.int32)

    def seq(self, i):
        return self.residues[self.offsets[i]:self.offsets[i + 1]].tobytes()

    def subset(self, idx):
        return SequenceSet.from_strings([self.seq(int(i)) for i in idx])


def parse_fasta(text):
    """FastaSequenceParser semantics (Salsa.CoreTPL/Bio/IO/Parser.cs:133-151):
    label = text after '>' trimmed; residue lines trimmed, upper-cased and
    concatenated."""
    labels, seqs = [], []
    cur = None
    for raw in text.replace("\r\n", "\n").replace("\r", "\n").split("\n"):
        line = raw.strip()
        if raw.lstrip().startswith(">"):
            labels.append(raw.lstrip()[1:].strip())
            seqs.append([])
            cur = seqs[-1]
        elif line and cur is not None:
            cur.append(line.upper())
    return SequenceSet.from_strings(["".join(s) for s in seqs], labels)


def synth_dna(seed, n, n_anc, sub_rate, indel_rate, len_kind, len_params):
    H = host()
    p = np.ascontiguousarray(len_params, dtype=np.float64)
    offs = np.zeros(n + 1, dtype=np.int64)
    need = -H.salsa_synth_dna(seed, n, n_anc, sub_rate, indel_rate, len_kind, _ptr(p), None, 0, _ptr(offs))
    res = np.zeros(max(need, 1), dtype=np.uint8)
    got = H.salsa_synth_dna(seed, n, n_anc, sub_rate, indel_rate, len_kind, _ptr(p), _ptr(res), need, _ptr(offs))
    assert got == need
    return SequenceSet(res[:need], offs)


def synth_protein(seed, n, len_kind, len_params):
    H = host()
    p = np.ascontiguousarray(len_params, dtype=np.float64)
    offs = np.zeros(n + 1, dtype=np.int64)
    need = -H.salsa_synth_protein(seed, n, len_kind, _ptr(p), None, 0, _ptr(offs))
    res = np.zeros(max(need, 1), dtype=np.uint8)
    got = H.salsa_synth_protein(seed, n, len_kind, _ptr(p), _ptr(res), need, _ptr(offs))
    assert got == need
    return SequenceSet(res[:need], offs)


# ---------------------------------------------------------------- partitioning
def range_partition(length, nparts, start=0):
    """RangePartitioner.Partition (Salsa.CoreTPL/RangePartitioner.cs:17-55): list of inclusive (start, end)."""
    s = np.zeros(nparts, dtype=np.int32)
    e = np.zeros(nparts, dtype=np.int32)
    rc = host().salsa_range_partition(start, length, nparts, _ptr(s), _ptr(e))
    if rc != SWG_OK:
        raise SwgError(rc, "bad partition request")
    return list(zip(s.tolist(), e.tolist()))


def block_is_computed(rank, j):
    """Checkerboard ownership (Salsa.SmithWatermanMPI/Program.cs:628-639)."""
    return bool(host().salsa_block_is_computed(rank, j))


# ---------------------------------------------------------------- aligner context
class Aligner:
    """One libswg context: the MI355X replacement of the reference's
    SmithWatermanGotoh + ComputeDistance per block (ComputeBlock,
    Salsa.SmithWatermanMPI/Program.cs:203-257)."""

    def __init__(self, scores, gap_open, gap_ext, alignment_type=ALIGN_NUCLEIC,
                 distance_fn=DIST_PERCENT_IDENTITY, tdistance=TDIST_INT16, diag_mode=DIAG_LOWER_MIRROR,
                 schedule=SCHED_PERSISTENT, mode=MODE_SW_LOCAL, device=0, program=PROGRAM_SWMPI):
        self._scores = np.ascontiguousarray(scores, dtype=np.float32).reshape(MATRIX_SIZE, MATRIX_SIZE)
        p = swg_params()
        p.mode, p.alignment_type, p.distance_fn = mode, alignment_type, distance_fn
        p.program = program
        p.tdistance, p.diag_mode, p.schedule = tdistance, diag_mode, schedule
        p.gap_open, p.gap_ext = float(gap_open), float(gap_ext)
        p.scores = self._scores.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        p.device = device
        self.tdistance = tdistance
        self.dtype = np.float64 if tdistance == TDIST_DOUBLE else np.int16
        h = ctypes.c_void_p()
        _check(lib().swg_create(ctypes.byref(h), ctypes.byref(p)))
        self._h = h
        self.seqs = None

    def close(self):
        if getattr(self, "_h", None):
            lib().swg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def set_sequences(self, seqs):
        self.seqs = seqs
        _check(lib().swg_set_sequences(self._h, _ptr(seqs.residues), _ptr(seqs.offsets), len(seqs)))

    def set_schedule(self, schedule, strip_rows=0):
        _check(lib().swg_set_schedule(self._h, schedule, strip_rows))

    def set_stream_chunk_bytes(self, nbytes):
        _check(lib().swg_set_option(self._h, OPT_STREAM_CHUNK_BYTES, int(nbytes)))

    def set_score_strip(self, rows):
        """Rows per strip of the int16 score pass (16, 24, 32; 0 = autotuned)."""
        _check(lib().swg_set_option(self._h, OPT_SCORE_STRIP_ROWS, int(rows)))

    def set_reverse_strip(self, rows):
        """Rows per strip of the reverse score pass (8 or 16; 0 = default)."""
        _check(lib().swg_set_option(self._h, OPT_REVERSE_STRIP_ROWS, int(rows)))

    def set_lps_lanes(self, lanes):
        """Lanes per pair of the lanes-per-pair score pass: 0 (by block), 16 or 32."""
        _check(lib().swg_set_option(self._h, OPT_LPS_LANES, int(lanes)))

    def set_score_max3(self, on):
        """The score pass's max3 form (values biased by gap_open, three-way
        maxima in one packed instruction) where the set allows it: True
        (default) or False; the results are identical."""
        _check(lib().swg_set_option(self._h, OPT_SCORE_MAX3, 1 if on else 0))

    def set_reverse_row_cap(self, rows):
        """Reverse rows per pair at most (a multiple of 16): pairs whose band
        reaches past it are re-run over their whole prefix."""
        _check(lib().swg_set_option(self._h, OPT_REVERSE_ROW_CAP, int(rows)))

    def set_tune_min_cells(self, cells):
        """DP cells a block needs for the autotuner to time it."""
        _check(lib().swg_set_option(self._h, OPT_TUNE_MIN_CELLS, int(cells)))

    def last_block_config(self):
        """Launch configuration of the last block: variant (SWG_SCHED_*), strip
        rows, re-run strip rows, pairs per wave, tuning state (0 fixed, 1
        autotuner trial, 2 autotuner choice, 3 the fastest configuration
        timed so far: a diagonal block while its key's search goes on) and
        the key's timed trials."""
        cfg = swg_block_config()
        _check(lib().swg_last_block_config(self._h, ctypes.byref(cfg)))
        return cfg.as_dict()

    def set_kernel_timing(self, on):
        """A wall-clock stamp kernel before the block's kernels and after each,
        inside the block's graph (see kernel_times)."""
        _check(lib().swg_set_option(self._h, OPT_KERNEL_TIMING, 1 if on else 0))

    def kernel_times(self):
        """Per-kernel [(name, ms, cells, cells_run)] of the last block run with
        kernel timing on."""
        buf = (swg_kernel_time * 16)()
        n = ctypes.c_int32()
        _check(lib().swg_last_kernel_times(self._h, buf, 16, ctypes.byref(n)))
        return [(KERNEL_NAMES.get(buf[k].kernel, str(buf[k].kernel)), buf[k].ms, buf[k].cells, buf[k].cells_run)
                for k in range(n.value)]

    def compute_block(self, row_start, row_end, col_start, col_end, is_diagonal, stats=False, out=None):
        """ComputeBlock: returns (distances[rows, cols], stats or None, pairs)."""
        rows, cols = row_end - row_start + 1, col_end - col_start + 1
        if out is None:
            out = np.zeros((max(rows, 0), max(cols, 0)), dtype=self.dtype)
        ld = _out_ld(out, self.dtype, rows, cols)
        st = np.zeros((max(rows, 0), max(cols, 0), 4), dtype=np.int32) if stats else None
        done = ctypes.c_int64(0)
        _check(lib().swg_compute_block(self._h, row_start, row_end, col_start, col_end, int(bool(is_diagonal)),
                                       _ptr(out), ld, _ptr(st) if stats else None, ctypes.byref(done)))
        return out, st, done.value

    def compute_block_device(self, row_start, row_end, col_start, col_end, is_diagonal, dev_out_ptr, ld,
                             dev_stats_ptr=None):
        done = ctypes.c_int64(0)
        _check(lib().swg_compute_block_device(self._h, row_start, row_end, col_start, col_end,
                                              int(bool(is_diagonal)), ctypes.c_void_p(dev_out_ptr), ld,
                                              ctypes.c_void_p(dev_stats_ptr) if dev_stats_ptr else None,
                                              ctypes.byref(done)))
        return done.value

    def compute_block_device_noalloc(self, row_start, row_end, col_start, col_end, is_diagonal, dev_out_ptr, ld,
                                     dev_stats_ptr=None):
        """swg_compute_block_device_noalloc: as compute_block_device, with
        every device buffer beyond the output taken from the library's static
        device scratch (no allocation)."""
        done = ctypes.c_int64(0)
        _check(lib().swg_compute_block_device_noalloc(self._h, row_start, row_end, col_start, col_end,
                                                      int(bool(is_diagonal)), ctypes.c_void_p(dev_out_ptr), ld,
                                                      ctypes.c_void_p(dev_stats_ptr) if dev_stats_ptr else None,
                                                      ctypes.byref(done)))
        return done.value

    def synchronize(self):
        _check(lib().swg_synchronize(self._h))

    def last_kernel_stats(self):
        """(kernel ms, DP cells, strip rows, variant SWG_SCHED_*) of the last block."""
        ms, cells, strip, var = ctypes.c_double(), ctypes.c_double(), ctypes.c_int32(), ctypes.c_int32()
        _check(lib().swg_last_kernel_stats(self._h, ctypes.byref(ms), ctypes.byref(cells), ctypes.byref(strip),
                                           ctypes.byref(var)))
        return ms.value, cells.value, strip.value, var.value

    def compute_full_matrix(self, world=1):
        """SWMPI end to end on one process (every rank's Compute + Scatter):
        the n x n matrix WriteFullMatrixOnRank0 writes."""
        n = len(self.seqs)
        full = np.zeros((n, n), dtype=self.dtype)
        pairs = ctypes.c_int64(0)
        rc = host().salsa_compute_full_matrix(self._h, n, world, _ptr(full), full.itemsize, ctypes.byref(pairs))
        _check(rc)
        return full, pairs.value

    def swms_compute_rank(self, world, rank, block_dir, done_dir=None, last_saved_block=-1,
                          block_write_frequency=1, project="salsa", sysname="mi355x", fasta="input.fasta",
                          labels=None, stop_after=-1):
        """SWMS Compute(lastSavedBlock) + WriteBlocks + done file for one rank;
        returns (return code, last saved block, pairs)."""
        n = len(self.seqs)
        labels = labels or ["seq%d" % i for i in range(n)]
        arr = (ctypes.c_char_p * n)(*[x.encode() for x in labels])
        saved, pairs = ctypes.c_int32(), ctypes.c_int64()
        rc = host().salsa_swms_compute_rank(self._h, n, world, rank, last_saved_block, block_write_frequency,
                                            block_dir.encode(), done_dir.encode() if done_dir else None,
                                            project.encode(), sysname.encode(), fasta.encode(), arr,
                                            np.dtype(self.dtype).itemsize, stop_after, ctypes.byref(saved),
                                            ctypes.byref(pairs))
        if rc not in (SWG_OK, SWG_E_CAPACITY):
            _check(rc)
        return rc, saved.value, pairs.value

    def compute_rank(self, world, rank):
        n = len(self.seqs)
        (rs, re) = range_partition(n, world)[rank]
        part = np.zeros((re - rs + 1, n), dtype=self.dtype)
        pairs = ctypes.c_int64(0)
        _check(host().salsa_compute_rank(self._h, n, world, rank, _ptr(part), part.itemsize, ctypes.byref(pairs)))
        return part, pairs.value


def write_swms_block(path, project, rank, sysname, fasta, row_block, col_block, row_range, col_range, names, data):
    arr = (ctypes.c_char_p * 4)(*[n.encode() for n in names])
    data = np.ascontiguousarray(data)
    rc = host().salsa_write_swms_block(path.encode(), project.encode(), rank, sysname.encode(), fasta.encode(),
                                       row_block, col_block, row_range[0], row_range[1], col_range[0], col_range[1],
                                       arr, _ptr(data), data.size, data.itemsize)
    if rc != SWG_OK:
        raise SwgError(rc, "cannot write %s" % path)


# ---------------------------------------------------------------- serving
class Server:
    """Continuous-batching query server over a device-resident panel
    (swg_serve_* in include/swg.h).  submit() returns a ticket; the result
    arrays are filled when the query completes."""

    def __init__(self, panel, scores, gap_open, gap_ext, alignment_type=ALIGN_NUCLEIC,
                 distance_fn=DIST_PERCENT_IDENTITY, tdistance=TDIST_INT16, max_query_len=2000, max_batch=16,
                 lanes=2, hbm_budget=1 << 30, stats=True, device=0, schedule=SCHED_PERSISTENT):
        self._scores = np.ascontiguousarray(scores, dtype=np.float32).reshape(MATRIX_SIZE, MATRIX_SIZE)
        p = swg_params()
        p.mode, p.alignment_type, p.distance_fn = MODE_SW_LOCAL, alignment_type, distance_fn
        p.tdistance, p.diag_mode, p.schedule = tdistance, DIAG_FULL, schedule
        p.gap_open, p.gap_ext = float(gap_open), float(gap_ext)
        p.scores = self._scores.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        p.device = device
        self.npanel = len(panel)
        self.dtype = np.float64 if tdistance == TDIST_DOUBLE else np.int16
        self.want_stats = stats
        self._panel = panel
        h = ctypes.c_void_p()
        _check(lib().swg_serve_create(ctypes.byref(h), ctypes.byref(p), _ptr(panel.residues), _ptr(panel.offsets),
                                      self.npanel, max_query_len, max_batch, lanes, int(hbm_budget), int(stats)))
        self._h = h
        self._bufs = {}

    def submit(self, query):
        q = np.frombuffer(bytes(query), dtype=np.uint8) if len(query) else np.zeros(1, np.uint8)
        out = np.zeros(self.npanel, dtype=self.dtype)
        st = np.zeros((self.npanel, 4), dtype=np.int32) if self.want_stats else None
        t = ctypes.c_int64()
        _check(lib().swg_serve_submit(self._h, _ptr(q), len(query), _ptr(out), _ptr(st), ctypes.byref(t)))
        self._bufs[t.value] = (out, st, q)
        return t.value

    def wait(self, ticket, timeout_ms=-1):
        done = ctypes.c_int32()
        _check(lib().swg_serve_wait(self._h, ticket, timeout_ms, ctypes.byref(done)))
        return bool(done.value)

    def result(self, ticket):
        out, st, _ = self._bufs[ticket]
        return out, st

    def times(self, ticket):
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        _check(lib().swg_serve_times(self._h, ticket, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def release(self, ticket):
        _check(lib().swg_serve_release(self._h, ticket))
        self._bufs.pop(ticket, None)

    def stats(self):
        n, cells, now = ctypes.c_int64(), ctypes.c_double(), ctypes.c_double()
        _check(lib().swg_serve_stats(self._h, ctypes.byref(n), ctypes.byref(cells), ctypes.byref(now)))
        return n.value, cells.value, now.value

    def close(self):
        if getattr(self, "_h", None):
            lib().swg_serve_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def write_matrix(path, data):
    """A row-major TDistance matrix file (WriteFullMatrixOnRank0, or a rank's
    partial stripe as PartialMatrixBinaryWriter writes it)."""
    data = np.ascontiguousarray(data)
    if host().salsa_write_matrix(path.encode(), _ptr(data), data.nbytes) != SWG_OK:
        raise SwgError(SWG_E_ARGUMENT, "cannot write %s" % path)


def write_index_file(path, seqs, labels=None, crlf=True):
    """WriteIndexFile (Salsa.SmithWatermanMPI/Program.cs:316-331): one line
    "<index>\t<label>\t<length>" per sequence."""
    n = len(seqs)
    labels = labels if labels is not None else (seqs.labels or [""] * n)
    arr = (ctypes.c_char_p * max(n, 1))(*[x.encode("utf-8") for x in labels])
    if host().salsa_write_index_file(path.encode(), arr, _ptr(seqs.offsets), n, int(bool(crlf))) != SWG_OK:
        raise SwgError(SWG_E_ARGUMENT, "cannot write %s" % path)


def read_matrix_rows(path, n_cols, start_row, row_count, dtype=np.int16):
    """MatrixBinaryReader.ReadInt16/ReadDouble(fileName, startRowIndex, rowCount)
    (Salsa.CoreTPL/Blas/MatrixBinaryReader.cs:224-311), as PWC and MDS load
    their distances."""
    out = np.zeros((row_count, n_cols), dtype=dtype)
    rc = host().salsa_read_matrix_rows(path.encode(), n_cols, start_row, row_count, out.itemsize, _ptr(out))
    if rc != SWG_OK:
        raise SwgError(rc, "cannot read %s" % path)
    return out


def read_partial_stripes(directory, prefix, ext, start_index, end_index, ranges, n_cols, dtype=np.int16):
    """MatrixBinaryReader.ReadInt16(path, prefix, ext, startIndex, endIndex, ranges)
    (MatrixBinaryReader.cs:191-217): the rank stripes "<prefix>_<i><ext>"."""
    starts = np.array([r[0] for r in ranges], dtype=np.int32)
    ends = np.array([r[1] for r in ranges], dtype=np.int32)
    rows = int(sum(ends[i] - starts[i] + 1 for i in range(start_index, end_index + 1)))
    out = np.zeros((rows, n_cols), dtype=dtype)
    rc = host().salsa_read_partial_stripes(directory.encode(), prefix.encode(), ext.encode(), start_index, end_index,
                                           _ptr(starts), _ptr(ends), n_cols, out.itemsize, _ptr(out))
    if rc != SWG_OK:
        raise SwgError(rc, "cannot read stripes %s_*%s" % (prefix, ext))
    return out


PWC, MDS = 0, 1


def distance_value(value, consumer=PWC, distance_cut=-1.0, undefined_value=-1.0, processing_option=0):
    """The distance a consumer derives from a stored Int16 value:
    PWCParallelism.getDistanceValue (PWCParallelism.cs:211-218) or
    SALSAParallelism.getDistanceValue (MDSasCHISQ/SALSAParallelism.cs:318-370)."""
    return host().salsa_distance_value(int(value), consumer, distance_cut, undefined_value, processing_option)


def read_restart_file(path, process_rank):
    """ReadRestartFile (Salsa.SmithWatermanMS/Program.cs:245-262): (world, rank, lastSavedBlock)."""
    w, r, ls = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    rc = host().salsa_read_restart_file(path.encode(), process_rank, ctypes.byref(w), ctypes.byref(r), ctypes.byref(ls))
    if rc != SWG_OK:
        raise SwgError(rc, "cannot read restart file %s" % path)
    return w.value, r.value, ls.value
