"""The HIP path (through the C ABI) against the committed golden fixtures
(tests/golden/pairs_v1.npz): every pair (i, j) of each fixture set, s1 = s_i,
for every distance configuration and schedule.  Int16 distances and the
statistics bit-exact; Double distances bit-exact where no logarithm is
involved, within 4 ulp for Kimura2 / Jukes-Cantor (device log vs host log)."""
import numpy as np
import pytest

import csharp_amd as ca
from tests.golden import make_golden as MG
from tests.test_golden import load

pytestmark = pytest.mark.gpu

SCHEDULES = [ca.SCHED_PERSISTENT, ca.SCHED_PLAIN, ca.SCHED_PERSISTENT_I32, ca.SCHED_PERSISTENT_I16,
             ca.SCHED_PERSISTENT_LPP16]


@pytest.mark.parametrize("name", [c[0] for c in MG.CONFIGS])
def test_gpu_matches_golden(name):
    z, _, sets = load()
    o, e, glob, program, nucleic, dist_fn, double = z[name + "_config"].tolist()
    seqs = sets[bytes(z[name + "_alphabet"]).decode()]
    M = ca.load_scoring_matrix(bytes(z[name + "_matrix"]).decode())
    want, wst = z[name + "_dist"], z[name + "_stats"]
    n = len(seqs)
    scheds = [ca.SCHED_PERSISTENT] if glob else SCHEDULES
    for sched in scheds:
        print("golden %s schedule=%d" % (name, sched), flush=True)
        al = ca.Aligner(M, o, e, alignment_type=ca.ALIGN_NUCLEIC if nucleic else ca.ALIGN_PROTEIN,
                        distance_fn=dist_fn, tdistance=ca.TDIST_DOUBLE if double else ca.TDIST_INT16,
                        mode=ca.MODE_NW_GLOBAL if glob else ca.MODE_SW_LOCAL, program=program,
                        diag_mode=ca.DIAG_FULL, schedule=sched)
        al.set_sequences(seqs)
        got, gst, pairs = al.compute_block(0, n - 1, 0, n - 1, False, stats=True)
        al.close()
        assert pairs == n * n
        # the oracle's statistics are (score, columns, identity, similarity);
        # the library's (score, columns, count, 0) with count = identity
        # (nucleic) or similarity (protein), the global kernel's fourth the
        # aligned-region length, and Kimura2 / Jukes-Cantor's the counts
        # behind the distance in both
        if glob or (nucleic and dist_fn in (ca.DIST_PERCENT_IDENTITY, ca.DIST_MINMAX_NORM)):
            assert np.array_equal(gst[..., :3], wst[..., :3]), (name, sched)
        elif not nucleic:
            assert np.array_equal(gst[..., :2], wst[..., :2]), (name, sched)
            assert np.array_equal(gst[..., 2], wst[..., 3]), (name, sched)
        else:
            assert np.array_equal(gst, wst), (name, sched)
        if not double:
            assert np.array_equal(got, want), (name, sched)
        elif dist_fn in (ca.DIST_KIMURA2, ca.DIST_JUKES_CANTOR):
            nan = np.isnan(want)
            assert np.array_equal(np.isnan(got), nan)
            assert np.all(np.abs(got[~nan] - want[~nan]) <= np.abs(want[~nan]) * 2.0 ** -50), (name, sched)
        else:
            same = (got.view(np.int64) == want.view(np.int64)) | (np.isnan(got) & np.isnan(want))
            assert same.all(), (name, sched)
