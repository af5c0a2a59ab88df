"""The C ABI library loads and exports every entry point declared in
include/*.h; argument validation that happens before any device call."""
import ctypes
import os
import re

import numpy as np
import pytest

import csharp_amd as ca

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:swg|salsa)_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    L = ca.lib()
    H = ca.host()
    for name in declared("swg.h"):
        assert hasattr(L, name), name
    for name in declared("salsa_host.h"):
        assert hasattr(H, name), name
    assert set(declared("swg.h")) == set(ca.ABI_SYMBOLS)
    assert set(declared("salsa_host.h")) == set(ca.HOST_SYMBOLS)


def test_nonintegral_scores_rejected(ednafull):
    M = ednafull.copy()
    M[ord("A"), ord("A")] = 5.5
    with pytest.raises(ca.SwgError) as ei:
        ca.Aligner(M, 16, 4)
    assert ei.value.code == ca.SWG_E_NONINTEGRAL
    with pytest.raises(ca.SwgError) as ei:
        ca.Aligner(ednafull, 16.5, 4)
    assert ei.value.code == ca.SWG_E_NONINTEGRAL


def test_unsupported_distance_rejected(ednafull):
    # checked before any device call: MinMaxNormScore exists in the SWMS and
    # global programs only; unknown distance functions and programs
    with pytest.raises(ca.SwgError) as ei:
        ca.Aligner(ednafull, 16, 4, distance_fn=ca.DIST_MINMAX_NORM)
    assert ei.value.code == ca.SWG_E_UNSUPPORTED
    with pytest.raises(ca.SwgError) as ei:
        ca.Aligner(ednafull, 16, 4, distance_fn=7)
    assert ei.value.code == ca.SWG_E_ARGUMENT
    with pytest.raises(ca.SwgError) as ei:
        ca.Aligner(ednafull, 16, 4, program=5)
    assert ei.value.code == ca.SWG_E_ARGUMENT


def test_null_arguments():
    L = ca.lib()
    assert L.swg_create(None, None) == ca.SWG_E_ARGUMENT
    assert L.swg_last_error()
    assert L.swg_synchronize(None) == ca.SWG_E_ARGUMENT
    L.swg_destroy(None)


def test_negative_gap_penalties_rejected(ednafull):
    # checked before any device call: penalties are costs the reference
    # subtracts (SmithWatermanGotoh.cs:159-187); the kernels need them >= 0
    for o, e in ((-1, 4), (16, -1), (-16, -4)):
        with pytest.raises(ca.SwgError) as ei:
            ca.Aligner(ednafull, o, e)
        assert ei.value.code == ca.SWG_E_UNSUPPORTED
        with pytest.raises(ca.SwgError) as ei:
            ca.Aligner(ednafull, o, e, mode=ca.MODE_NW_GLOBAL)
        assert ei.value.code == ca.SWG_E_UNSUPPORTED
