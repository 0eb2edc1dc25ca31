"""The CPU oracle against the reference's own outputs.

Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py from the reference
headers compiled in plain mode) hold seeded inputs and the reference's outputs for every hot-path
function, plus whole driver sequences.  The oracle must reproduce them value-exactly (+0 == -0);
it is compiled without FMA contraction like the reference's plain x86-64 build and calls the same
glibc pow().  Where the reference harness is built (oracle/_ref), more seeds are checked live.
"""
import glob
import os

import numpy as np
import pytest

import refcases
import sequences
from _oracle import ref

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _cases():
    out = []
    for case in refcases.CASES_2D + refcases.CASES_3D:
        for op in refcases.ops_for(case):
            out.append((case, op))
    return out


def _exact(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, what
    bad = ~((a == b) | (np.isnan(a) & np.isnan(b)))
    assert not bad.any(), "%s: %d values differ, max rel %.3g" % (
        what, int(bad.sum()), float(np.max(np.abs(a - b)[bad] / (np.abs(a[bad]) + 1e-300))))


@pytest.mark.parametrize("case,op", _cases(), ids=lambda x: getattr(x, "name", x))
def test_oracle_matches_golden(case, op):
    blob = np.load(os.path.join(GOLDEN, case.name + ".npz"))
    inp = {k[3:]: blob[k] for k in blob.files if k.startswith("in/")}
    pre = "out/%s/" % op
    want = {k[len(pre):]: blob[k] for k in blob.files if k.startswith(pre)}
    assert want, "no golden outputs for " + op
    got = refcases.run_oracle(case, op, inp)
    for k, v in want.items():
        _exact(got[k], v, "%s %s [%s]" % (case.name, op, refcases.OPS_2D.get(op, "")))


@pytest.mark.parametrize("seq", sequences.SEQUENCES, ids=lambda s: s.name)
def test_oracle_sequence_matches_golden(seq):
    blob = np.load(os.path.join(GOLDEN, "seq_" + seq.name + ".npz"))
    got = seq.run(sequences.OracleBackend())
    for k in blob.files:
        _exact(got[k], blob[k], seq.name + " " + k)


@pytest.mark.skipif(ref() is None, reason="reference harness not built here (oracle/_ref)")
@pytest.mark.parametrize("seed", [101, 202, 303])
def test_oracle_matches_live_reference(seed):
    for base in refcases.CASES_2D + refcases.CASES_3D:
        case = refcases.Case(base.name, base.kind, (base.lx + seed % 5, base.ly + seed % 3, base.lz), seed, base.cite)
        if case.lz == 1:
            case.lz = 1
        inp = refcases.make_inputs(case)
        for op in refcases.ops_for(case):
            a = refcases.run_ref(case, op, inp)
            b = refcases.run_oracle(case, op, inp)
            for k in a:
                _exact(b[k], a[k], "%s seed %d %s %s" % (case.name, seed, op, k))


def test_golden_files_present():
    files = glob.glob(os.path.join(GOLDEN, "*.npz"))
    assert len(files) >= 4
