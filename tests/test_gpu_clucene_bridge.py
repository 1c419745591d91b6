"""GPU run of the CLucene-bridge check (tests/cpp/clucene_bridge_check.cpp):
ConjunctionAdapter over ConjunctionBatch results against a CPU restatement of
ConjunctionScorer's leapfrog (src/core/CLucene/search/ConjunctionScorer.cpp:52-121),
400 queries of 1-4 terms, next() and skipTo() walks, docIDs and scores equal."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "clucene_bridge_check")


def test_clucene_bridge_matches_leapfrog():
    assert os.path.exists(BIN), "build first: make -C tests/cpp"
    print("case clucene_bridge_check", flush=True)
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=100)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0 and "all equal" in r.stdout
