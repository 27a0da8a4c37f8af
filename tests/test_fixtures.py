"""The oracle's frozen bytes (tests/fixtures/oracle_fixtures.json, written by
tests/fixtures/make_fixtures.py): plans, level-0 hashes and oracle output hashes
of the seeded benchmark configurations, recomputed and compared.  A change to
the oracle, the planner or the input generator fails here, so it cannot move
together with the kernels the oracle checks."""
import json
import os

import pytest

from tests.fixtures import make_fixtures as M

FIX = json.load(open(os.path.join(os.path.dirname(M.__file__), "oracle_fixtures.json")))


@pytest.mark.parametrize("name,w,h,seed", M.SRGBA8, ids=[c[0] for c in M.SRGBA8])
def test_srgba8_oracle_bytes_frozen(name, w, h, seed, orc):
    assert M.srgba8_entry(w, h, seed) == FIX[name]


@pytest.mark.parametrize("name,w,h,seed", M.RGBA16F, ids=[c[0] for c in M.RGBA16F])
def test_rgba16f_oracle_bytes_frozen(name, w, h, seed, orc):
    assert M.rgba16f_entry(w, h, seed) == FIX[name]


def test_mixed_workload_frozen():
    assert M.mixed_entry(*M.MIXED) == FIX["512 mixed"]
