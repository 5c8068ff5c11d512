"""The oracle reproduces its committed digests (tests/golden/oracle_digests.json)
for every golden scene: the small cases, configs 1-3, all 256 frames of
config 4 and all 64 requests of config 5. Regenerate with tests/golden/make_golden.py only
when a deliberate change to the oracle or a generator is made."""
import json
import os

import pytest

from oracle import refcpu
from parity import digest
from scene_cases import golden_cases

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "oracle_digests.json")))


def test_every_case_has_a_digest():
    assert sorted(GOLDEN) == sorted(n for n, _ in golden_cases())


@pytest.mark.parametrize("name,build", golden_cases(), ids=[n for n, _ in golden_cases()])
def test_oracle_digest(name, build):
    rgba, depth, st = refcpu.draw(build())
    assert digest(rgba, depth, st) == GOLDEN[name]
