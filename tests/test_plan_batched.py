"""The batched refinement loop of the interaction plan (the one the device patch
estimator feeds: every pass estimates the pending patches of all unfinished near
pairs in one call) gives the same plan as refining each pair on its own, when both
take the host estimates (HSBEM_PLAN_BATCHED=1 switches the host plan to the loop).
Covers tensor refinement, degree promotion, projected fans and the all-pairs mode."""
import math

import numpy as np
import pytest

from hornsculpt_amd import bem, meshes
from test_host import PLAN_CASES


@pytest.mark.parametrize("name,make,frequency,opts", PLAN_CASES, ids=[c[0] for c in PLAN_CASES])
def test_batched_loop_gives_the_same_plan(monkeypatch, name, make, frequency, opts):
    mesh = make()
    k = 2 * math.pi * frequency / 343.0
    num = bem.BemNumerics(near_relative_tolerance=opts.get("tolerance", 1e-4),
                          near_distance_ratio=opts.get("near_ratio", 1.0), near_maximum_depth=opts.get("depth", 6))
    monkeypatch.delenv("HSBEM_PLAN_BATCHED", raising=False)
    direct = bem.interaction_plan(mesh, num, k)
    monkeypatch.setenv("HSBEM_PLAN_BATCHED", "1")
    batched = bem.interaction_plan(mesh, num, k)
    assert len(direct["near"]) > 0 or name == "tetra"  # the tetrahedron's pairs all touch
    for key in ("touching", "near", "patches", "whole", "near_offsets", "near_trials"):
        assert np.array_equal(batched[key], direct[key]), key
    for key in ("candidate_pairs", "maximum_depth", "maximum_estimated_error"):
        assert batched[key] == direct[key], key


def test_batched_loop_reports_the_failing_pair(monkeypatch):
    """A pair that cannot reach its tolerance within the depth limit is refused with
    the pair's identity, as in the per-pair path."""
    mesh = meshes.icosphere(2)
    k = 2 * math.pi * 2400 / 343.0
    num = bem.BemNumerics(near_relative_tolerance=1e-12, near_maximum_depth=1)
    monkeypatch.delenv("HSBEM_PLAN_BATCHED", raising=False)
    with pytest.raises(RuntimeError, match="near panels") as direct:
        bem.interaction_plan(mesh, num, k)
    monkeypatch.setenv("HSBEM_PLAN_BATCHED", "1")
    with pytest.raises(RuntimeError, match="near panels") as batched:
        bem.interaction_plan(mesh, num, k)
    print(direct.value, "|", batched.value)
