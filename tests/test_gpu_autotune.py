"""The autotuner on a small problem: every candidate converges to the same
interface solution, the ranking is sorted by time and carries its options."""
import sys

import pytest

from bddcml_amd.autotune import autotune, candidate_sets
from bddcml_amd.cube import CubeProblem

pytestmark = pytest.mark.gpu


def test_autotune_ranks_candidates():
    print("case autotune 4x3 poisson", flush=True)
    sys.stdout.flush()
    p = CubeProblem(4, 3, kind="poisson", rhs="random", seed=41)
    cands = [{"nd_leaf": 128, "task_order": 1}, {"nd_leaf": 256, "task_order": 0},
             {"nd_leaf": 256, "explicit_inverse": 1}]
    ranking = autotune(p.subdomains(), p.nsub, p.numbase, candidates=cands, repeats=1,
                       solve_args=(0, 1e-10, 300, 50), agree=1e-8)
    assert len(ranking) == 3
    assert [r["total_s"] for r in ranking] == sorted(r["total_s"] for r in ranking)
    assert sorted(map(str, (r["options"] for r in ranking))) == sorted(map(str, cands))
    its = {r["iterations"] for r in ranking}
    assert max(its) - min(its) <= 1


def test_autotune_coordinate_search_over_launch_shapes():
    """The default search: every knob's values timed with the others at the
    best so far (the library defaults timed once), the launch-shape options
    accepted, one solution throughout."""
    from bddcml_amd.autotune import DEFAULT_KNOBS
    print("case autotune coordinate search 4x3 poisson", flush=True)
    sys.stdout.flush()
    p = CubeProblem(4, 3, kind="poisson", rhs="random", seed=42)
    ranking = autotune(p.subdomains(), p.nsub, p.numbase, repeats=1,
                       solve_args=(0, 1e-10, 300, 50), agree=1e-8)
    distinct = 1 + sum(len(v) - 1 for v in DEFAULT_KNOBS.values())
    assert len(ranking) == distinct
    assert [r["total_s"] for r in ranking] == sorted(r["total_s"] for r in ranking)
    for r in ranking:
        assert set(r["options"]) == set(DEFAULT_KNOBS)
    seen = {k: {r["options"][k] for r in ranking} for k in DEFAULT_KNOBS}
    for k, v in DEFAULT_KNOBS.items():
        assert seen[k] == set(v), k
    its = {r["iterations"] for r in ranking}
    assert max(its) - min(its) <= 1
