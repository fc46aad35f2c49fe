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
