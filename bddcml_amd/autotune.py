"""Autotuner for the launch plan and launch shapes of one problem.

The device path has knobs that change how the work is cut into launches and
how each launch is shaped, not what is computed: the largest part that
nested dissection leaves whole (``nd_leaf``: the tile structure of the
interior factors and the number of elimination levels, hence of launches),
the XCD task order of the tile launches, the waves per workgroup of the
interior sweeps (``sweep_waves``), the form of the local preconditioner
launch (``pc_form``: two or four tiles in flight per workgroup) and when the
explicit inverse of the local operator is formed.  The tile edge itself
stays 64 (the unit of the matrix-core tiling, the LDS layout and every plan
array).  Which setting is fastest depends on the mesh, the subdomain sizes
and the iteration count, so this module times setup + solve on the caller's
own subdomains: by default a coordinate search (each knob in turn over its
values, the others at their best so far), or a given list of candidate
option sets.  Every candidate solves the same problem; the solutions must
agree (they differ only by rounding), which the tuner checks.

    from bddcml_amd.autotune import autotune
    ranking = autotune(subdomains, nsub)      # coordinate search
    best_options = ranking[0]["options"]

``python -m bddcml_amd.autotune --nel 16 --nsub 4`` tunes a cube problem.
"""
from __future__ import annotations

import argparse
import itertools
import json
import time

import numpy as np

from . import Bddcml

# option sets of a full grid: nested-dissection leaf size x task order x
# explicit-inverse mode (the ordering stays nested dissection, RCM is slower
# for every mesh measured)
DEFAULT_GRID = {
    "nd_leaf": (128, 256, 512),
    "task_order": (1, 0),
    "explicit_inverse": (2, 0),
}

# the coordinate search's knobs in the order searched, the library's
# default first: the plan (leaf size), then the launch shapes, then when the
# explicit inverse is formed
DEFAULT_KNOBS = {
    "nd_leaf": (512, 256, 128),
    "task_order": (1, 0),
    "sweep_waves": (0, 4, 16),
    "pc_form": (0, 1, 2),
    "explicit_inverse": (2, 0, 1),
}


def candidate_sets(grid: dict | None = None) -> list[dict]:
    """Cartesian product of a {option: values} grid as a list of dicts."""
    grid = grid or DEFAULT_GRID
    keys = list(grid)
    return [dict(zip(keys, vals)) for vals in itertools.product(*(grid[k] for k in keys))]


def _run(subdomains, nsub, numbase, options, levels, nsub2, setup_args, solve_args):
    if isinstance(levels, (list, tuple)):
        nsublev = list(levels)
    elif levels == 2:
        nsublev = [nsub, 1]
    elif levels == 3:
        nsublev = [nsub, nsub2, 1]
    else:
        raise ValueError(f"levels={levels}: an integer level count must be 2 or 3; pass the full "
                         "nsublev list (subdomains of every level, ending with 1) for more levels")
    b = Bddcml()
    b.init(len(nsublev), nsublev, numbase=numbase)
    for k, v in options.items():
        if b.set_option(k, v) != 0:
            b.finalize()
            raise ValueError(f"option {k}={v} rejected")
    for d in subdomains:
        b.upload_subdomain_data(d)
    t0 = time.perf_counter()
    b.setup_preconditioner(*setup_args)
    t1 = time.perf_counter()
    it, reason, cond = b.solve(*solve_args)
    t2 = time.perf_counter()
    u = b.interface_solution()
    b.finalize()
    return dict(setup_s=t1 - t0, solve_s=t2 - t1, total_s=t2 - t0, iterations=it,
                reason=reason), u


def autotune(subdomains, nsub, numbase=1, candidates=None, repeats=2, levels=2, nsub2=0,
             setup_args=(1, 0, 1, 1, 1, 0, 0, 0), solve_args=(0, 1e-8, 1000, 100),
             agree=1e-6, verbose=False, knobs=None):
    """levels: 2, 3 (with nsub2 second-level subdomains) or the full
    nsublev list of bddcml_init.
    With `candidates` (a list of option dicts) times setup + solve for each;
    without, runs a coordinate search over `knobs` ({option: values}, the
    first value the starting point; DEFAULT_KNOBS by default).  Each option
    set is timed as the best of `repeats` runs, after one untimed run that
    pays the one-time costs.  Returns every option set timed, sorted by time:
    [{"options": {...}, "total_s", "setup_s", "solve_s", "iterations"}, ...].
    Raises if a run does not converge or its interface solution differs
    from the first run's by more than `agree` (relative)."""
    subdomains = list(subdomains)
    if candidates is not None and not candidates:
        raise ValueError("no candidates")
    knobs = knobs if knobs is not None else DEFAULT_KNOBS
    if candidates is None and not knobs:
        raise ValueError("no knobs")
    state = dict(u_ref=None)
    timed = {}

    def measure(opts):
        key = tuple(sorted(opts.items()))
        if key in timed:
            return timed[key]
        best = None
        for _ in range(max(1, repeats)):
            r, u = _run(subdomains, nsub, numbase, opts, levels, nsub2, setup_args, solve_args)
            if r["reason"] != 0:
                raise RuntimeError(f"options {opts}: the solve did not converge ({r['reason']})")
            if state["u_ref"] is None:
                state["u_ref"] = u
            u_ref = state["u_ref"]
            err = np.abs(u - u_ref).max() / max(np.abs(u_ref).max(), 1e-300)
            if err > agree:
                raise RuntimeError(f"options {opts}: solution differs by {err:.2e}")
            if best is None or r["total_s"] < best["total_s"]:
                best = r
        best = dict(options=dict(opts), **{k: best[k] for k in ("total_s", "setup_s", "solve_s",
                                                                 "iterations")})
        if verbose:
            print(json.dumps(best), flush=True)
        timed[key] = best
        return best

    if candidates is not None:
        _run(subdomains, nsub, numbase, candidates[0], levels, nsub2, setup_args, solve_args)
        for opts in candidates:
            measure(opts)
    else:
        current = {k: v[0] for k, v in knobs.items()}
        _run(subdomains, nsub, numbase, current, levels, nsub2, setup_args, solve_args)
        for k, values in knobs.items():
            results = [measure(dict(current, **{k: v})) for v in values]
            current = min(results, key=lambda r: r["total_s"])["options"]
    return sorted(timed.values(), key=lambda r: r["total_s"])


def main():
    from .cube import CubeProblem
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--nel", type=int, default=16, help="elements per subdomain edge")
    ap.add_argument("--nsub", type=int, default=4, help="subdomains per cube edge")
    ap.add_argument("--kind", default="poisson", choices=("poisson", "elasticity"))
    ap.add_argument("--repeats", type=int, default=2)
    args = ap.parse_args()
    p = CubeProblem(args.nel, args.nsub, kind=args.kind, rhs="random", seed=2)
    ranking = autotune(p.subdomains(), p.nsub, p.numbase, repeats=args.repeats, verbose=True)
    print(f"{len(ranking)} option sets timed")
    print("best:", json.dumps(ranking[0]))


if __name__ == "__main__":
    main()
