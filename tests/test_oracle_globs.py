"""Glob / corner classification of the oracle against (a) hand counts on
structured cubes and (b) an independent pure-Python restatement of the
reference's dd_create_globs (src/module_dd.f90:8925-10554) written from
the reference text, sharing no code with the oracle."""
import numpy as np
import pytest

from bddcml_amd.cube import CubeProblem


def py_globs(problem):
    """Returns (corner global node set, set of edge keys, set of face keys)
    where a glob key is the frozenset of its global node ids."""
    b = problem.numbase
    subs = [problem.subdomain(s) for s in range(problem.nsub)]
    node_subs = {}
    for s, d in enumerate(subs):
        for g in d["isngn"] - b:
            node_subs.setdefault(int(g), []).append(s)
    data = []
    for s, d in enumerate(subs):
        isngn = [int(g) for g in d["isngn"] - b]
        xyz = np.asarray(d["xyz"])
        iface = [l for l, g in enumerate(isngn) if len(node_subs[g]) > 1]
        adj = sorted({t for g in isngn for t in node_subs[g] if t != s})
        # sharers of each interface node in neighbour order (all components 1)
        key = {l: tuple(t for t in adj if t in node_subs[isngn[l]]) for l in iface}
        # first pass: vertices
        gtype = {}
        seen = set()
        for l in iface:
            if l in seen:
                continue
            grp = [m for m in iface if key[m] == key[l] and m not in seen]
            seen.update(grp)
            if len(key[l]) == 1:
                t = 1
            elif len(grp) > 1:
                t = 2
            else:
                t = 3
            for m in grp:
                gtype[m] = t
        for l in iface:
            if len(key[l]) == 1:
                gtype[l] = 1
        # face triangle search per neighbour sharing a face
        for t in adj:
            shared = sorted((isngn[l], l) for l in iface if t in key[l])
            shared = [l for _, l in shared]
            if not any(gtype[l] == 1 for l in shared):
                continue
            X = xyz[shared]
            d0 = ((X - X[0]) ** 2).sum(1)
            c1 = int(np.argmax(d0))
            d1 = ((X - X[c1]) ** 2).sum(1)
            c2 = int(np.argmax(d1))
            picks = [c1, c2]
            if len(shared) > 2:
                p1, p2 = X[c1], X[c2]
                cr = np.cross(p2 - p1, X - p1)
                area = (cr ** 2).sum(1)
                if area.max() > 0:
                    picks.append(int(np.argmax(area)))
                    picks.append(len(area) - 1 - int(np.argmax(area[::-1])))
            for c in picks:
                gtype[shared[c]] = 3
        data.append(dict(isngn=isngn, iface=iface, adj=adj, key=key,
                         lcorners={isngn[l] for l in iface if gtype[l] == 3},
                         fixed={l for l in iface if d["ifix"][l] > 0}))
    corners = set()
    edges, faces = set(), set()
    for s, dd in enumerate(data):
        allc = set(dd["lcorners"])
        for t in dd["adj"]:
            allc |= data[t]["lcorners"]
        ign = {dd["isngn"][l] for l in dd["iface"]}
        mine = allc & ign
        corners |= mine
        rest = [l for l in dd["iface"] if dd["isngn"][l] not in mine and l not in dd["fixed"]]
        groups = {}
        for l in rest:
            groups.setdefault(dd["key"][l], set()).add(dd["isngn"][l])
        for k, nodes in groups.items():
            (faces if len(k) == 1 else edges).add(frozenset(nodes))
    return corners, edges, faces


@pytest.mark.parametrize("nel,ns", [(2, 2), (3, 2), (2, 3), (4, 3), (3, 4)])
def test_globs_match_python_restatement(oracle_lib, nel, ns):
    p = CubeProblem(nel, ns)
    o, _ = oracle_lib.solve_cube(p)
    corners, edges, faces = py_globs(p)
    c = o.counts()
    assert set(int(v) for v in o.corner_nodes()) == corners
    assert c["nedge"] == len(edges)
    assert c["nface"] == len(faces)
    o.close()


def test_hand_counts_2x2x2():
    """2x2x2 subdomains of 2^3 elements: the centre node plus the corners of
    every face square (nodes with one coordinate 2 and the others in {0,2,4},
    27 - 8 = 19), 6 half-edges and 12 faces after Dirichlet removal."""
    p = CubeProblem(2, 2)
    corners, edges, faces = py_globs(p)
    assert len(corners) == 19 and len(edges) == 6 and len(faces) == 12


def test_cfg1_counts(oracle_lib):
    """32^3 elements, 4^3 subdomains of 8^3: 27 interior vertices plus face
    triangle corners; 108 edges and 144 faces as in structured counting."""
    p = CubeProblem(8, 4)
    o, (it, reason, cond) = oracle_lib.solve_cube(p)
    c = o.counts()
    assert (c["nedge"], c["nface"]) == (108, 144)
    assert c["ncorner"] == 117
    assert c["nig"] == 8937
    assert reason == 0
    o.close()
