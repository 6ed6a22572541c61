"""Pins the oracle's integer topology (edges, dihedrals, BSR pattern, slot
tables) with closed forms and brute-force enumerations written
independently of the oracle's algorithms."""
import itertools

import numpy as np
import pytest

import pyoracle
from chysx import scenes


def _sizes(n):
    pos, tris = scenes.cloth_grid(n)
    o = pyoracle.OracleSim(32)
    o.set_state(pos, np.zeros_like(pos), np.ones(len(pos)))
    o.set_mesh(tris)
    o.build_stretch(1.0)
    o.build_shear(1.0)
    o.build_bending(1.0)
    o.set_pins(scenes.top_corners(n), pos[scenes.top_corners(n)], 1.0)
    return o, tris


@pytest.mark.parametrize("n", [2, 3, 5, 17, 32])
def test_grid_closed_forms(n):
    o, tris = _sizes(n)
    s = o.sizes()
    c = scenes.grid_counts(n)
    assert (s["n"], s["tris"], s["edges"], s["dihedrals"], s["nnz"]) == (c["N"], c["T"], c["E"], c["D"], c["nnz"])
    ro, ci = o.pattern()
    lens = np.diff(ro)
    assert ro[-1] == c["nnz"]
    # interior rows of the Newton grid couple 6 one-ring + 6 two-ring vertices
    if n >= 5:
        interior = [y * n + x for y in range(2, n - 2) for x in range(2, n - 2)]
        assert np.all(lens[interior] == 12)


def _brute_dihedrals(tris):
    """Interior edges from an edge -> incident-triangles map (no hashing of
    directed edges): canonical (min, max, apex of min->max, apex of max->min)."""
    inc = {}
    for t, (a, b, c) in enumerate(tris):
        for u, v, w in ((a, b, c), (b, c, a), (c, a, b)):
            inc.setdefault((min(u, v), max(u, v)), []).append((u, v, w))
    out = set()
    for (lo, hi), lst in inc.items():
        if len(lst) != 2:
            continue
        fwd = [w for (u, v, w) in lst if (u, v) == (lo, hi)]
        bwd = [w for (u, v, w) in lst if (u, v) == (hi, lo)]
        if len(fwd) == 1 and len(bwd) == 1:
            out.add((lo, hi, fwd[0], bwd[0]))
    return out


def _random_mesh(seed, nv=60, nt=90):
    rng = np.random.default_rng(seed)
    pos, tris = scenes.cloth_grid(8)
    keep = rng.choice(len(tris), size=len(tris) - 20, replace=False)
    return tris[np.sort(keep)]


@pytest.mark.parametrize("n", [3, 9, 32])
def test_dihedrals_bruteforce_grid(n):
    _, tris = scenes.cloth_grid(n)
    d = pyoracle.detect_dihedrals(tris)
    assert set(map(tuple, d.tolist())) == _brute_dihedrals(tris.tolist())
    assert len(d) == len(set(map(tuple, d.tolist())))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_dihedrals_bruteforce_ragged(seed):
    tris = _random_mesh(seed)
    d = pyoracle.detect_dihedrals(tris)
    assert set(map(tuple, d.tolist())) == _brute_dihedrals(tris.tolist())


def test_edges_bruteforce():
    tris = _random_mesh(3)
    e = pyoracle.build_edges(tris)
    ref = sorted({(min(a, b), max(a, b)) for t in tris.tolist() for a, b in itertools.combinations(t, 2)})
    assert [tuple(x) for x in e.tolist()] == ref


def test_pattern_and_slots_bruteforce():
    n = 9
    pos, tris = scenes.cloth_grid(n)
    o = pyoracle.OracleSim(64)
    o.set_state(pos, np.zeros_like(pos), np.ones(len(pos)))
    o.set_mesh(tris)
    o.build_stretch(1.0)
    o.build_shear(1.0)
    o.build_bending(1.0)
    pins = np.array([0, 5, 80], np.int32)
    o.set_pins(pins, pos[pins], 1.0)
    dih, _ = o.dihedrals()
    pairs = set()
    for el in list(tris.tolist()) + list(dih.tolist()):
        for a, b in itertools.permutations(el, 2):
            if a != b:
                pairs.add((a, b))
    pairs = sorted(pairs)
    ro, ci = o.pattern()
    ref_ro = np.zeros(n * n + 1, np.int64)
    for r, _ in pairs:
        ref_ro[r + 1] += 1
    assert np.array_equal(ro, np.cumsum(ref_ro))
    assert np.array_equal(ci, np.array([c for _, c in pairs]))
    # slot semantics: diag -> -i-1; off-diag -> position of (i_a, i_b)
    for kind, elems, K in (("stretch", tris, 3), ("shear", tris, 3), ("bending", dih, 4)):
        sl = o.slots(kind, len(elems) * K * K).reshape(len(elems), K, K)
        for e in range(len(elems)):
            for a in range(K):
                for b in range(K):
                    i, j = elems[e][a], elems[e][b]
                    s = sl[e, a, b]
                    if i == j:
                        assert s == -i - 1
                    else:
                        assert ro[i] <= s < ro[i + 1] and ci[s] == j
    assert np.array_equal(o.slots("pin", 3), -pins - 1)


def test_pattern_out_of_range_raises():
    pos, tris = scenes.cloth_grid(3)
    o = pyoracle.OracleSim(32)
    o.set_state(pos, np.zeros_like(pos), np.ones(len(pos)))
    o.set_mesh(tris)
    o.build_stretch(1.0)
    o.set_pins(np.array([99], np.int32), np.zeros((1, 3)), 1.0)
    with pytest.raises(RuntimeError, match="out of range"):
        o.sizes()
    o2 = pyoracle.OracleSim(32)
    o2.set_state(pos[:5], np.zeros((5, 3)), np.ones(5))
    o2.set_mesh(tris)
    with pytest.raises(RuntimeError, match="out of range"):
        o2.build_stretch(1.0)
