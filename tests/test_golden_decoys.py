"""Decoy / entrapment draws against vectors written by the reference's own functions
(tests/golden/make_decoy_vectors.py -> decoys_vectors.json; reference vulcan/index/decoys.py:28-168): the
pseudo-reverse, the seeded C-term-fixed shuffle permutations, and the decoy, entrapment and shuffled-decoy sets
with their redraw counts (CPU)."""
import json
import os

import numpy as np

from vulcan.index import decoys as D

VEC = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "decoys_vectors.json")))


def _perm(q):
    return None if q is None else np.asarray(q).tolist()


def test_pseudo_reverse_and_shuffle_permutations():
    assert [D.pseudo_reverse(p) for p in VEC["peptides"]] == VEC["pseudo_reverse"]
    for v in VEC["hash_shuffle_perm"]:
        assert D.hash_shuffle_perm(v["pep"], v["seed"], v["attempt"]).tolist() == v["perm"], v


def test_decoy_set():
    g = VEC["make_decoys"]
    out, n_shuf = D.make_decoys(VEC["peptides"], g["seed"])
    assert n_shuf == g["n_shuffled"]
    assert {k: [d, _perm(q)] for k, (d, q) in out.items()} == g["out"]


def test_entrapment_and_shuffled_decoys():
    g = VEC["make_entrapment"]
    taken = set(g["targets"])
    ent, nr, nd = D.make_entrapment(g["targets"], g["seed"], taken)
    assert (nr, nd) == (g["n_redrawn"], g["n_dropped"])
    assert {k: [e, _perm(q)] for k, (e, q) in ent.items()} == g["out"]
    s = VEC["make_shuffled_decoys"]
    sd, nr2, nd2 = D.make_shuffled_decoys(s["forward"], s["seed"], set())
    assert (nr2, nd2) == (s["n_redrawn"], s["n_dropped"])
    assert {k: [d, _perm(q)] for k, (d, q) in sd.items()} == s["out"]
