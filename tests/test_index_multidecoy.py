"""Multi-replicate L1 tables (build_multi_decoy_tables; reference build.py:257-418, decoys.py:72-168): pair
partners per decoy set, family sizes, entrapment / decoy rules, order independence, on-disk round trip (CPU)."""
import numpy as np

from vulcan.index import build as B
from vulcan.index import decoys as D
from vulcan.index import synth

PARAMS = {"entrap_seeds": [2001, 2002], "decoy_seeds": [1001, 1002],
          **synth.WORKLOADS[1].index_params()}


def _table(n=120, seed=3):
    return B.build_multi_decoy_tables(synth.proteome(n, seed), PARAMS)


def test_partners_and_families():
    """The reference's structure checks (tests/gpu/test_family.py:29-42)."""
    l1, lab = _table()
    assert len(l1.partners) == 2
    side = lab["decoy_set"]
    n = np.arange(len(l1))
    for j, pj in enumerate(l1.partners):
        fwd = side < 0
        assert (side[pj[fwd]] == j).all()                 # forward -> its set-j decoy
        assert (pj[pj[fwd]] == n[fwd]).all()              # symmetric
        other = (side >= 0) & (side != j)
        assert (pj[other] == n[other]).all()              # rows of other decoy sets: self
        assert np.array_equal(l1.length[pj], l1.length) and np.array_equal(l1.n_var[pj], l1.n_var)
    assert len(np.unique(l1.pf_id)) == len(l1)
    fam = np.bincount(lab["family"], minlength=4)
    assert fam[2] == 2 * fam[0] and fam[3] == 2 * fam[1] and fam[1] == 2 * fam[0] > 0
    assert np.all(np.diff(l1.mass32) >= 0)
    # decoy_seed / is_decoy / is_entrapment agree with the family and side codes
    assert np.array_equal(lab["is_decoy"], side >= 0)
    assert np.array_equal(lab["decoy_seed"][side >= 0], np.array([1001, 1002])[side[side >= 0]])
    assert set(np.unique(lab["entrap_seed"][lab["is_entrapment"]])) == {2001, 2002}


def test_entrapment_and_decoy_rules():
    targets, _ = B.digest(synth.proteome(60, 5), PARAMS["enzyme"])
    taken = set(targets)
    e1, _, _ = D.make_entrapment(targets, 2001, taken)
    e2, _, _ = D.make_entrapment(targets, 2002, taken)
    tset, rev = set(targets), {D.pseudo_reverse(t) for t in targets}
    ents = [e for e, _ in e1.values()] + [e for e, _ in e2.values()]
    assert len(set(ents)) == len(ents)                                  # no entrapment repeats
    for t, (e, perm) in list(e1.items()) + list(e2.items()):
        assert e != t and e not in tset and e not in rev and D.pseudo_reverse(e) != e
        assert e[-1] == t[-1] and sorted(e) == sorted(t) and D.apply_perm(t, perm) == e
    fwd = sorted(tset | set(ents))
    used: set = set()
    d1, _, _ = D.make_shuffled_decoys(fwd, 1001, used)
    d2, _, _ = D.make_shuffled_decoys(fwd, 1002, used)
    decs = [d for d, _ in d1.values()] + [d for d, _ in d2.values()]
    assert len(set(decs)) == len(decs) and not set(decs) & set(fwd)
    # order free: the same input in another order gives the same draws
    assert D.make_entrapment(list(reversed(targets)), 2001, set(targets))[0].keys() == e1.keys()


def test_save_and_load_keep_every_partner(tmp_path):
    l1, lab = _table(60, 7)
    l1.save(str(tmp_path / "m"), lab)
    back = B.L1Index(str(tmp_path / "m"))
    assert len(back.partners) == 2
    for a, b in zip(back.partners, l1.partners):
        assert np.array_equal(a, b)
    lb = B.read_labels(str(tmp_path / "m"))
    assert np.array_equal(lb["decoy_set"], lab["decoy_set"]) and np.array_equal(lb["family"], lab["family"])
