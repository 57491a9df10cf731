"""Write tests/golden/decoys_vectors.json: input/output vectors of the reference's peptide decoy and entrapment
functions (vulcan/index/decoys.py:28-168 of the reference: pseudo_reverse, hash_shuffle_perm, make_decoys,
make_entrapment, make_shuffled_decoys), which no closed form pins. Run once where the reference checkout is
present (REFERENCE_ROOT, default /root/reference); the tests read only the JSON it writes.
"""
import importlib.util
import json
import os

import numpy as np

REF = os.environ.get("REFERENCE_ROOT", "/root/reference")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "decoys_vectors.json")


def peptides(n, seed):
    rng = np.random.default_rng(seed)
    aa = np.array(list("ACDEFGHKLMNPQRSTVWY"))
    out = []
    for _ in range(n):
        body = "".join(rng.choice(aa, size=int(rng.integers(6, 20))))
        out.append(body + rng.choice(["K", "R"]))
    # palindromes and pairs whose pseudo-reverse collides with another forward peptide
    out += ["ABCBAK", "LLLLLLK", "PEPTIDEK", "EDITPEPK", "GGAGGR", "MMAMMR"]
    return sorted(set(out))


def main():
    spec = importlib.util.spec_from_file_location("ref_decoys", os.path.join(REF, "vulcan", "index", "decoys.py"))
    D = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(D)
    fwd = peptides(60, 7)
    vec = {"peptides": fwd}
    vec["pseudo_reverse"] = [D.pseudo_reverse(p) for p in fwd]
    vec["hash_shuffle_perm"] = [{"pep": p, "seed": s, "attempt": a, "perm": D.hash_shuffle_perm(p, s, a).tolist()}
                                for p in fwd[:20] for s in (1001, 2001) for a in (0, 1, 3)]
    dec, n_shuf = D.make_decoys(fwd, 1001)
    vec["make_decoys"] = {"seed": 1001, "n_shuffled": int(n_shuf),
                          "out": {k: [d, None if q is None else np.asarray(q).tolist()] for k, (d, q) in dec.items()}}
    targets = fwd[:40]
    taken = set(targets)
    ent, nr, nd = D.make_entrapment(targets, 2001, taken)
    vec["make_entrapment"] = {"targets": targets, "seed": 2001, "n_redrawn": int(nr), "n_dropped": int(nd),
                              "out": {k: [e, np.asarray(q).tolist()] for k, (e, q) in ent.items()}}
    forward = sorted(set(targets) | {e for e, _ in ent.values()})
    used: set = set()
    sd, nr2, nd2 = D.make_shuffled_decoys(forward, 1002, used)
    vec["make_shuffled_decoys"] = {"forward": forward, "seed": 1002, "n_redrawn": int(nr2), "n_dropped": int(nd2),
                                   "out": {k: [d, np.asarray(q).tolist()] for k, (d, q) in sd.items()}}
    with open(OUT, "w") as fh:
        json.dump(vec, fh, indent=0, sort_keys=True)
    print(f"wrote {OUT}: {len(fwd)} peptides")


if __name__ == "__main__":
    main()
