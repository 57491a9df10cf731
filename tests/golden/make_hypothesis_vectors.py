"""Write tests/golden/hypothesis_vectors.json: precursor hypotheses and candidate row windows from the
reference's host rule (reference vulcan/gpu/host.py:64-91, make_hypotheses: charges tried when none is given,
isotope errors, Sage f32 mass arithmetic, windows by searchsorted on mass32) on a small synthetic table and
run -- unknown charges, NaN precursors, invalid spectra, masses at the table's ends and on duplicate masses.
Run once where the reference checkout is present (REFERENCE_ROOT, default /root/reference).
"""
import importlib.util
import json
import os
import sys

import numpy as np

REF = os.environ.get("REFERENCE_ROOT", "/root/reference")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hypothesis_vectors.json")


def main():
    spec = importlib.util.spec_from_file_location("ref_host", os.path.join(REF, "vulcan", "gpu", "host.py"))
    H = importlib.util.module_from_spec(spec)
    sys.modules["ref_host"] = H          # dataclasses look their module up while the class is built
    spec.loader.exec_module(H)
    rng = np.random.default_rng(5)
    mass32 = np.sort(np.concatenate([rng.uniform(500, 5000, 1500), np.full(5, 1500.25), [500.0, 5000.0]])
                     ).astype(np.float32)
    n = 150
    charge = rng.integers(0, 6, n).astype(np.int8)
    pmz = rng.uniform(300, 1800, n)
    pmz[::17] = np.nan
    # precursors whose neutral mass lands on a table mass (window edges) and at the table's ends
    for i, (m, z) in enumerate([(1500.25, 2), (500.0, 1), (5000.0, 3), (float(mass32[1234]), 2)]):
        charge[i] = z
        pmz[i] = (m + z * 1.007276) / z
    valid = rng.random(n) > 0.1
    off = np.zeros(n + 1, np.int64)
    sp = H.Spectra("golden", np.arange(n), np.arange(n), np.zeros(n), pmz, charge, np.zeros(0, np.float32),
                   np.zeros(0, np.float32), off)
    cases = []
    for kw in ({}, {"precursor_ppm": (-20.0, 5.0), "isotope_errors": (-1, 3), "precursor_charge": (2, 6)}):
        p = H.SearchParams(**kw)
        h = H.make_hypotheses(sp, valid, mass32, p)
        used = ("precursor_ppm", "isotope_errors", "precursor_charge", "max_fragment_charge")
        cases.append({"params": {k: getattr(p, k) for k in used},
                      "out": {k: np.asarray(v).tolist() for k, v in h.items()}})
    with open(OUT, "w") as fh:
        json.dump({"mass32": mass32.tolist(), "precursor_mz": [None if np.isnan(x) else x for x in pmz],
                   "charge": charge.tolist(), "valid": valid.tolist(), "cases": cases}, fh)
    print(f"wrote {OUT}: {[len(c['out']['spec']) for c in cases]} hypotheses")


if __name__ == "__main__":
    main()
