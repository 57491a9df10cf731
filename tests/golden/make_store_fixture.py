"""Write tests/golden/ref_run.vspec.arrow (+ .json sidecar) with the reference's own store writer (reference
vulcan/io/store.py:62-152, schema.py), and ref_run_inputs.json with the arrays it was given: a fixture of the
store format as the reference's converters produce it (full schema, two record batches, MS1 rows, unknown
charges, vendor columns this build does not write). Run once where the reference checkout is present
(REFERENCE_ROOT, default /root/reference); the tests read only the files it writes.
"""
import json
import os
import sys

import numpy as np

REF = os.environ.get("REFERENCE_ROOT", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))


def batch(rng, n, scan0):
    npk = rng.integers(0, 12, n)
    off = np.concatenate([[0], np.cumsum(npk)]).astype(np.int64)
    level = np.where(np.arange(n) % 4 == 0, 1, 2).astype(np.uint8)
    z = rng.integers(0, 5, n).astype(np.int8)
    cols = {"scan": np.arange(scan0, scan0 + n, dtype=np.uint32), "ms_level": level,
            "rt_min": np.round(rng.uniform(0, 90, n), 4), "injection_time_ms": rng.uniform(5, 50, n).astype(np.float32),
            "precursor_mz": np.where(level == 2, np.round(rng.uniform(350, 1500, n), 5), np.nan),
            "precursor_charge": z, "precursor_charge_mask": (z == 0) | (level == 1),
            "activation": np.where(level == 2, 2, 0).astype(np.uint8), "activation_mask": level == 1,
            "analyzer": np.ones(n, np.uint8), "polarity": np.ones(n, np.int8),
            "tic": rng.uniform(1e5, 1e7, n).astype(np.float32), "peak_source": np.zeros(n, np.uint8),
            "peak_offsets": off, "mz": np.sort(rng.uniform(100, 2000, off[-1])).astype(np.float32),
            "intensity": rng.lognormal(8, 1, off[-1]).astype(np.float32)}
    return cols


def main():
    sys.path.insert(0, REF)
    from vulcan.io import store as RS
    rng = np.random.default_rng(11)
    path = os.path.join(HERE, "ref_run.vspec.arrow")
    meta = {"run_name": "ref_run", "vendor": "synthetic", "created_utc": "2026-01-01T00:00:00",
            "instrument_model": "none", "source_path": "synthetic", "source_size": 0, "converter": "fixture",
            "converter_version": "0", "peak_order": "mz_ascending", "acquisition": "DDA"}
    w = RS.StoreWriter(path, meta)
    parts = [batch(rng, 21, 100), batch(rng, 13, 121)]
    for c in parts:
        w.write(c)
    w.close()
    inputs = {}
    for k in ("scan", "ms_level", "rt_min", "precursor_mz", "precursor_charge", "mz", "intensity"):
        inputs[k] = np.concatenate([c[k] for c in parts]).tolist()
    inputs["peak_counts"] = np.concatenate([np.diff(c["peak_offsets"]) for c in parts]).tolist()
    inputs["precursor_charge_null"] = np.concatenate([c["precursor_charge_mask"] for c in parts]).tolist()
    with open(os.path.join(HERE, "ref_run_inputs.json"), "w") as fh:
        json.dump(inputs, fh)
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
