"""Shared fixtures. `gpu` tests need a ROCm device and the built libvulcan_hip.so; the rest run on CPU."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "vulcan-proteomics_amd"))
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and libvulcan_hip.so")


STATIC = {"C": 57.021464}


@pytest.fixture(scope="session")
def small_plain():
    """small_plain-shaped table (2,000 proteins) and 2,000 of its spectra."""
    from vulcan.index import synth
    l1, lab = synth.build_index(0)
    sp, truth, _ = synth.spectra(l1, lab, 0, n=2000)
    w = synth.WORKLOADS[0]
    return {"l1": l1, "labels": lab, "sp": sp, "truth": truth, "p": w.search_params(), "static": STATIC,
            "var": {}}


@pytest.fixture(scope="session")
def proteome_small():
    """workload-1 semantics (1 missed cleavage, Met-ox <= 2, 2+ fragments for z >= 3) on 3,000 proteins."""
    from vulcan.index import synth
    l1, lab = synth.build_index(1, n_proteins=3000)
    sp, truth, _ = synth.spectra(l1, lab, 1, n=3000)
    w = synth.WORKLOADS[1]
    return {"l1": l1, "labels": lab, "sp": sp, "truth": truth, "p": w.search_params(), "static": STATIC,
            "var": dict(w.var_mods)}


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
