"""Pins of the PSM-feature oracle (oracle/psm.py) by closed forms, and the Parquet writer (CPU only)."""
import math

import numpy as np
import pytest

from oracle import psm as OP


def test_lnfact_is_stirling_plus_half_log():
    """The reference's lnfact (engine.py:354-358) is n ln n - n + ln(n)/2 + ln(2 pi n)/2: Stirling's series of
    ln n! plus ln(n)/2 without the 1/(12 n) term, so it falls short of lgamma(n + 1) + ln(n)/2 by 1/(12 n) up to
    the next term of the series, 1/(360 n^3) (and lnfact(0) = 1)."""
    assert OP.lnfact(0) == 1.0
    for n in (1, 2, 3, 7, 20, 150, 1000):
        d = float(OP.lnfact(n)) - (math.lgamma(n + 1) + 0.5 * math.log(n))
        assert -1.0 / (12 * n) - 1e-12 <= d <= -1.0 / (12 * n) + 1.0 / (360 * n ** 3) + 1e-12


@pytest.mark.parametrize("lam,k", [(0.5, 0), (0.5, 3), (2.0, 2), (7.5, 20), (40.0, 12), (0.01, 5)])
def test_poisson_is_log10_pmf_with_the_lnfact_term(lam, k):
    """log10(lam^k e^-lam / e^lnfact(k)) = (k ln lam - lam - lnfact(k)) / ln 10, i.e. scipy's Poisson log pmf with
    ln k! replaced by lnfact(k)."""
    from scipy.stats import poisson
    got = float(OP.poisson(lam * 1000, 1000, k))
    want = (poisson.logpmf(k, lam) + math.lgamma(k + 1) - float(OP.lnfact(k))) / math.log(10)
    assert abs(got - want) <= 1e-9 * max(1.0, abs(want))


def test_poisson_edges():
    """An overflowing lam^k (finite denominator) becomes 1e-325, which is 0 in f64: log10 = -inf (the reference's
    underflow edge); no scored candidates gives NaN."""
    assert OP.poisson(300.0, 1.0, 130) == -np.inf
    assert np.isnan(OP.poisson(0, 0, 3)) and np.isnan(OP.poisson(5, 0, 3))


def test_features_wiring():
    """Hand-built selections: two groups (spectra 0 and 2), report 2, brute-force expectations per slot."""
    report = 2
    sel = np.array([1, 0, 2, -1])                    # group 0: candidates 1 then 0; group 1: candidate 2 only
    next_hs = np.array([30.0, 0.0, 0.0, 0.0])
    grp_spec = np.array([0, 2])
    c_row = np.array([5, 3, 7])
    c_hyp = np.array([0, 1, 3])
    c_cnt = np.array([9, -1, 4])
    k2 = {"hs": np.array([30.0, 41.5, 12.25]), "sb": np.float32([1.5, 2.0, 0.25]), "sy": np.float32([0.5, 4.0, 1.0]),
          "ppm": np.float32([1.0, 2.0, 3.0]), "mb": np.int16([2, 3, 1]), "my": np.int16([1, 4, 0]),
          "lb": np.int16([2, 3, 1]), "ly": np.int16([1, 3, 0])}
    hyp_off = np.array([0, 2, 2, 4])                 # spectrum 0: hypotheses 0, 1; spectrum 2: 2, 3
    h_z = np.int8([2, 3, 2, 4])
    h_iso = np.int8([0, 1, 0, -1])
    mass32 = np.float32(np.arange(10) * 100.0 + 500.0)
    length = np.uint8([7] * 10)
    calc5 = np.float64(mass32[5])
    h_pm = np.float32([calc5 * (1 + 10e-6), 800.0 + 1.00335, 1000.0, 1200.0 - 1.00335])
    h_scored = np.array([10, 30, 5, 15])
    h_matched = np.array([20, 40, 1, 9])
    tic = np.float32([8.0, 1.0, 5.0])
    c = OP.features(sel, next_hs, grp_spec, c_row, c_hyp, c_cnt, k2, hyp_off, h_z, h_iso, h_pm, h_scored, h_matched,
                    tic, mass32, length, report)
    assert c["rank"].tolist() == [1, 2, 1, 0]
    assert c["spec_idx"][:3].tolist() == [0, 0, 2] and c["pf_row"][:3].tolist() == [3, 5, 7]
    assert c["delta_best"][:3].tolist() == [0.0, 41.5 - 30.0, 0.0]
    assert c["delta_next"][:3].tolist() == [41.5 - 30.0, 30.0, 12.25]
    assert c["scored_candidates"][:3].tolist() == [40, 40, 20]
    assert c["matched_peaks"][:3].tolist() == [7, 3, 1] and c["k1_count"][:3].tolist() == [-1, 9, 4]
    assert c["summed_intensity"][:3].tolist() == [6.0, 2.0, 1.25]
    assert c["matched_intensity_pct"][:3].tolist() == [75.0, 25.0, 25.0]
    assert c["longest_y_pct"][1] == np.float32(1 / 7)
    # symmetric ppm, isotope 0; pm carries its f32 rounding (half an ulp at 1000 Da is 0.03 ppm)
    assert abs(float(c["delta_mass_ppm"][1]) - 2e6 * 10e-6 / (2 + 10e-6)) < 0.07
    lam = (20 + 40) / (10 + 30)
    assert abs(c["poisson"][0] - math.log10(lam ** 7 * math.exp(-lam) / math.exp(float(OP.lnfact(7))))) < 1e-12
    assert c["charge"][:3].tolist() == [3, 2, 4] and c["isotope_error"][:3].tolist() == [1, 0, -1]


def test_write_psms_roundtrip(tmp_path):
    import pyarrow as pa
    import pyarrow.parquet as pq

    from vulcan.gpu.engine import write_psms
    tbl = pa.table({"run": ["r1", "r1"], "spec_idx": np.int64([0, 1]), "hyperscore": np.float64([1.5, 2.5])})
    path = write_psms(tbl, str(tmp_path))
    assert path.endswith("r1.psms.parquet")
    assert pq.read_table(path).equals(tbl)
