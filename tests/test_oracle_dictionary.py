"""Dictionary construction (src/gabor.cpp:213-252) of the oracle and of the
HIP library's host code, against the closed-form table (golden fixture)."""
import json
import os

import pytest

import pyoracle as po

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "dictionary_closed_form.json")))


def test_config1_table_exact():
    d = po.Dictionary(GOLD["fs"], GOLD["N"], GOLD["energyError"])
    assert len(d.scales) == GOLD["L"]
    assert d.entries == GOLD["entries"]
    assert abs(d.root - GOLD["root"]) < 1e-11
    for row, sc in zip(GOLD["scales"], d.scales):
        assert (sc.Ngauss, sc.Nfft, sc.T, sc.F) == (row["Ngauss"], row["Nfft"], row["T"], row["F"]), row
        assert sc.T * sc.F == row["entries"]
        assert abs(sc.s - row["s"]) <= 5e-7, (sc.s, row["s"])


def test_scales_geometric():
    d = po.Dictionary(GOLD["fs"], GOLD["N"], GOLD["energyError"])
    ratios = [b.s / a.s for a, b in zip(d.scales, d.scales[1:])]
    assert max(ratios) - min(ratios) < 1e-12
    assert abs(d.scales[0].s - 2.0 / GOLD["fs"]) < 1e-15
    assert abs(d.scales[-1].s - GOLD["N"] / GOLD["fs"]) < 1e-12


@pytest.mark.parametrize("row", GOLD["totals"])
def test_other_workload_totals(row):
    d = po.Dictionary(row["fs"], row["N"], row["energyError"])
    assert len(d.scales) == row["L"]
    if "entries" in row:
        assert d.entries == row["entries"]
    else:
        assert abs(d.entries - row["entries_rounded"]) / row["entries_rounded"] < 0.002
    if "Nfft_max" in row:
        assert d.scales[-1].Nfft == row["Nfft_max"]
    if "Nfft_min" in row:
        assert d.scales[0].Nfft == row["Nfft_min"]
    if "T_max" in row:
        assert d.scales[0].T == row["T_max"] and d.scales[-1].T == row["T_min"]


def test_freq_max_clip():
    # the highest bin is the first at or above maxAtomFrequency (gabor.cpp:242-245)
    d = po.Dictionary(128.0, 1280, 0.01, freq_max=20.3)
    for sc in d.scales:
        df = 128.0 / sc.Nfft
        if sc.F < sc.Nfft // 2 + 1:
            assert (sc.F - 2) * df < 20.3 <= (sc.F - 1) * df, (sc.F, df)


def test_scale_limits_and_errors():
    d = po.Dictionary(256.0, 2048, 0.02, scale_min=0.05, scale_max=1.0)
    assert abs(d.scales[0].s - 0.05) < 1e-15 and abs(d.scales[-1].s - 1.0) < 1e-12
    with pytest.raises(po.OracleError):
        po.Dictionary(256.0, 2048, 1.5)


def test_library_dictionary_matches_oracle():
    """The engine's host dictionary is bit-identical (no device call)."""
    import empi_amd as E
    for fs, N, eps in [(128.0, 1280, 0.01), (256.0, 5120, 0.01), (512.0, 10240, 0.005), (256.0, 3001, 0.03),
                       (1000.0, 777, 0.1)]:
        d = po.Dictionary(fs, N, eps)
        scales, entries = E.dictionary(fs, N, eps)
        assert entries == d.entries
        assert [(s.s, s.Nfft, s.T, s.F) for s in d.scales] == [tuple(x) for x in scales]
