"""The campaign CSV writer (campaign.py) against the reference's recorded raw
CSV lines (instrument/csv.hpp format, tests/golden/reference_csv_rows.json,
taken from experiments/c1/internal_calibration/internal_calibration_raw.csv):
the header and every recorded line come back character for character."""
import json
import os

from conftest import GOLDEN
import campaign


def _golden():
    return json.load(open(os.path.join(GOLDEN, "reference_csv_rows.json")))


def test_header_is_the_reference_schema():
    g = _golden()
    assert campaign.csv_header() == g["header"]
    assert len(campaign.CSV_COLUMNS) == 43


def test_recorded_lines_round_trip():
    g = _golden()
    assert len(g["lines"]) >= 40
    problems = set()
    for line in g["lines"]:
        row = campaign.parse_line(line)
        problems.add(row["problem"])
        assert campaign.csv_line(row) == line
    assert problems == {"bratu", "burgers", "chandra", "brusselator"}


def test_append_writes_header_once(tmp_path):
    g = _golden()
    path = str(tmp_path / "out.csv")
    for line in g["lines"][:3]:
        campaign.append_csv(path, campaign.parse_line(line))
    text = open(path).read().splitlines()
    assert text == [g["header"]] + g["lines"][:3]


def test_fold_csv_lines_reproduce_the_reference_rows():
    # calibrate_b6.cpp:128-137 writes doubles in the default stream format
    # (6 significant digits): the recorded rows of bratu_n31_coarse.csv
    import json
    import numa_hip as H
    from conftest import GOLDEN
    rows = json.load(open(GOLDEN + "/reference_data.json"))["fold_n31_coarse"]
    assert H.FOLD_CSV_HEADER.split(",") == ["lambda", "spectral_gap", "eigen_residual",
                                            "continuation_iters", "n_F", "n_JVP", "n_prec"]
    p = H.ContinuationPoint(rows[0]["lambda"], rows[0]["spectral_gap"],
                            rows[0]["eigen_residual"], rows[0]["continuation_iters"],
                            rows[0]["n_F"], rows[0]["n_JVP"], rows[0]["n_prec"])
    assert H.fold_csv_line(p) == "0,29.585,7.5108e-08,0,1,95,108"
    for r in rows[:10]:
        p = H.ContinuationPoint(r["lambda"], r["spectral_gap"], r["eigen_residual"],
                                r["continuation_iters"], r["n_F"], r["n_JVP"], r["n_prec"])
        f = H.fold_csv_line(p).split(",")
        assert float(f[0]) == r["lambda"] and float(f[1]) == r["spectral_gap"]
