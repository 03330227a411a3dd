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
