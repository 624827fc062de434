"""The thpack parser (C ABI plp_parse_instance and the plp_gpu CLI) against the
reference CLI's verdicts on a corpus of malformed and odd texts
(tests/golden/parse_cases.json, written by make_parse_cases.py from the
reference's g++ build of gpu/plp_gpu.cpp).

A rejected text must be rejected with the reference's message and exit code
(the CLI prints `instance error: <message>` and exits 1 before any device is
touched); an accepted text must parse, with the int-cast field values of
gpu/plp_common.hpp:128-136.
"""
import json
import os
import subprocess
import tempfile

import pytest

from conftest import REPO, golden_path

with open(golden_path("parse_cases.json")) as _f:
    CASES = json.load(_f)["cases"]

CLI = os.path.join(REPO, "plp-gpu_amd", "build", "plp_gpu")


@pytest.fixture(scope="module")
def plp():
    import plp_amd
    return plp_amd


def _ids(cases):
    return [c["name"] for c in cases]


REJECTED = [c for c in CASES if c["exit"] != 0]
ACCEPTED = [c for c in CASES if c["exit"] == 0]


def test_corpus_covers_every_message():
    msgs = {c["stderr"].split(": ", 1)[1].split(":")[0] for c in REJECTED}
    for m in ("empty file", "instance count must be >= 1", "not an integer",
              "malformed problem line", "container line must have 3 dimensions",
              "container dimensions must be positive", "type-count line must be a single integer",
              "instance truncated", "item line must have 8 tokens", "item index must be sequential",
              "item dimensions must be positive", "orientation flags must be 0/1",
              "item qty must be >= 0", "trailing data after last instance",
              "expected exactly 1 instance, got 2"):
        assert m in msgs, m


@pytest.mark.parametrize("case", REJECTED, ids=_ids(REJECTED))
def test_cli_rejects_like_reference(case):
    with tempfile.NamedTemporaryFile("w", suffix=".txt", delete=False) as f:
        f.write(case["text"])
        path = f.name
    try:
        p = subprocess.run([CLI, path, "--single-pass", "--eps", "0", "--no-theta", "--json"],
                           capture_output=True, text=True, timeout=60)
    finally:
        os.remove(path)
    assert (p.returncode, p.stderr.strip()) == (case["exit"], case["stderr"])


@pytest.mark.parametrize("case", REJECTED, ids=_ids(REJECTED))
def test_capi_rejects_with_reference_message(plp, case):
    with pytest.raises(plp.PlpError) as ei:
        plp.parse_instance(case["text"])
    msg = case["stderr"].split("instance error: ", 1)[1]
    assert str(ei.value) == f"{msg} (status {ei.value.code})"


@pytest.mark.parametrize("case", ACCEPTED, ids=_ids(ACCEPTED))
def test_capi_accepts_like_reference(plp, case):
    L, W, H, types = plp.parse_instance(case["text"])
    assert L > 0 and W > 0 and H > 0
    if case["n_boxes"] > 0:
        assert len(types) == 1 and int(types[0][6]) == 5


def test_field_values_follow_strtoll_and_int_cast(plp):
    by = {c["name"]: c for c in CASES}
    L, W, H, types = plp.parse_instance(by["tabs_and_signs"]["text"])
    assert (L, W, H) == (587, 233, 220) and types[0].tolist() == [10, 1, 20, 1, 30, 1, 5]
    L, W, H, types = plp.parse_instance(by["leading_zeros"]["text"])
    assert (L, W, H) == (587, 233, 220) and types[0].tolist() == [10, 1, 20, 1, 30, 1, 5]
    # 99999999999999999999 saturates at LLONG_MAX, whose int cast is -1
    _, _, _, types = plp.parse_instance(by["saturated_dimension"]["text"])
    assert int(types[0][0]) == -1
    _, _, _, types = plp.parse_instance(by["negative_type_count"]["text"])
    assert len(types) == 0
