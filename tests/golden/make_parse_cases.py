#!/usr/bin/env python3
"""Write tests/golden/parse_cases.json: the reference CLI's verdict on a corpus
of thpack texts (run in the build container).

Each case is run through the reference's own g++ CLI (gpu/plp_gpu.cpp, built
from its sources by `make -C oracle ref` into oracle/_ref/plp_gpu_ref) as
`plp_gpu_ref <file> --single-pass --eps 0 --no-theta --json`; the fixture keeps
the text, the exit code, the stderr line for rejected texts, and the parsed
container and types for accepted ones (read back from the reference's own
JSON: n_boxes / placements need a decoder and are not kept).  The corpus
covers every message of the format's grammar (plp/parser.py:67-139 rules, as
gpu/plp_common.hpp:68-141 words them), the order in which two faults are
reported, and the strtoll-level details of a field (sign, saturation, blank
lines, carriage returns).

usage: python3 tests/golden/make_parse_cases.py
"""
from __future__ import annotations

import json
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.abspath(os.path.join(HERE, "..", ".."))
REF_CLI = os.path.join(REPO, "oracle", "_ref", "plp_gpu_ref")

ITEM = "1 10 1 20 1 30 1 5"


def one(*body: str) -> str:
    return "\n".join(["1", "1 1001", *body]) + "\n"


CASES = {
    "empty": "",
    "whitespace_only": "   \n\t\n \r\n",
    "count_zero": "0\n",
    "count_negative": "-3\n1 1\n587 233 220\n1\n" + ITEM + "\n",
    "count_not_int": "one\n",
    "count_only_line": "1\n",
    "count_line_extra_fields": "1 extra fields\n1 1001\n587 233 220\n1\n" + ITEM + "\n",
    "count_plus_sign": "+1\n1 1001\n587 233 220\n1\n" + ITEM + "\n",
    "problem_line_three_fields": one().replace("1 1001", "1 1001 7") + "587 233 220\n1\n" + ITEM,
    "missing_container": "1\n1 1001\n",
    "container_two_fields": one("587 233", "1", ITEM),
    "container_four_fields": one("587 233 220 1", "1", ITEM),
    "container_zero": one("587 0 220", "1", ITEM),
    "container_negative": one("-587 233 220", "1", ITEM),
    "container_not_int": one("587 23x3 220", "1", ITEM),
    "container_hex": one("0x10 233 220", "1", ITEM),
    "container_float": one("587.0 233 220", "1", ITEM),
    "container_not_int_and_zero": one("0 abc 220", "1", ITEM),
    "missing_type_count": "1\n1 1001\n587 233 220\n",
    "type_count_two_fields": one("587 233 220", "1 2", ITEM),
    "type_count_not_int": one("587 233 220", "x", ITEM),
    "truncated_items": one("587 233 220", "2", ITEM),
    "item_seven_tokens": one("587 233 220", "1", "1 10 1 20 1 30 1"),
    "item_nine_tokens": one("587 233 220", "1", ITEM + " 9"),
    "item_index_wrong": one("587 233 220", "1", "2 10 1 20 1 30 1 5"),
    "item_index_not_int": one("587 233 220", "1", "a 10 1 20 1 30 1 5"),
    "item_dim_zero": one("587 233 220", "1", "1 0 1 20 1 30 1 5"),
    "item_dim_negative": one("587 233 220", "1", "1 10 1 -20 1 30 1 5"),
    "item_dim_float": one("587 233 220", "1", "1 10.5 1 20 1 30 1 5"),
    "item_flag_two": one("587 233 220", "1", "1 10 2 20 1 30 1 5"),
    "item_flag_plus_one": one("587 233 220", "1", "1 10 +1 20 1 30 1 5"),
    "item_flag_word": one("587 233 220", "1", "1 10 1 20 yes 30 1 5"),
    "item_qty_negative": one("587 233 220", "1", "1 10 1 20 1 30 1 -5"),
    "item_qty_not_int": one("587 233 220", "1", "1 10 1 20 1 30 1 5x"),
    "item_dim_zero_and_bad_flag": one("587 233 220", "1", "1 0 7 20 1 30 1 5"),
    "item_bad_flag_and_bad_qty": one("587 233 220", "1", "1 10 7 20 1 30 1 -1"),
    "item_bad_dim_token_and_bad_index": one("587 233 220", "1", "3 x 1 20 1 30 1 5"),
    "trailing_data": one("587 233 220", "1", ITEM, "999"),
    "two_problems_declared_one_given": "2\n1 1001\n587 233 220\n1\n" + ITEM + "\n",
    "two_problems": "2\n1 1\n587 233 220\n1\n" + ITEM + "\n2 2\n100 100 100\n1\n" + ITEM + "\n",
    # accepted texts
    "blank_and_crlf_lines": "1\r\n\r\n1 1001\r\n  587 233 220 \r\n\n1\r\n" + ITEM + "\r\n\n",
    "tabs_and_signs": "1\n1\n\t+587\t233 +220\n+1\n+1 +10 1 20 1 30 1 +5\n",
    "problem_line_one_field": one().replace("1 1001", "7") + "587 233 220\n1\n" + ITEM + "\n",
    "zero_types": one("587 233 220", "0"),
    "negative_type_count": one("587 233 220", "-2"),
    "zero_qty": one("587 233 220", "1", "1 10 1 20 1 30 1 0"),
    "negative_zero": one("587 233 220", "1", "1 10 1 20 1 30 1 -0"),
    "leading_zeros": one("0587 0233 0220", "1", "01 010 1 020 1 030 1 05"),
    "saturated_dimension": one("587 233 220", "1", "1 99999999999999999999 1 20 1 30 1 5"),
}


def run_ref(text: str) -> dict:
    with tempfile.NamedTemporaryFile("w", suffix=".txt", delete=False) as f:
        f.write(text)
        path = f.name
    try:
        p = subprocess.run([REF_CLI, path, "--single-pass", "--eps", "0", "--no-theta", "--json"],
                           capture_output=True, text=True, timeout=120)
    finally:
        os.remove(path)
    out = {"exit": p.returncode}
    if p.returncode != 0:
        out["stderr"] = p.stderr.strip()
    else:
        out["n_boxes"] = json.loads(p.stdout)["n_boxes"]
    return out


def main() -> None:
    cases = []
    for name, text in CASES.items():
        r = run_ref(text)
        cases.append({"name": name, "text": text, **r})
        print(name, r)
    with open(os.path.join(HERE, "parse_cases.json"), "w") as f:
        json.dump({"source": "reference CLI gpu/plp_gpu.cpp (g++ build) on each text",
                   "cases": cases}, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
