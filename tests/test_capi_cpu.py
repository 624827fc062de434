"""CPU-side checks of the product: the C-ABI library loads and exports every
symbol include/plp_hip.h declares, the host parser and CLI behave like the
reference's (no compute call needs a GPU here)."""
import ctypes as C
import os
import random
import re
import struct
import subprocess

import pytest

from conftest import REPO, golden_path

LIB = os.path.join(REPO, "plp-gpu_amd", "build", "libplp_hip.so")
CLI = os.path.join(REPO, "plp-gpu_amd", "build", "plp_gpu")
HEADER = os.path.join(REPO, "include", "plp_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(plp_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(LIB)
    syms = declared_symbols()
    assert len(syms) >= 25
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_lists_every_declared_symbol():
    import plp_amd
    assert sorted(plp_amd.EXPORTS) == declared_symbols()


def test_status_strings():
    import plp_amd
    lib = plp_amd._lib
    for code in range(8):
        assert lib.plp_status_string(code)
    assert lib.plp_status_string(99) == b"unknown status"


def test_capi_parser_matches_reference_fixture():
    import plp_amd
    L, W, H, types = plp_amd.parse_instance(open(golden_path("data", "instances", "BR1_050.txt")).read())
    assert (L, W, H) == (587, 233, 220)
    assert types.tolist() == [[60, 1, 40, 1, 32, 1, 64], [98, 1, 75, 1, 55, 1, 40],
                              [60, 1, 59, 1, 39, 1, 64]]


@pytest.mark.parametrize("bad", ["", "0", "1\n1 0\n587 233\n1\n1 1 1 1 1 1 1 1",
                                 "1\n1 0\n587 233 220\n1\n1 10 2 20 1 30 1 5",
                                 "1\n1 0\n587 233 220\n1\n1 10 1 20 1 30 1 5\n7"])
def test_capi_parser_rejects(bad):
    import plp_amd
    with pytest.raises(plp_amd.PlpError) as ei:
        plp_amd.parse_instance(bad)
    assert ei.value.code == plp_amd.PLP_E_FORMAT


def test_device_count_does_not_fail():
    import plp_amd
    assert plp_amd.device_count() >= 0


def test_engine_needs_a_device_or_reports_it():
    """Without a GPU the context fails with a HIP status instead of running
    anything on the CPU."""
    import plp_amd
    if plp_amd.device_count() > 0:
        pytest.skip("device present")
    with pytest.raises(plp_amd.PlpError) as ei:
        plp_amd.Ctx(0)
    assert ei.value.code == plp_amd.PLP_E_HIP


def test_cli_selftest_matches_cpython():
    out = subprocess.run([CLI, "--selftest"], capture_output=True, text=True, check=True).stdout
    want = []
    for s in (0, 1, 7, 123456789):
        want.append(f"SEED {s}")
        r = random.Random(s)
        want += [format(struct.unpack("<Q", struct.pack("<d", r.random()))[0], "016x")
                 for _ in range(16)]
        want += [str(r.getrandbits(13)) for _ in range(16)]
    for v in (0.0, 1.0, 0.5, 0.1, 2.0 / 3.0, 0.5785123966942148, 0.6132075471698113,
              0.6595744680851063, 1e-05, 1e-04, 1e15, 1e16, 0.001, 123456789.5, 1.0 / 121.0,
              1.0 / 106.0):
        want.append(f"REPR {v!r}")
    assert out.splitlines() == want


@pytest.mark.parametrize("args,code", [
    ([], 2),
    (["--bogus"], 2),
    (["a.txt", "b.txt"], 2),
    (["{BR1}", "--order", "weird"], 2),
    (["{BR1}", "--batch", "4"], 2),            # --batch requires --json
    (["{BR1}", "--device", "4"], 2),           # --device requires --json
    (["{BR1}", "--batch", "2", "--device", "2", "--json"], 2),
    (["{BR1}", "--passes", "2", "--single-pass"], 2),
    (["/nonexistent/file.txt"], 1),
    (["{BAD}"], 1),
])
def test_cli_usage_and_instance_errors(tmp_path, args, code):
    bad = tmp_path / "bad.txt"
    bad.write_text("1\n1 0\n587 233 220\n1\n1 10 2 20 1 30 1 5\n")
    br1 = golden_path("data", "instances", "BR1_050.txt")
    argv = [CLI] + [x.replace("{BR1}", br1).replace("{BAD}", str(bad)) for x in args]
    p = subprocess.run(argv, capture_output=True, text=True)
    assert p.returncode == code, p.stderr


def test_synth_generator_is_deterministic_and_follows_qty_rule():
    from plp_amd import synth
    assert synth.multi_instance() == synth.multi_instance()
    L, W, H, types = synth.parse_types(synth.plain_instance())
    assert (L, W, H) == synth.CONTAINER and len(types) == 3
    for (l, lf, w, wf, h, hf, q) in types:
        assert 30 <= min(l, w, h) and max(l, w, h) <= 120 and lf == wf == hf == 1
        assert q == max(1, round(1.1 * L * W * H / (3 * l * w * h)))
    _, _, _, types = synth.parse_types(synth.multi_instance())
    assert len(types) == 20 and all(lf or wf or hf for (_, lf, _, wf, _, hf, _) in types)
    batch = synth.mixed_batch(32)
    ns = [len(synth.parse_types(t)[3]) for t in batch]
    assert all(3 <= n <= 100 for n in ns) and len(set(ns)) > 10
    _, _, _, types = synth.parse_types(synth.large_instance())
    assert len(types) == 100 and sum(t[6] for t in types) == 600
    assert synth.mixed_batch(16) == [open(golden_path("synth", f"mixed_{i:03d}.txt")).read()
                                     for i in range(16)]
