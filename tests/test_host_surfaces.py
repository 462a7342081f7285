"""Config parser, subset parser and MP5 book writer of the C++ host against
the reference's own conf.cpp / io.cpp (compiled from the reference sources
into oracle/_ref where the reference tree is present) and documented behaviour."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOOL = os.path.join(ROOT, "empi_amd", "empi-host-tool")
REF = os.path.join(ROOT, "oracle", "_ref")

pytestmark = pytest.mark.skipif(not os.path.exists(TOOL), reason="host tool not built")

CONFIGS = [
    "energyError 0.01\nmaximalNumberOfIterations 50\nenergyPercent 99.0\n\nMP SMP\n",
    "# comment\n   \t# indented comment\nkey\tvalue with  spaces   \n  leading 1\nnovalue\nnovalue2   \n",
    "a 1\na 2\nb\t\t  3 4\t\n#c 5\nc 6 # not a comment\n",
    "MP\tmmp3\r\nsamplingFrequency 128.0\r\n",
]
KEYS = ["energyError", "maximalNumberOfIterations", "energyPercent", "MP", "key", "leading", "novalue",
        "novalue2", "a", "b", "c", "samplingFrequency", "#c"]


def run(cmd):
    return subprocess.run(cmd, capture_output=True, text=True)


@pytest.mark.parametrize("text", CONFIGS)
def test_config_parser_matches_reference(tmp_path, text):
    f = tmp_path / "x.cfg"
    f.write_bytes(text.encode())
    mine = run([TOOL, "conf", str(f)]).stdout.splitlines()
    got = {line.split("\t", 1)[0]: line.split("\t", 1)[1] for line in mine}
    if os.path.exists(os.path.join(REF, "conf_driver")):
        ref = run([os.path.join(REF, "conf_driver"), str(f)] + KEYS).stdout.splitlines()
        want = {line.split("\t", 1)[0]: line.split("\t", 1)[1] for line in ref}
        assert {k: v for k, v in got.items() if k in KEYS} == want
    assert all(not k.startswith("#") for k in got)


def test_config_missing_file(tmp_path):
    r = run([TOOL, "conf", str(tmp_path / "nope.cfg")])
    assert r.returncode == 1 and r.stdout.strip() == "ERROR\tcouldNotOpenConfigFile"


@pytest.mark.parametrize("text,mx,want", [("1", 3, "1"), ("1-3", 3, "1 2 3"), ("1-2,5,8-10", 10, "1 2 5 8 9 10"),
                                         ("3,1", 3, "3 1")])
def test_subset_parser(text, mx, want):
    assert run([TOOL, "subset", text, str(mx)]).stdout.split() == want.split()


@pytest.mark.parametrize("text,mx", [("0", 3), ("4", 3), ("1-4", 3), ("x", 3), ("", 3)])
def test_subset_parser_errors(text, mx):
    r = run([TOOL, "subset", text, str(mx)])
    assert r.returncode == 1 and r.stdout.startswith("ERROR\t")


def _description(path, fs, C, N, epochs, seed):
    rng = np.random.default_rng(seed)
    lines = [f"{fs!r} {C} {N} {epochs}"]
    truth = []
    for _ in range(epochs):
        sig = rng.standard_normal((C, N)).astype(np.float32)
        for c in range(C):
            lines.append(" ".join(repr(float(v)) for v in sig[c]))
        atoms = []
        for c in range(C):
            n = int(rng.integers(0, 5))
            p = rng.standard_normal((n, 6)) * 100
            atoms.append(p)
            lines.append(str(n))
            lines.extend(" ".join(repr(float(v)) for v in row) for row in p)
        truth.append((sig, atoms))
    path.write_text("\n".join(lines) + "\n")
    return truth


@pytest.mark.parametrize("fs,C,N,epochs", [(128.0, 1, 16, 1), (256.0, 3, 7, 2), (500.0, 2, 1, 3)])
def test_book_bytes_match_reference(tmp_path, fs, C, N, epochs):
    desc = tmp_path / "d.txt"
    truth = _description(desc, fs, C, N, epochs, C * 100 + N)
    mine = tmp_path / "mine.b"
    assert run([TOOL, "book", str(desc), str(mine)]).returncode == 0
    data = mine.read_bytes()
    assert data[:6] == b"MPv5.0"
    if os.path.exists(os.path.join(REF, "book_driver")):
        ref = tmp_path / "ref.b"
        assert run([os.path.join(REF, "book_driver"), str(desc), str(ref)]).returncode == 0
        assert data == ref.read_bytes()
    # round trip through the host reader
    out = run([TOOL, "readbook", str(mine)]).stdout.splitlines()
    assert out[0].startswith(f"fs {fs:.9g} channels {C} epochs {epochs}")
    rows = [line for line in out if not line.startswith(("fs", "epoch"))]
    want = []
    for sig, atoms in truth:
        for c, p in enumerate(atoms):
            for row in p:
                want.append([c + 1] + [float(np.float32(v)) for v in row])
    got = [[int(r.split()[0])] + [float(np.float32(v)) for v in r.split()[1:]] for r in rows]
    assert got == want


# ------------------------------------------------------------ signal readers
# src/io.cpp:19-85 and src/empi.cpp:82-98: whole signal, every epoch (a
# partial last epoch throws fileIsTruncated after the full ones), selected
# epochs in the listed order (a listed epoch past the end throws).

READ_CASES = [
    # (mode, args, epochs expected as 0-based sample ranges, error or None)
    ("whole", [], [(0, 23)], None),
    ("all", ["5"], [(0, 5), (5, 10), (10, 15), (15, 20)], "fileIsTruncated"),
    ("all", ["23"], [(0, 23)], None),
    ("all", ["1"], [(i, i + 1) for i in range(23)], None),
    ("selected", ["5", "3,1"], [(10, 15), (0, 5)], None),
    ("selected", ["5", "2,2,4"], [(5, 10), (5, 10), (15, 20)], None),
    ("selected", ["5", "4,5"], [(15, 20)], "fileIsTruncated"),
    ("selected", ["23", "1"], [(0, 23)], None),
]


def _expected(x, sel, ranges, error):
    lines = []
    for k, (a, b) in enumerate(ranges):
        lines.append(f"epoch {k + 1} samples {b - a}")
        for c in sel:
            lines.append(" ".join(f"{float(v):.9g}" for v in x[a:b, c - 1]))
    if error:
        lines.append(f"ERROR\t{error}")
    return lines


@pytest.mark.parametrize("mode,args,ranges,error", READ_CASES)
def test_signal_readers_match_reference(tmp_path, mode, args, ranges, error):
    rng = np.random.default_rng(len(args) * 7 + len(ranges))
    x = rng.standard_normal((23, 4)).astype(np.float32)        # [t][channel], 4 channels
    f = tmp_path / "sig.bin"
    x.tofile(f)
    sel = [4, 1, 3]
    cmd = [str(f), "4", ",".join(map(str, sel)), mode] + args
    mine = run([TOOL, "read"] + cmd)
    assert mine.stdout.splitlines() == _expected(x, sel, ranges, error)
    assert mine.returncode == (1 if error else 0)
    driver = os.path.join(REF, "reader_driver")
    if os.path.exists(driver):
        ref = run([driver] + cmd)
        assert mine.stdout == ref.stdout and mine.returncode == ref.returncode


def test_signal_reader_missing_file(tmp_path):
    r = run([TOOL, "read", str(tmp_path / "none.bin"), "1", "1", "whole"])
    assert r.returncode == 1 and r.stdout.strip() == "ERROR\tcouldNotOpenSignalFile"
