"""Host-side logic of the C++ interface layer: the .ini rules of the
reference's macros and the track file format.  CPU only (no GPU call)."""
import os
import subprocess

import numpy as np
import pytest

from supertrack_amd import trackfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.environ.get("SUPERTRACK_AMD_CLI") or os.path.join(ROOT, "supertrack_amd", "supertrack")


def ini_dump(tmp_path, text, section, name):
    p = tmp_path / "m.ini"
    p.write_bytes(text.encode() if isinstance(text, str) else text)
    out = subprocess.run([CLI, "--ini-dump", str(p), section, name], capture_output=True, text=True, check=True)
    d = {}
    for line in out.stdout.splitlines():
        k, v = line.split("=", 1)
        d[k] = v
    return d


def test_reference_macro_keys(tmp_path):
    # macros/default.ini style: inline comments after white space, case-insensitive lookups
    text = """; Example macro file
[Run]
Type = Folder
Path = /some/where/
FirstFile = -1 ; #-1 is the default case
Threads = 10
Oversamples = 10

[VoxelConstrainedSphere]
ScoringRegionHalfLength = 2.5e3    ; side length in nanometers
ScoringSphereDiameter = 10        ; sphere diameter
ShiftTrack = false\t\t\t\t   ; turn off track shifting

[Histogram]
Type = lin 
BinLower = 0.01
"""
    d = ini_dump(tmp_path, text, "VoxelConstrainedSphere", "ScoringRegionHalfLength")
    assert d["get"] == "[2.5e3]" and float(d["real"]) == 2500.0
    assert d["integer"] == "2"          # strtol stops at '.'
    d = ini_dump(tmp_path, text, "voxelconstrainedsphere", "SHIFTTRACK")
    assert d["get"] == "[false]" and d["bool"] == "00"
    d = ini_dump(tmp_path, text, "Run", "FirstFile")
    assert d["get"] == "[-1]" and d["integer"] == "-1"
    d = ini_dump(tmp_path, text, "Histogram", "Type")
    assert d["get"] == "[lin]"           # right-stripped
    d = ini_dump(tmp_path, text, "Histogram", "BinLower")
    assert float(d["float"]) == float(np.float32(0.01))
    d = ini_dump(tmp_path, text, "Run", "Missing")
    assert d["get"] == "[<default>]" and d["integer"] == "-999" and d["bool"] == "10"


def test_ini_rules(tmp_path):
    text = """# hash comment
[A]
hexval = 0x10
oct = 010
hash_in_value = a#b ; comment
no_space;comment = 5
colon : 7
dup = one
dup = two
multi = first
  second ; trailing comment
yes = YES
on = on
bad
[B
"""
    assert ini_dump(tmp_path, text, "A", "hexval")["integer"] == "16"
    assert ini_dump(tmp_path, text, "A", "oct")["integer"] == "8"
    assert ini_dump(tmp_path, text, "A", "hash_in_value")["get"] == "[a#b]"
    assert ini_dump(tmp_path, text, "A", "no_space;comment")["get"] == "[5]"
    assert ini_dump(tmp_path, text, "A", "colon")["get"] == "[7]"
    assert ini_dump(tmp_path, text, "A", "dup")["get"] == "[one\\ntwo]"
    assert ini_dump(tmp_path, text, "A", "multi")["get"] == "[first\\nsecond]"
    assert ini_dump(tmp_path, text, "A", "yes")["bool"] == "11"
    assert ini_dump(tmp_path, text, "A", "on")["bool"] == "11"
    # first error line: "bad" (line 14)
    assert ini_dump(tmp_path, text, "A", "yes")["error"] == "14"


def test_ini_missing_file(tmp_path):
    out = subprocess.run([CLI, "--ini-dump", str(tmp_path / "nope.ini"), "A", "b"], capture_output=True, text=True)
    assert "error=-1" in out.stdout


def test_ini_long_lines_split_like_fgets(tmp_path):
    # a 250-character value is read in pieces of 199 characters
    val = "x" * 250
    d = ini_dump(tmp_path, f"[S]\nk = {val}\n", "S", "k")
    assert d["get"] == "[" + ("k = " + val)[4:199] + "]"


def test_trackfile_roundtrip(tmp_path):
    rng = np.random.default_rng(0)
    off = np.array([0, 3, 3, 10], np.int64)
    cols = [rng.normal(size=10) for _ in range(4)]
    p = str(tmp_path / "a.sttrk")
    trackfile.write(p, *cols, off)
    x, y, z, e, o = trackfile.read(p)
    np.testing.assert_array_equal(o, off)
    for a, b in zip((x, y, z, e), cols):
        np.testing.assert_array_equal(a, b)
