"""CPU checks of the drop-in boundary: the C ABI library exports every entry
point its header declares (no compute calls), the ctypes table matches the
header, the dataset file format, and the command lines' argument and file
errors (exit status 1 before any device work, as TC.cu:223-233 and
TC.cu:62-66 behave)."""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "reset_amd.h")
BIN = os.path.join(ROOT, "reset_amd", "bin")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(rs_[a-z0-9_]+)\s*\(", text))
    return names


@pytest.fixture(scope="module")
def built_lib():
    from reset_amd import build, lib

    if not os.path.exists(lib.LIB_PATH):
        build.build_lib()
    return lib


@pytest.fixture(scope="module")
def built_tools(built_lib):
    from reset_amd import build

    if not all(os.path.exists(os.path.join(BIN, t)) for t in build.TOOLS):
        build.build_tools()
    return BIN


def test_header_entry_points_are_all_exported(built_lib):
    import ctypes

    declared = _declared()
    assert len(declared) > 40
    so = ctypes.CDLL(built_lib.LIB_PATH)
    missing = sorted(n for n in declared if not hasattr(so, n))
    assert missing == []


def test_ctypes_table_matches_header(built_lib):
    assert set(built_lib.SIGNATURES) == _declared()
    assert built_lib.missing_symbols() == []


def test_host_only_entry_points(built_lib):
    L = built_lib.load()
    assert L.rs_version().decode().startswith("reset_amd")
    assert L.rs_status_string(3).decode()
    # calculate_length is evaluated on the host: the closed form at 0.2 / 5
    for d, want in [(0, 8), (1, 8), (7, 8), (8, 16), (1000, 1024), (1024, 2048)]:
        assert L.rs_calculate_length(d, 0.2, 5) == want


def test_bin_format_round_trip(tmp_path):
    from reset_amd import binfmt

    v = np.array([3, -1, 7, 2**31 - 1], np.int32)
    p = tmp_path / "v.bin"
    binfmt.write_i32_vec(p, v)
    raw = p.read_bytes()
    assert raw[:8] == (4).to_bytes(8, "little") and len(raw) == 8 + 16
    assert raw[8:12] == (3).to_bytes(4, "little")
    assert np.array_equal(binfmt.read_i32_vec(p), v)
    binfmt.write_int(tmp_path / "s.bin", 123456)
    assert (tmp_path / "s.bin").read_bytes() == (123456).to_bytes(4, "little")
    assert binfmt.read_int(tmp_path / "s.bin") == 123456
    (tmp_path / "t.bin").write_bytes(raw[:12])
    with pytest.raises(ValueError):
        binfmt.read_i32_vec(tmp_path / "t.bin")


@pytest.mark.parametrize("tool", ["tc", "sss", "ir"])
def test_cli_argument_and_file_errors(built_tools, tmp_path, tool):
    exe = os.path.join(built_tools, tool)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "Usage" in r.stderr
    r = subprocess.run([exe, str(tmp_path), "--bogus=1"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "Unknown argument: --bogus=1" in r.stderr
    r = subprocess.run([exe, str(tmp_path), "--alpha=x"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "Bad value" in r.stderr
    r = subprocess.run([exe, str(tmp_path / "absent")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "Error opening file" in r.stderr


@pytest.mark.parametrize("tool,bad", [("tc", "cuckoo"), ("sss", "bogus"), ("ir", "roaring")])
def test_cli_unknown_method(built_tools, tmp_path, tool, bad):
    """--method= lists the tool's methods (tc and ir: original, cuco; sss:
    original, cuco, cuckoo, hopscotch, roaring); others exit 1 before any
    file or device is touched (baseline/tc/main.cu:55-70)."""
    exe = os.path.join(built_tools, tool)
    r = subprocess.run([exe, str(tmp_path), f"--method=original,{bad}"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1 and f"Unknown method: {bad}" in r.stderr


def test_intersection_cli_usage(built_tools):
    r = subprocess.run([os.path.join(built_tools, "intersection"), "5"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1 and "Usage" in r.stderr
