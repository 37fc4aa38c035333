"""CPU checks of the drop-in boundary: the C ABI library exports every entry
point its header declares (no compute calls), the ctypes table matches the
header, the dataset file format, and the command lines' argument and file
errors (exit status 1 before any device work, as TC.cu:223-233 and
TC.cu:62-66 behave)."""
import os
import re
import subprocess
import sys

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

    objs = [os.path.join(build.STANDALONE_DIR, f"{a}_hashtable.o") for a in build.STANDALONE_APPS]
    if not all(os.path.exists(os.path.join(BIN, t)) for t in build.TOOLS) or not all(map(os.path.exists, objs)):
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


def test_bin64_format_round_trip(tmp_path):
    """The 64-bit files of the multi-GPU drivers (multi-GPU/sss-multi-GPU/
    main.cu:25-57): uint64 count + int64 values, int64 scalar; the readers
    narrow to int32 and refuse a value that does not fit."""
    from reset_amd import binfmt

    v = np.array([0, -5, 2**31 - 1, -2**31], np.int64)
    p = tmp_path / "v64.bin"
    binfmt.write_i64_vec(p, v)
    raw = p.read_bytes()
    assert raw[:8] == (4).to_bytes(8, "little") and len(raw) == 8 + 32
    assert raw[16:24] == (-5).to_bytes(8, "little", signed=True)
    assert np.array_equal(binfmt.read_i64_vec(p), v)
    got = binfmt.read_vec_64(p)
    assert got.dtype == np.int32 and np.array_equal(got, v.astype(np.int32))
    binfmt.write_int64(tmp_path / "n64.bin", 4_000_000)
    assert (tmp_path / "n64.bin").read_bytes() == (4_000_000).to_bytes(8, "little")
    assert binfmt.read_num_64(tmp_path / "n64.bin") == 4_000_000
    binfmt.write_i64_vec(tmp_path / "big.bin", [1, 2**31])
    with pytest.raises(ValueError):
        binfmt.read_vec_64(tmp_path / "big.bin")
    (tmp_path / "t64.bin").write_bytes(raw[:20])
    with pytest.raises(ValueError):
        binfmt.read_i64_vec(tmp_path / "t64.bin")


def test_bin64_cpp_mirror_reads_and_writes_the_same_bytes(built_lib, tmp_path):
    """reset_amd.hpp's read_vec_64 / read_num_64 / write_vec_64 / write_num_64
    against the Python writer and reader (host code only, no device call)."""
    from reset_amd import binfmt

    src = tmp_path / "rt.cpp"
    src.write_text(r"""
#include "reset_amd.hpp"
#include <cstdio>
int main(int argc, char** argv) {
  std::vector<int> v = reset_amd::read_vec_64(argv[1]);
  int n = reset_amd::read_num_64(argv[2]);
  std::vector<int64_t> w;
  for (int x : v) w.push_back((int64_t)x * 2);
  reset_amd::write_vec_64(argv[3], w);
  reset_amd::write_num_64(argv[4], (int64_t)n + 1);
  try { reset_amd::read_vec_64(argv[5]); return 2; } catch (const std::out_of_range&) {}
  std::printf("%zu %d\n", v.size(), n);
  return 0;
}
""")
    exe = tmp_path / "rt"
    lib_dir = os.path.dirname(built_lib.LIB_PATH)
    cmd = ["g++", "-std=c++17", "-D__HIP_PLATFORM_AMD__", f"-I{ROOT}/include", "-I/opt/rocm/include", str(src),
           "-o", str(exe), f"-L{lib_dir}", "-lreset_amd", "-L/opt/rocm/lib", "-lamdhip64",
           f"-Wl,-rpath,{lib_dir}", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    binfmt.write_i64_vec(tmp_path / "a.bin", [5, -7, 2**31 - 1, 0])
    binfmt.write_int64(tmp_path / "n.bin", 77)
    binfmt.write_i64_vec(tmp_path / "bad.bin", [-2**31 - 1])
    r = subprocess.run([str(exe), str(tmp_path / "a.bin"), str(tmp_path / "n.bin"), str(tmp_path / "b.bin"),
                        str(tmp_path / "m.bin"), str(tmp_path / "bad.bin")], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.split() == ["4", "77"], r.stdout + r.stderr
    assert np.array_equal(binfmt.read_i64_vec(tmp_path / "b.bin"), [10, -14, 2 * (2**31 - 1), 0])
    assert np.frombuffer((tmp_path / "m.bin").read_bytes(), "<i8")[0] == 78


@pytest.mark.parametrize("app,inputs", [
    ("sss", "std::vector<int> vertexs, csr_cols, csr_offsets; int num_nodes, num_edges;"),
    ("ir", "std::vector<int> inverted_index_offsets, inverted_index; int inverted_index_num;"),
    ("tc", "std::vector<int> csr_row, csr_cols; int num_nodes, num_edges;")])
def test_standalone_objects_link_against_an_app_main(built_tools, tmp_path, app, inputs):
    """The free-function form (make_hash_table over the app's globals): an
    app main that defines the reference's input globals links with
    reset_amd/standalone/<app>_hashtable.o, which defines the outputs."""
    macro = {"tc": "RESET_AMD_APP_TC", "sss": "RESET_AMD_APP_SSS", "ir": "RESET_AMD_APP_IR"}[app]
    src = tmp_path / "main.cpp"
    src.write_text(f"#define {macro}\n#include \"reset_amd_standalone.hpp\"\n{inputs}\n"
                   "int main(int argc, char**) { if (argc > 5) { make_hash_table(0.2f, 5); free_hash_table(); }\n"
                   "  return reset_amd_standalone_tables() == nullptr && max_length == 0 ? 0 : 1; }\n")
    lib_dir = os.path.join(ROOT, "reset_amd")
    obj = os.path.join(lib_dir, "standalone", f"{app}_hashtable.o")
    exe = tmp_path / "app"
    r = subprocess.run(["g++", "-std=c++17", "-D__HIP_PLATFORM_AMD__", f"-I{ROOT}/include", "-I/opt/rocm/include",
                        str(src), obj, "-o", str(exe), f"-L{lib_dir}", "-lreset_amd", "-L/opt/rocm/lib",
                        "-lamdhip64", f"-Wl,-rpath,{lib_dir}", "-Wl,-rpath,/opt/rocm/lib"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    assert subprocess.run([str(exe)]).returncode == 0


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


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus 2 exits with status 2 and says so,
    before touching a device."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 2
    assert "exactly one GPU" in (r.stdout + r.stderr)
