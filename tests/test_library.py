"""The C ABI library loads and exports exactly what include/csm.h declares (no GPU needed)."""
import ctypes
import os
import re
import subprocess

import util


def header_functions():
    text = open(util.HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.findall(r"\b(csm_[a-z_]+)\s*\(", text)


def test_header_and_binding_agree():
    ffi = util.ffi()
    assert header_functions() == ffi.EXPORTED


def test_library_exports_every_declared_symbol():
    util.ensure_built()
    lib = ctypes.CDLL(os.path.join(util.PKG, "lib", "libcsm.so"))
    for name in header_functions():
        assert hasattr(lib, name), name
    nm = subprocess.run(["nm", "-D", "--defined-only", os.path.join(util.PKG, "lib", "libcsm.so")],
                        capture_output=True, text=True).stdout
    exported = {l.split()[-1] for l in nm.splitlines() if " T " in l and l.split()[-1].startswith("csm_")}
    assert exported == set(header_functions())


def test_session_without_gpu_fails_cleanly():
    import pytest
    ffi = util.ffi()
    try:
        s = ffi.Session(device=0)
    except ffi.CsmError as e:
        assert e.status == 5  # CSM_ERR_DEVICE
        return
    s.close()
    pytest.skip("a GPU is present")


def test_csmg_reports_errors_cleanly(tmp_path):
    util.ensure_built()
    r = subprocess.run([util.CSMG, "--query", str(tmp_path / "missing"), "--data", "x", "--update", "y"],
                       capture_output=True, text=True)
    assert r.returncode == 1
    assert "error (" in r.stdout


def test_csmg_prints_the_semi_bfs_threshold_in_use(tmp_path):
    """csmg prints the reference's settings block (main.cpp:22) with the threshold the session will
    use: the header's default, or --semi_bfs_threshold."""
    util.ensure_built()
    text = open(util.HEADER).read()
    default = int(re.search(r"#define CSM_DEFAULT_SEMI_BFS_THRESHOLD (\d+)u", text).group(1))
    for extra, want in (([], default), (["--semi_bfs_threshold", "4096"], 4096)):
        r = subprocess.run([util.CSMG, "--query", str(tmp_path / "missing"), "--data", "x", "--update", "y"] + extra,
                           capture_output=True, text=True)
        assert f"MIN_NUM_RESULTS_TO_GPU: {want}\n" in r.stdout


def header_struct_fields(name):
    """The field names of `typedef struct { ... } name;` in include/csm.h, in order."""
    text = re.sub(r"/\*.*?\*/", "", open(util.HEADER).read(), flags=re.S)
    m = re.search(r"typedef struct \{([^}]*)\}\s*" + name + r"\s*;", text)
    assert m, name
    fields = []
    for decl in m.group(1).split(";"):
        decl = decl.strip()
        if decl:  # "<type> a" or "<type> a, b, c"
            fields += [x.strip() for x in decl.split(None, 1)[1].split(",")]
    return fields


def test_ctypes_structs_match_the_header():
    """The ctypes mirrors of the ABI's structs have the header's fields in the header's order (a
    missing or reordered field would shift every later one)."""
    ffi = util.ffi()
    for name, cls in [("csm_options", ffi.Options), ("csm_technique", ffi.Technique), ("csm_query_result", ffi.QueryResult),
                      ("csm_batch_info", ffi.BatchInfo), ("csm_tunables", ffi.Tunables)]:
        assert header_struct_fields(name) == [f[0] for f in cls._fields_], name
    assert header_struct_fields("csm_tunables") == list(ffi.TUNABLES)
    assert ctypes.sizeof(ffi.Tunables) == 4 * len(ffi.TUNABLES)
