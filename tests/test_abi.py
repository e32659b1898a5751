"""The C-ABI library loads and exports every symbol include/gpm.h declares
(no compute calls: no GPU needed)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "gpm.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(gpm_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_core_api():
    names = declared_functions()
    for must in ("gpm_graph_create", "gpm_tc", "gpm_clique", "gpm_pattern",
                 "gpm_match_batch", "gpm_last_error", "gpm_mlp_predict"):
        assert must in names


def test_library_exports_every_declared_symbol(gpm):
    lib = gpm.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert missing == []


def test_error_path_without_gpu(gpm):
    # Argument errors are reported through status + message, not aborts.
    lib = gpm.lib()
    st = lib.gpm_csr_orient(ctypes.c_uint32(0), None, None, None, None, None)
    assert st == 1
    assert b"null" in lib.gpm_last_error()


def test_noalloc_limit_rejected_without_gpu(gpm):
    # The edge-count limit of the no-alloc triangle count is checked before
    # anything touches a device.
    lib = gpm.lib()
    st = lib.gpm_tc_edge_noalloc(ctypes.c_void_p(16), ctypes.c_void_p(16), ctypes.c_uint32(10),
                                 ctypes.c_uint64((16 << 20) + 1), ctypes.c_void_p(16), None)
    assert st == 1
    assert b"no-alloc limit" in lib.gpm_last_error()
    st = lib.gpm_tc_edge_noalloc(None, None, ctypes.c_uint32(0), ctypes.c_uint64(0), None, None)
    assert st == 1
