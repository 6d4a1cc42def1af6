"""The C-ABI library loads without a GPU and exports every symbol that
include/jfem_hip.h declares.  No compute calls."""
import ctypes
import os
import subprocess

import pytest


def test_library_exports_declared_symbols():
    from jfem_amd import _native as N
    L = N.lib()
    declared = N.declared_symbols()
    assert len(declared) >= 30
    for name in declared:
        assert hasattr(L, name), name
        assert name in N._SIGS, f"{name} has no ctypes signature"


def test_header_is_plain_c():
    """The boundary header compiles as C (no torch or C++ types in signatures)."""
    from jfem_amd import _native as N
    r = subprocess.run(["gcc", "-fsyntax-only", "-x", "c", "-std=c99", N.HEADER], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_last_error_without_device():
    from jfem_amd import _native as N
    assert isinstance(N.lib().jfem_last_error(), bytes)


def test_native_library_missing_fails_loudly(tmp_path, monkeypatch):
    from jfem_amd import _native as N
    monkeypatch.setattr(N, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(N, "_lib", None)
    with pytest.raises(N.NativeLibraryError):
        N.lib()


def test_workload_generator_is_deterministic():
    from jfem_amd.workloads import SplitMix64, sinusoid_terms
    a = sinusoid_terms(2, 16, 3)
    b = sinusoid_terms(2, 16, 3)
    assert (a[0] == b[0]).all() and (a[1] == b[1]).all()
    assert a[0].min() >= 1 and a[0].max() <= 8
    # SplitMix64 reference value for seed 0 (first output of the published generator)
    assert SplitMix64(0).next_u64() == 0xE220A8397B1DCDAF


def test_oracle_splitmix_matches_python():
    from oracle import oracle as O
    from jfem_amd.workloads import SplitMix64
    st = ctypes.c_uint64(12345)
    s = SplitMix64(12345)
    for _ in range(5):
        assert O.lib().orc_splitmix64(ctypes.byref(st)) == s.next_u64()


def test_cell_arrays_are_checked_against_the_grid():
    """per-cell coefficient arrays must have one entry per finest cell: the C ABI
    reads exactly that many, so the mirror rejects other lengths first"""
    import numpy as np
    from jfem_amd import api, _native as N

    class _Space:
        cells = np.array([4, 3], dtype=np.int32)

    assert api._cell_array(None, _Space(), "c") is None
    assert api._cell_array(np.ones((3, 4)), _Space(), "c").shape == (12,)
    with pytest.raises(N.IllegalArgumentException):
        api._cell_array(np.ones(16), _Space(), "c")


def test_mixed_integral_kinds_match_header():
    """the Python mirror passes the JFEM_MIXED_* kinds the header declares; other names are refused"""
    import re

    from jfem_amd import mixed as M, _native as N
    hdr = open(os.path.join(os.path.dirname(__file__), "..", "include", "jfem_hip.h")).read()
    kinds = {m.group(1): int(m.group(2)) for m in re.finditer(r"#define JFEM_MIXED_(\w+)\s+(\d+)", hdr)}
    assert kinds == {"VELOCITY_GRAD_GRAD": 0, "DIV_VALUE": 1, "VALUE_GRAD": 2}
    vi = M.TPVectorCellIntegral(1.0, M.TPVectorCellIntegral.GRAD_GRAD)
    assert M.MixedCellIntegral.fromVelocityIntegral(vi).kind == kinds["VELOCITY_GRAD_GRAD"]
    assert M.MixedTPCellIntegral(M.MixedTPCellIntegral.DIV_VALUE).kind == kinds["DIV_VALUE"]
    assert M.MixedTPCellIntegral(-2.0, M.MixedTPCellIntegral.VALUE_GRAD).kind == kinds["VALUE_GRAD"]
    with pytest.raises(N.UnsupportedOperationException):
        M.MixedTPCellIntegral("GradValue")


def test_csr_import_argument_errors_without_device():
    """jfem_mat_from_csr refuses a row pointer that does not start at 0 or is not monotone with
    JFEM_ERR_ARG before it touches the context or the device (the context pointer here is a
    dummy that is never dereferenced); the Python mirror refuses CSR arrays shorter than
    rowptr[-1] before calling the library"""
    import numpy as np
    from jfem_amd import api, _native as N
    L = N.lib()
    ctx_storage = ctypes.create_string_buffer(64)
    dummy_ctx = ctypes.c_void_p(ctypes.addressof(ctx_storage))
    cols = np.zeros(4, dtype=np.int32)
    vals = np.ones(4)
    out = ctypes.c_void_p()
    for rp in ([1, 2, 4], [-1, 2, 4], [0, 3, 2]):
        rp = np.array(rp, dtype=np.int64)
        rc = L.jfem_mat_from_csr(dummy_ctx, 2, 2, N.lptr(rp), N.iptr(cols), N.dptr(vals), ctypes.byref(out))
        assert rc == 1, (rp, rc)      # JFEM_ERR_ARG
        assert not out.value
    assert b"row pointer" in L.jfem_last_error()

    class _Ctx:
        h = dummy_ctx

    with pytest.raises(N.IllegalArgumentException):
        api.SparseMatrix.from_csr([0, 2, 5], [0, 1, 0, 1], [1.0, 2.0, 3.0, 4.0, 5.0], ctx=_Ctx())
    with pytest.raises(N.IllegalArgumentException):
        api.SparseMatrix.from_csr([0, 2, 4], [0, 1, 0, 1], [1.0, 2.0, 3.0], ctx=_Ctx())
