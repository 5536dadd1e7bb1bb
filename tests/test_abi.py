"""The C ABI library without a GPU: it loads, exports every function the headers declare,
and its host-only pieces (status strings, transform operators) are right.  No call here
touches the device."""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

import crm_hip
import oracle_lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = sorted(glob.glob(os.path.join(ROOT, "include", "*.h")))


def _declared():
    names = set()
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names |= set(re.findall(r"\b(crm_(?:hip|mlp|hist)_\w+)\s*\(", text))
    return names


def test_headers_declare_the_abi():
    import crm_glue
    import crm_history
    import crm_mlp
    names = _declared()
    bound = (set(crm_hip.ABI_SYMBOLS) | set(crm_hip.DEBUG_SYMBOLS) | set(crm_mlp.SYMBOLS) | set(crm_glue.SYMBOLS)
             | set(crm_history.SYMBOLS))
    assert bound == names


def test_library_exports_every_declared_symbol():
    lib = crm_hip.load()
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing


def test_product_path_fails_loudly_without_library(monkeypatch, tmp_path):
    monkeypatch.setattr(crm_hip, "_lib", None)
    monkeypatch.setattr(crm_hip, "LIB_PATH", str(tmp_path / "absent.so"))
    with pytest.raises(RuntimeError, match="missing"):
        crm_hip.CrmHip(0)


def test_status_strings():
    seen = set()
    for code in range(11):
        s = crm_hip.errstr(code)
        assert s and s != "unknown error"
        seen.add(s)
    assert len(seen) == 11
    assert crm_hip.errstr(99) == "unknown error"
    assert "16" in crm_hip.errstr(crm_hip.CRM_ERR_CFL)


def test_kernel_names():
    lib = crm_hip.load()
    assert [lib.crm_hip_kernel_name(i).decode() for i in range(len(crm_hip.KERNELS))] == \
        crm_hip.KERNELS
    assert lib.crm_hip_kernel_name(99).decode() == "unknown"


def test_null_arguments_rejected_before_device_work():
    lib = crm_hip.load()
    assert lib.crm_hip_run(None, None, None, 0.0, None, None, None, None) == crm_hip.CRM_ERR_ARG
    assert lib.crm_hip_get_stats(None, None) == crm_hip.CRM_ERR_ARG
    assert lib.crm_hip_finalize(None) == crm_hip.CRM_OK


def _slot(n, s):
    return 0 if s == 0 else (n if s == 1 else s)


@pytest.mark.parametrize("n", [8, 16, 32, 64])
def test_transform_operators_reproduce_reference_transform(n):
    """The dense operators the pressure kernel multiplies by are the reference's own
    transform (fft.F90:5-153, restated bit-exactly by the oracle) applied to unit vectors;
    on any row they agree with it to FP64 rounding, forward and inverse."""
    Ft, Gt = crm_hip.fft991_operators(n)
    lib = oracle_lib.oracle().lib
    dp = C.POINTER(C.c_double)
    lib.oracle_fft991.argtypes = [dp, C.c_int, C.c_int]
    rng = np.random.default_rng(n)
    for _ in range(10):
        x = rng.standard_normal(n)
        a = np.zeros(n + 2)
        a[:n] = x
        lib.oracle_fft991(a.ctypes.data_as(dp), n, -1)
        coef = np.array([a[_slot(n, s)] for s in range(n)])
        assert np.abs(x @ Ft - coef).max() <= 1e-14 * np.abs(coef).max()
        b = np.zeros(n + 2)
        for s in range(n):
            b[_slot(n, s)] = coef[s]
        lib.oracle_fft991(b.ctypes.data_as(dp), n, 1)
        assert np.abs(coef @ Gt - b[:n]).max() <= 1e-14 * np.abs(b[:n]).max()
    # forward then inverse is the identity to the transform's single-precision twiddles
    assert np.abs(Ft @ Gt - np.eye(n)).max() < 1e-6
    # the slots that the inverse ignores (b0, b_{n/2}) are not carried: n slots for n values
    assert Ft.shape == (n, n)
