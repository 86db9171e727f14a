"""The C ABI library loads and exports every symbol include/airhip.h declares;
host-side argument validation (no GPU needed)."""
import os
import re

import numpy as np
import pytest
import scipy.sparse as sp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "airhip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(air_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import ctypes
    from airtoolsii_amd import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    assert sorted(_lib.EXPORTS) == syms


def test_last_error_is_a_string():
    from airtoolsii_amd import _lib
    L = _lib.load()
    assert isinstance(L.air_last_error(), bytes)


def test_calczeta_entry_point_matches_oracle(oracle):
    from airtoolsii_amd import calczeta
    z = calczeta(60)
    np.testing.assert_array_equal(z, oracle.calczeta(np.arange(2, 60)))
    from airtoolsii_amd._lib import AirError
    with pytest.raises(AirError):
        calczeta(2)


def _tiny():
    return sp.random(12, 9, density=0.5, random_state=1, format="csr")


@pytest.mark.parametrize("kwargs,msg", [
    (dict(K=[3, 2]), "strictly increasing"),
    (dict(K=[]), "ALWAYS"),
    (dict(b=np.ones(5)), "do not match"),
    (dict(x0=np.ones(4)), "x0 does not match"),
    (dict(options={"stoprule": {"type": "DP"}}), "taudelta"),
    (dict(options={"stoprule": {"type": "XX"}}), "not valid"),
    (dict(options={"lbound": np.ones(3)}), "lbound"),
    (dict(options={"relaxpar": "foo"}), "not valid"),
    (dict(method="kaczmarz"), "not defined"),
])
def test_sirt_argument_errors_before_any_device_work(kwargs, msg):
    from airtoolsii_amd import sirt
    args = dict(method="cimmino", A=_tiny(), b=np.ones(12), K=5, x0=None, options=None)
    args.update(kwargs)
    with pytest.raises(ValueError, match=msg):
        sirt(args["method"], args["A"], args["b"], args["K"], args["x0"], args["options"])


def test_ptr_keeps_converted_arrays_alive():
    """ptr(np.ascontiguousarray(a, dtype)) of an array that needs converting
    points at a temporary copy: the pointer object must keep it alive."""
    import ctypes as C
    import gc
    from airtoolsii_amd import _lib
    p = _lib.ptr(np.ascontiguousarray(np.arange(1000, dtype=np.int32), np.int64))
    junk = [np.full(1000, -7, np.int64) for _ in range(50)]
    gc.collect()
    got = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_int64)), (1000,))
    assert np.array_equal(got, np.arange(1000)) and len(junk) == 50
