"""C-ABI library and host logic checks that need no GPU."""

import ctypes
import os
import re

import pytest
import torch

from conftest import LIB, ROOT

HEADER = os.path.join(ROOT, "include", "laker_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(laker_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    syms = declared_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, f"missing exports: {missing}"


def test_binding_covers_header():
    import laker._native as N

    assert set(N.EXPORTED) == set(declared_symbols())


def test_version_and_error_channel():
    import laker._native as N

    assert b"laker-hip" in N.lib.laker_version()
    # invalid arguments are rejected on the host before any device work
    rc = N.lib.laker_operator_create(None, 0, 4, 0, 0.1, 1.0, 0, None, ctypes.byref(ctypes.c_void_p()))
    assert rc == N.LAKER_EINVAL
    assert "NULL" in N.last_error() or "row" in N.last_error()
    with pytest.raises(ValueError):
        N.check(rc)
    with pytest.raises(RuntimeError):
        N.check(N.LAKER_EBREAKDOWN)


def test_regressor_validation_errors():
    import laker

    with pytest.raises(ValueError):
        laker.LAKERRegressor(embedding_dim=0)
    with pytest.raises(ValueError):
        laker.LAKERRegressor(lambda_reg=0.0)
    with pytest.raises(ValueError):
        laker.LAKERRegressor(kernel_approx="bogus")
    with pytest.raises(ValueError):
        laker.LAKERRegressor(precision="fp16")
    m = laker.LAKERRegressor(verbose=False)
    with pytest.raises(RuntimeError, match="not been fitted"):
        m.predict(torch.zeros(3, 2))
    m.set_params(lambda_reg=0.5, pcg_tol=1e-8)
    assert m.lambda_reg == 0.5 and m.core.pcg_tol == 1e-8
    with pytest.raises(ValueError):
        m.set_params(nonsense=1)
    p = m.get_params()
    assert p["lambda_reg"] == 0.5 and p["precision"] == "fp32"


def test_host_rules():
    from laker.core import auto_chunk
    from laker.utils import adaptive_shrinkage_rho, default_num_probes

    assert auto_chunk(2000, None) is None
    assert auto_chunk(100000, None) == 8192
    assert auto_chunk(50000, None) == 5000
    assert auto_chunk(20000, None) == 2000
    assert adaptive_shrinkage_rho(200, 2000, 0.1) == 0.5
    assert default_num_probes(100000) == 632


def test_pathloss_problem_deterministic():
    import laker

    x1, y1 = laker.pathloss_problem(500, seed=7)
    x2, y2 = laker.pathloss_problem(500, seed=7)
    assert torch.equal(x1, x2) and torch.equal(y1, y2)
    assert abs(float(y1.mean())) < 1e-5 and abs(float(y1.std()) - 1.0) < 1e-5
    _, yk = laker.pathloss_problem(500, seed=7, k=3)
    assert yk.shape == (500, 3) and torch.equal(yk[:, 0], y1)
