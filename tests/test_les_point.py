"""The hand-derived vector-Jacobian products of the pointwise subgrid models
(diablo3_test_amd/csrc/les_point.h), used by the adjoint substep, against
central finite differences of the same point models on random inputs and
random per-plane constants (tests/les_point_check.cpp, built with g++).
Directions whose AMD clip max(0, .) switches inside the difference stencil
are skipped (the derivative is one-sided there)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_point_model_vjps_match_finite_differences(tmp_path):
    exe = str(tmp_path / "les_point_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", exe, os.path.join(HERE, "les_point_check.cpp")], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, check=True, timeout=120).stdout
    rows = [ln.split() for ln in out.strip().splitlines()]
    assert [(r[0], r[1]) for r in rows] == [("vel", "1"), ("vel", "4"), ("vel", "5"), ("th", "5")]
    for model, typ, err, trials in rows:
        assert int(trials) >= 300, (model, typ)
        assert float(err) < 1e-5, (model, typ, err)
