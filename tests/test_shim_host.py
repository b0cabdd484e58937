"""Reference-signature shim, host parts (no device): getIndex is the
reference's linear scan of the Green CDF (randomWalk/CPU/mainFunctions.cpp:
40-46): the panel i with intg[i] <= r < intg[i+1], -1 outside the table."""
import os
import subprocess

import numpy as np
import pytest

import oracle_ffi as O

DRIVER = os.path.join(os.path.dirname(__file__), "shim", "shim_driver")


def test_get_index_scans_the_cdf(tmp_path):
    if not os.path.exists(DRIVER):
        pytest.fail("tests/shim/shim_driver is not built (make -C tests/shim)")
    intg = O.green_cdf()
    rng = np.random.default_rng(5)
    r = np.concatenate([rng.uniform(0, 1, 2000), intg[:-1], intg[1:] - 1e-12, [-0.5, 1.0, 1.5]])
    (tmp_path / "cdf").write_text("\n".join(repr(float(v)) for v in intg))
    (tmp_path / "vals").write_text("\n".join(repr(float(v)) for v in r))
    out = subprocess.run([DRIVER, "getindex", str(tmp_path / "cdf"), str(tmp_path / "vals")],
                         capture_output=True, text=True, timeout=60, check=True)
    got = np.array([int(x) for x in out.stdout.split()])
    want = np.searchsorted(intg, r, side="right") - 1
    want[(r < intg[0]) | (r >= intg[-1])] = -1
    assert np.array_equal(got, want)
