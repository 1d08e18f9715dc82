"""The oracle against the reference's own outputs frozen in
tests/golden/ref_kernels.npz (written by tests/golden/make_golden.py from
oracle/_ref, COSA's compiled routines): identical binary64 results on every
fixture.  Runs everywhere, also where /root/reference is absent."""
import ast
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

from make_golden import run, set_cfg  # noqa: E402
from cosa_amd.config import default_config  # noqa: E402
from oracle_bind import Oracle  # noqa: E402

PATH = os.path.join(HERE, "golden", "ref_kernels.npz")
DATA = np.load(PATH, allow_pickle=False)
NAMES = [str(n) for n in DATA["names"]]


@pytest.fixture(scope="module")
def oracle():
    return Oracle()


@pytest.mark.parametrize("name", NAMES)
def test_oracle_matches_reference_fixture(oracle, name):
    # fixtures store the configuration they were made with; fields added to
    # the configuration since then take their defaults
    cfg = default_config()
    cfg.update(dict(ast.literal_eval(bytes(DATA[f"{name}/cfg"]).decode())))
    inputs = {k.split("/")[-1]: DATA[k] for k in DATA.files if k.startswith(f"{name}/in/")}
    want = {k.split("/")[-1]: DATA[k] for k in DATA.files if k.startswith(f"{name}/out/")}
    set_cfg(oracle, cfg)
    got = run(oracle, None, name, cfg, inputs, oracle=True)
    assert set(got) == set(want)
    for k in want:
        np.testing.assert_array_equal(got[k], want[k], err_msg=f"{name}:{k}")
