"""The C oracle against the golden fixtures written from the reference's compiled modules."""
import pytest

from golden_util import NAMES, load, load_diags
from pyoracle import CpuOracle


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_golden_bitwise(name):
    import numpy as np

    cfg, grid, st_in, st_out, steps = load(name)
    o = CpuOracle(cfg, grid)
    o.start(st_in)
    o.run(0, steps)
    got = o.get_state(steps % 2)
    for k in st_out:
        assert np.array_equal(got[k], st_out[k]), k
    for k, v in load_diags(name).items():
        assert np.array_equal(o.get_diag(k, v.shape), v), k
