"""Host-side encoding of the surrogate's velocity targets (cism_amd/surrogate.py)."""
import numpy as np
import pytest

from cism_amd import surrogate as SG


@pytest.mark.parametrize("target", ["linear", "log1p"])
def test_encode_decode_inverse(target):
    v = np.array([-9000.0, -100.0, -1e-3, 0.0, 2e-3, 50.0, 100.0, 12000.0])
    t = SG.encode(v, 100.0, target)
    np.testing.assert_allclose(SG.decode(t, 100.0, target), v, rtol=1e-12, atol=1e-12)
    assert np.all(np.diff(t) > 0)                     # monotone
    assert np.all(np.sign(t) == np.sign(v))


def test_log1p_compresses_and_linear_scales():
    v = np.array([100.0, 10000.0])
    assert np.allclose(SG.encode(v, 100.0, "linear"), [1.0, 100.0])
    t = SG.encode(v, 100.0, "log1p")
    assert t[0] == pytest.approx(np.log(2.0)) and t[1] == pytest.approx(np.log(101.0))
    with pytest.raises(ValueError):
        SG.encode(v, 100.0, "sqrt")
