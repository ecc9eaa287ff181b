"""The present-pass restatement (oracle/present.py) against the reference's own
vectors (packages/shared-samplers/src/__tests__/tonemap.test.ts:17-136) and
the closed forms of the operators."""
import numpy as np
import pytest

from oracle import present as P


def test_none_and_linear_exposure_vectors():
    assert P.apply_tonemap([0.5] * 3, "none", 2).tolist() == [1, 1, 1]
    assert P.apply_tonemap([0.6] * 3, "none", 2).tolist() == pytest.approx([1.2] * 3)
    assert P.apply_tonemap([0.6] * 3, "linear", 2).tolist() == [1, 1, 1]


def test_hdr_kept_and_saturated_only_at_finite_f32():
    assert P.apply_tonemap([65504, 32752, -65504], "none", 2).tolist() == [131008, 65504, -131008]
    out = P.apply_tonemap([2, 1, 0.5], "none", P.F32_MAX)
    assert out.tolist() == [P.F32_MAX, P.F32_MAX, P.F32_MAX * 0.5]
    for mode in ("aces", "agx", "reinhard", "linear"):
        assert np.isfinite(P.apply_tonemap([2, 1, 0.5], mode, P.F32_MAX)).all()
    with pytest.raises(ValueError):
        P.apply_tonemap([1, 1, 1], "none", -1)


def test_operator_vectors():
    assert float(P.aces(0)) == pytest.approx(0, abs=1e-6)
    assert float(P.aces(1e6)) <= 1
    assert float(P.aces(0.5)) > float(P.aces(0.2))
    assert float(P.reinhard(0)) == 0
    assert float(P.reinhard(1)) == pytest.approx(0.5)
    assert float(P.reinhard(3)) == pytest.approx(0.75)
    assert float(P.reinhard(-0.5)) == 0
    assert float(P.agx(0)) >= 0 and float(P.agx(1e6)) <= 1 and float(P.agx(1)) > float(P.agx(0.05))
    assert P.TONEMAP_MODE_INDEX == {"aces": 0, "agx": 1, "reinhard": 2, "linear": 3, "none": 4}


def test_aces_is_the_narkowicz_rational_and_monotone():
    x = np.linspace(0.001, 50, 20000)
    y = P.aces(x)
    assert (np.diff(y) >= 0).all() and y.max() <= 1
    np.testing.assert_allclose(y, np.minimum(1, x * (2.51 * x + 0.03) / (x * (2.43 * x + 0.59) + 0.14)))


def test_srgb_oetf_vectors():
    assert float(P.linear_to_srgb(0)) == pytest.approx(0, abs=1e-6)
    assert float(P.linear_to_srgb(1)) == pytest.approx(1, abs=1e-6)
    assert float(P.linear_to_srgb(0.5)) == pytest.approx(0.7353569, abs=5e-5)
    for x in (0.001, 0.05, 0.2, 0.8):
        assert float(P.srgb_to_linear(P.linear_to_srgb(x))) == pytest.approx(x, abs=5e-6)
    # the two pieces meet at the breakpoint to within the standard's rounding
    lo = 0.0031308 * 12.92
    hi = 1.055 * 0.0031308 ** (1 / 2.4) - 0.055
    assert abs(lo - hi) < 1e-7


def test_present_rounds_to_binary16_and_clamps():
    out = P.present(np.array([[1e9, 0.5, -1.0]]), "none", 1.0, "linear")
    assert out.tolist() == [[65504.0, 0.5, 0.0]]
    v = P.present(np.array([[0.1234567, 0.0, 2.0]]), "reinhard", 1.0, "srgb")
    assert (v.astype(np.float16).astype(np.float64) == v).all()
