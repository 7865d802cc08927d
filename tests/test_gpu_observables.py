"""Static correlations and local measurements on the device against the
oracle: Observables::TwoPointCorr (src/Observables.h:132-155) for spin and
current operators, relative 1e-12 of the largest entry."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

esc = pytest.importorskip("energyspincurrents_amd")


M12 = O.Model(3, 4, K=(1.0, 1.0, 1.0), B=(0.3, 0.3, 0.3))


@pytest.fixture(scope="module")
def ground():
    return O.ground_pair(M12)[1]


@pytest.mark.parametrize("op_a,op_b", [("Sx", "Sx"), ("Sx", "Sz"), ("Sy", "Sz"), ("Jsx", "Jsy"), ("Jex", "Jex"),
                                       ("Jey", "Sz")])
def test_two_point_matches_oracle(ground, op_a, op_b):
    m, gs = M12, ground
    lat = esc.Honeycomb(m.Lx, m.Ly)
    eng = esc.Engine(esc.SpinModel(lat, K=m.K, B=m.B), "c128")
    C = eng.two_point(op_a, op_b, eng.upload(gs))
    ref = O.two_point_ref(m, op_a, op_b, gs)
    assert np.abs(C - ref).max() < 1e-12 * max(1.0, np.abs(ref).max())
