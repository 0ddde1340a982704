"""Host logic of the inelastic models: the library's K tables
(hipnegf_inel_kmat_po / _nonpo, prepare_POKmat / prepare_NonPOKmat of
elphinel.F90:318-519) equal the oracle's, which the q_z quadrature pins
(tests/test_oracle_inelastic.py).  Host-only calls, no GPU needed."""
import numpy as np
import pytest

from oracle import pyoracle
from tests.inel_model import LATT


@pytest.fixture(scope="module")
def libs():
    from libnegf_amd import hip
    pyoracle.build()
    return hip


def test_kmat_tables_match_oracle(libs):
    hip = libs
    kfrac = np.array([[0.0, 0.0, 0.0], [0.2, 0.1, 0.0], [0.45, -0.3, 0.0], [0.1, 0.4, 0.0]])
    kw = np.array([0.1, 0.2, 0.3, 0.4])
    for latt in (LATT, None, np.array([[2.0, 0.5, 0.0], [0.0, 1.7, 0.0], [0.0, 0.0, 8.0]])):
        a = hip.inel_kmat_po(7, 0.4, kfrac, kw, 1.3, 0.06, 2.5, 12.0, 10.0, 0.7, latt=latt)
        b = pyoracle.inel_kmat_po(7, 0.4, kfrac, kw, 1.3, 0.06, 2.5, 12.0, 10.0, 0.7, latt=latt)
        np.testing.assert_allclose(a, b, rtol=1e-14, atol=0)
        a = hip.inel_kmat_nonpo(5, 0.5, kfrac, kw, 0.8, 2.0, latt=latt)
        b = pyoracle.inel_kmat_nonpo(5, 0.5, kfrac, kw, 0.8, 2.0, latt=latt)
        np.testing.assert_allclose(a, b, rtol=1e-14, atol=0)
    # all lattice vectors 0: every k is the origin, the non-polar table vanishes
    assert np.all(hip.inel_kmat_nonpo(5, 0.5, kfrac, kw, 0.8, 2.0) == 0.0)
