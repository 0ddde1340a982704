"""Oracle pins of contacts with a surface block (surface_green,
contselfenergy.F90:186-239: GS = inv([[z Sss - Hss, z Ss1 - Hs1],
[z S1s - H1s, inv(g_PL)]]), then Sigma = Tlc GS Tcl over surface + PL1).

No reference-held vector covers this path; two exact identities pin it:
a surface coupled to the device only through itself is the same system as
that surface moved into the device as an extra layer (with the device at the
contacts' complex energy), and a surface identical to a lead layer leaves the
self-energy of the plain contact unchanged."""
import numpy as np
import pytest

from tests.surfaces import lead_copy_surface, surface_pair


@pytest.fixture(scope="module")
def oracle():
    from oracle import pyoracle
    pyoracle.build()
    pyoracle.set_backend("builtin")
    return pyoracle


@pytest.mark.parametrize("overlap", [False, True])
@pytest.mark.parametrize("E", [-1.3, -0.2, 0.7])
def test_surface_equals_device_layer(oracle, overlap, E):
    A, B = surface_pair(overlap=overlap)
    d = 1e-3  # contacts and device at E + i d in both
    ra = oracle.transmission_dos(A, E, d, d, full=True)
    rb = oracle.transmission_dos(B, E, d, d, full=True)
    assert abs(ra["T"][0] - rb["T"][0]) <= 1e-10 * max(1.0, abs(rb["T"][0]))
    n = A.ndev
    assert np.abs(ra["diag"] - rb["diag"][:n]).max() <= 1e-10 * np.abs(rb["diag"]).max()
    assert ra["ncyc"].tolist() == rb["ncyc"].tolist()


def test_surface_equal_to_lead_layer(oracle):
    A, B = lead_copy_surface()
    nd, ns = A.bsz[-1], A.contacts[0].ns
    for E in (-1.1, 0.3, 2.6):
        z = E + 1e-4j
        sa, gsa, _ = oracle.contact_sigma(A.contacts[0], nd, z)
        sb, gsb, _ = oracle.contact_sigma(B.contacts[0], nd, z)
        assert gsa.shape == (ns + B.contacts[0].nc,) * 2
        # the surface block's function is the lead's own surface function
        assert np.abs(gsa[:ns, :ns] - gsb).max() <= 1e-7 * np.abs(gsb).max()
        assert np.abs(sa - sb).max() <= 1e-7 * np.abs(sb).max()
    Ta = oracle.transmission(A, 0.3, 1e-4)
    Tb = oracle.transmission(B, 0.3, 1e-4)
    assert abs(Ta[0] - Tb[0]) <= 1e-7 * max(1.0, Tb[0])


def test_surface_from_dense_matches_builder(oracle):
    """The dense-matrix extraction (surfstart..surfend, ln_extract.F90:41-76)
    reproduces the structure built block by block."""
    from tests.chains import from_dense
    A, _ = surface_pair(W=4, nbl=3, ns=2, seed=7)
    W, nbl, ns = 4, 3, 2
    cR, cL = A.contacts
    # dense order: device (12), right contact [surface 2, PL 4, PL 4], left contact [PL 4, PL 4]
    N = W * nbl + ns + 4 * W
    H = np.zeros((N, N), dtype=np.complex128)
    for i in range(nbl):
        H[i * W:(i + 1) * W, i * W:(i + 1) * W] = A.Hd[i]
    for i in range(nbl - 1):
        H[i * W:(i + 1) * W, (i + 1) * W:(i + 2) * W] = A.Hu[i]
        H[(i + 1) * W:(i + 2) * W, i * W:(i + 1) * W] = A.Hl[i]
    r0 = W * nbl
    s_, p1, p2 = r0, r0 + ns, r0 + ns + W
    H[s_:p1, s_:p1] = cR.Hss
    H[s_:p1, p1:p2] = cR.Hs1
    H[p1:p2, s_:p1] = cR.H1s
    H[p1:p2, p1:p2] = H[p2:p2 + W, p2:p2 + W] = cR.H00
    H[p1:p2, p2:p2 + W] = cR.H01
    H[p2:p2 + W, p1:p2] = cR.H01.conj().T
    d0 = (nbl - 1) * W
    H[d0:d0 + W, s_:p2] = cR.HMC
    H[s_:p2, d0:d0 + W] = cR.HMC.conj().T
    l1 = p2 + W
    H[l1:l1 + W, l1:l1 + W] = H[l1 + W:l1 + 2 * W, l1 + W:l1 + 2 * W] = cL.H00
    H[l1:l1 + W, l1 + W:l1 + 2 * W] = cL.H01
    H[l1 + W:l1 + 2 * W, l1:l1 + W] = cL.H01.conj().T
    H[0:W, l1:l1 + W] = cL.HMC
    H[l1:l1 + W, 0:W] = cL.HMC.conj().T
    plend = [W * (i + 1) for i in range(nbl)]
    D = from_dense(H, None, plend, [s_ + 1, l1 + 1], [l1, N], [nbl, 1], surfend=[p1, l1])
    assert D.contacts[0].ns == ns and D.contacts[1].ns == 0
    for E in (-0.4, 1.1):
        Ta = oracle.transmission(A, E, 1e-4)
        Td = oracle.transmission(D, E, 1e-4)
        assert abs(Ta[0] - Td[0]) <= 1e-12 * max(1.0, Ta[0])
