"""Surface-function disk cache of the C API (SGFcache = 0, readOldT_SGFs 0/2):
file names, the reference's record format (ln_cache.F90:458-594,
outmatrix.F90:187-245), the values written against the oracle's g_s, and
that a load solve uses what the files hold."""
import os
import shutil

import numpy as np
import pytest

from tests.chains import chain_dense, from_dense
from tests.test_api import _chain_session, gpu_negf  # noqa: F401  (fixture)

REC = np.dtype([("l0", "<i4"), ("i", "<i4"), ("j", "<i4"), ("re", "<f8"), ("im", "<f8"),
                ("l1", "<i4")])


def read_records(path, n):
    """Fortran unformatted sequential records (int32 i, int32 j, complex(dp)),
    each framed by its byte length, into an n x n matrix."""
    raw = open(path, "rb").read()
    assert len(raw) % 32 == 0
    r = np.frombuffer(raw, dtype=REC)
    assert (r["l0"] == 24).all() and (r["l1"] == 24).all()
    g = np.zeros((n, n), dtype=np.complex128)
    g[r["i"] - 1, r["j"] - 1] = r["re"] + 1j * r["im"]
    return g, r


def _session(NEGF, H, scratch, readold, cache=0):
    n = _chain_session(NEGF, H, [15, 30, 45, 60], [4, 1], -1.0, 1.0, 0.25)
    n.params.sgfcache = cache
    n.params.readold_t_sgfs = readold
    n.set_scratch(str(scratch))
    return n


def _H():
    ons = np.zeros(60)
    ons[25] = ons[35] = 1.0
    return chain_dense(onsite=ons)


@pytest.mark.gpu
def test_sgf_files_written_and_reloaded(gpu_negf, tmp_path):
    NEGF, orc = gpu_negf
    H = _H()
    a = _session(NEGF, H, tmp_path, 2)  # compute and save
    a.solve_landauer()
    E, _ = a.get_energies()
    Ta = a.get_transmission()
    assert (a.get_item_ncyc() > 0).all()
    a.close()
    names = sorted(os.listdir(tmp_path))
    want = sorted(f"Matu{c:02d}_00_01_{p:02d}.dat" for c in (1, 2) for p in range(1, len(E) + 1))
    assert names == want
    # values against the oracle's g_s at E + i delta; records in column order,
    # zero entries left out
    s = from_dense(H, None, [15, 30, 45, 60], [61, 81], [80, 100], [4, 1])
    for p in (1, 5, 9):
        for c in (0, 1):
            g, r = read_records(tmp_path / f"Matu{c + 1:02d}_00_01_{p:02d}.dat", 10)
            order = (r["j"].astype(np.int64) - 1) * 10 + (r["i"] - 1)
            assert (np.diff(order) > 0).all()
            assert ((r["re"] != 0) | (r["im"] != 0)).all()
            nd = s.bsz[s.contacts[c].cblk - 1]
            _, GS, _ = orc.contact_sigma(s.contacts[c], nd, complex(E[p - 1], 1e-4))
            assert np.abs(g - GS).max() <= 1e-10 * np.abs(GS).max()
    # a load solve reads every file: same T bitwise, no decimation
    b = _session(NEGF, H, tmp_path, 0)
    b.solve_landauer()
    np.testing.assert_array_equal(b.get_transmission(), Ta)
    assert (b.get_item_ncyc() == 0).all() and (b.get_item_status() == 0).all()
    b.close()


@pytest.mark.gpu
def test_sgf_load_uses_file_contents(gpu_negf, tmp_path):
    """Point 5's files replaced by point 6's: with S = I the coupling does not
    depend on E, so the load solve's T(E5) is the two-contact T at E5 with the
    self-energies of E6 (oracle).  Point 7's files removed: recomputed and
    written again."""
    NEGF, orc = gpu_negf
    H = _H()
    a = _session(NEGF, H, tmp_path, 2)
    a.solve_landauer()
    E, _ = a.get_energies()
    Ta = a.get_transmission()
    a.close()
    for c in (1, 2):
        shutil.copy(tmp_path / f"Matu{c:02d}_00_01_06.dat", tmp_path / f"Matu{c:02d}_00_01_05.dat")
        os.remove(tmp_path / f"Matu{c:02d}_00_01_07.dat")
    b = _session(NEGF, H, tmp_path, 0)
    b.set_batch(1)  # one point per batch: point 7 alone is recomputed
    b.solve_landauer()
    Tb = b.get_transmission()
    b.close()
    s = from_dense(H, None, [15, 30, 45, 60], [61, 81], [80, 100], [4, 1])
    sig6 = [orc.contact_sigma(s.contacts[c], s.bsz[s.contacts[c].cblk - 1],
                              complex(E[5], 1e-4))[0] for c in (0, 1)]
    t5 = orc.transmission_sigma(s, E[4], sig6)
    t6 = orc.transmission_sigma(s, E[5], sig6)
    assert abs(Ta[0, 5] - t6) <= 1e-10 * abs(t6)  # the oracle's T with given self-energies
    assert abs(Tb[0, 4] - t5) <= 1e-10 * abs(t5)
    assert abs(Tb[0, 4] - Ta[0, 4]) > 1e-6 * abs(Ta[0, 4])  # the planted functions were used
    assert abs(Tb[0, 6] - Ta[0, 6]) <= 1e-12 * abs(Ta[0, 6])
    assert all(os.path.exists(tmp_path / f"Matu{c:02d}_00_01_07.dat") for c in (1, 2))
    keep = [i for i in range(len(E)) if i != 4]
    np.testing.assert_allclose(Tb[0, keep], Ta[0, keep], rtol=1e-12, atol=0)


@pytest.mark.gpu
def test_sgf_memory_and_dummy_modes_write_nothing(gpu_negf, tmp_path):
    NEGF, _ = gpu_negf
    H = _H()
    ref = _session(NEGF, H, tmp_path, 1)
    ref.solve_landauer()
    T0 = ref.get_transmission()
    ref.close()
    for cache in (1, 2):
        n = _session(NEGF, H, tmp_path, 2, cache=cache)
        n.solve_landauer()
        np.testing.assert_allclose(n.get_transmission(), T0, rtol=1e-12, atol=0)
        n.close()
    assert os.listdir(tmp_path) == []


@pytest.mark.gpu
def test_sgf_files_ldos_dephasing_and_surface(gpu_negf, tmp_path):
    """The LDOS and the dephasing (SCBA) solves save and reload the same
    files with identical results; a contact with a surface block writes g_s of
    order ns + nc, equal to the oracle's."""
    NEGF, orc = gpu_negf
    H = _H()
    for mode in ("ldos", "deph"):
        d = tmp_path / mode
        d.mkdir()
        out = []
        for readold in (2, 0):
            if mode == "ldos":
                n = _session(NEGF, H, d, readold)
                n.set_ldos_intervals(np.array([1, 31]), np.array([30, 60]))
            else:  # the SCBA path wants PLs of the contacts' size: 6 PLs of 10 sites
                n = _chain_session(NEGF, chain_dense(), [10, 20, 30, 40, 50, 60], [6, 1], -1.0, 1.0, 0.25)
                n.params.sgfcache, n.params.readold_t_sgfs = 0, readold
                n.set_scratch(str(d))
                n.set_elph_dephasing(np.full(60, 0.05), 5)
            n.solve_landauer()
            out.append((n.get_transmission(), n.get_ldos() if mode == "ldos" else n.get_energy_current()))
            if readold == 0 and mode == "ldos":
                assert (n.get_item_ncyc() == 0).all()
            n.close()
        assert len(os.listdir(d)) == 18
        np.testing.assert_array_equal(out[0][0], out[1][0])
        np.testing.assert_array_equal(out[0][1], out[1][1])
    # a surface block on contact 1
    import scipy.sparse as sps
    from tests.surfaces import surface_pair
    from tests.test_gpu_surface import _dense_with_surface
    A, _ = surface_pair(W=4, nbl=4, ns=3, seed=5)
    Hs, plend, surfstart, surfend, contend, cblk = _dense_with_surface(A)
    d = tmp_path / "surf"
    d.mkdir()
    n = NEGF()
    n.set_h(sps.csr_matrix(Hs))
    n.set_s_id(Hs.shape[0])
    n.init_structure(2, np.array(contend), np.array(surfend), len(plend), np.array(plend), np.array(cblk),
                     surfstart=np.array(surfstart))
    p = n.params
    p.emin, p.emax, p.estep, p.delta = -1.5, 1.5, 0.5, 1e-3
    p.sgfcache, p.readold_t_sgfs = 0, 2
    n.set_scratch(str(d))
    n.solve_landauer()
    E, _ = n.get_energies()
    n.close()
    D = from_dense(Hs, None, plend, surfstart, contend, cblk, surfend=surfend)
    c0 = D.contacts[0]
    ng = c0.nc + c0.ns
    g, _ = read_records(d / "Matu01_00_01_03.dat", ng)
    _, GS, _ = orc.contact_sigma(c0, D.bsz[c0.cblk - 1], complex(E[2], 1e-3))
    assert np.abs(g - GS).max() <= 1e-10 * np.abs(GS).max()
