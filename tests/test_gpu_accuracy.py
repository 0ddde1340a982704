"""Accuracy of the GPU transmission against the algorithm's exact value, at
the full configuration sizes.

The truth is the double-double evaluation of the same algorithm on the same
double inputs (oracle/negf_dd.c; fixtures written by
tests/golden/make_dd_fixtures.py): C2, the disordered ribbon at all 2000
energies, and C3, the k-resolved slab at 40 PLs, 24 energies over 4 k-points.
Both also hold the FP64 reference CPU path's own rounding error at each point
(built-in LAPACK-style loops and MKL), which is what "as accurate as the
reference" means here: at eta = 1e-6 (C2) and 1e-5 (C3) the blocks have
condition numbers up to ~1/eta and the 40-100 block recursion amplifies
rounding, so at a few percent of the energies any FP64 evaluation -- MKL's
included -- is off by 1e-8 to 1e-4, and per point the errors of two correct
FP64 implementations are independent random draws (MKL fails "within 2x of
the built-in loops' error" at ~5 % of the C2 points).  The bars are
therefore:

* where the reference path is accurate (its error below 1e-11), the GPU is
  within 1e-10 relative -- the 1e-10 target, everywhere it is attainable;
* over the ill-conditioned points (the reference path's error above 1e-10)
  the GPU's error per point, relative to the geometric mean of the two
  backends' errors there, has a mean log10 below LOG_MEAN and a 90th
  percentile below LOG_P90 -- the spread MKL and the built-in loops show
  against each other (on C2 their log10 ratio has mean 0.02 and 90th
  percentile 0.89; on C3 0.58 and 1.17);
* the GPU's largest error is within MAX_RATIO of the reference path's;
* the same decimation iteration counts everywhere.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
LOG_MEAN = 0.35
LOG_P90 = 1.0
MAX_RATIO = 10.0


def _check(T, fx, label):
    Tt = fx["T_hi"] + fx["T_lo"]
    big = Tt > 1e-6 * Tt.max()
    e_gpu = np.abs(T - Tt) / Tt
    e_b = np.abs(fx["T_builtin"] - Tt) / Tt
    e_m = np.abs(fx["T_mkl"] - Tt) / Tt if np.isfinite(fx["T_mkl"]).all() else e_b
    e_cpu = np.maximum(e_b, e_m)
    good = big & (e_cpu < 1e-11)
    assert good.sum() >= (0.5 * big.sum() if label == "c2" else 4)
    assert e_gpu[good].max() <= 1e-10, (label, e_gpu[good].max())
    ill = big & (e_cpu >= 1e-10)
    geo = np.sqrt(np.maximum(e_b[ill], 1e-17) * np.maximum(e_m[ill], 1e-17))
    lr = np.log10(np.maximum(e_gpu[ill], 1e-17) / geo)
    stats = dict(n_good=int(good.sum()), n_ill=int(ill.sum()), log_mean=float(lr.mean()),
                 log_p90=float(np.quantile(lr, 0.9)), max_gpu=float(e_gpu[big].max()),
                 max_cpu=float(e_cpu[big].max()))
    assert stats["log_mean"] <= LOG_MEAN, (label, stats)
    assert stats["log_p90"] <= LOG_P90, (label, stats)
    assert stats["max_gpu"] <= MAX_RATIO * stats["max_cpu"], (label, stats)
    return stats


def test_c2_against_double_double():
    from libnegf_amd import hip, workloads as wl
    fx = np.load(os.path.join(HERE, "golden", "dd_c2.npz"))
    s, E, eta = wl.config2()
    assert np.array_equal(E, fx["E"]) and eta == fx["eta"]
    eng = hip.Engine(0)
    ds = hip.DeviceStructure(s)
    T, ncyc, st = hip.solve_transmission(eng, ds, E, delta=eta)
    assert (st == 0).all()
    assert np.array_equal(ncyc, fx["ncyc"])
    print(_check(T, fx, "c2"))


def test_c3_nbl40_against_double_double():
    from libnegf_amd import hip, workloads as wl
    fx = np.load(os.path.join(HERE, "golden", "dd_c3.npz"))
    slab, kg, E, eta = wl.config3()
    ks = sorted(set(fx["k"].tolist()))
    ds = hip.DeviceStructure([slab.at_k(*kg[k]) for k in ks])
    slot = np.array([ks.index(k) for k in fx["k"]], dtype=np.int32)
    assert np.array_equal(E[fx["idx"]], fx["E"])
    eng = hip.Engine(0)
    T, ncyc, st = hip.solve_transmission(eng, ds, fx["E"], slot=slot, delta=eta)
    assert (st == 0).all()
    assert np.array_equal(ncyc, fx["ncyc"])
    print(_check(T, fx, "c3"))
