"""The results file (safire_amd.results over libafq_h5.so): the reference's layout
(format_version on the root, Measurements/Stage<N>/<observable>/bins extendable along the bin
axis, complex values as a trailing axis of 2 with the __complex__ attribute), appends that grow
the datasets, the bins of binsize consecutive measurements, and the rebinned averages and
jackknife read back from it."""
import os
import subprocess

import numpy as np
import pytest

from safire_amd import h5inputs, stats

pytestmark = pytest.mark.skipif(not os.path.exists(h5inputs.LIB_PATH),
                                reason="HDF5 reader library not built")
H5DUMP = "/opt/conda/bin/h5dump"


def _trace(n, rng):
    from safire_amd.driver import Trace
    tr = Trace(list(range(0, 2 * n, 2)), [], [], [], [])
    comps = rng.normal(size=(n, 3)) + 1j * rng.normal(size=(n, 3))
    tr.components = list(comps)
    tr.energy = list(comps.sum(1))
    tr.total_weight = list(rng.normal(size=n) + 0j)
    tr.eff_walkers = list(rng.uniform(10, 20, size=n))
    tr.overlap = list(rng.normal(size=n) + 1j * rng.normal(size=n))
    tr.phase = list(np.exp(1j * rng.normal(size=n)))
    tr.backprop = {4: {"OneRDM": list(rng.normal(size=(n // 2, 1, 3, 3)) + 0j)}}
    return tr


def test_layout_append_and_analysis(tmp_path):
    from safire_amd.results import Results, write_results
    rng = np.random.default_rng(3)
    tr = _trace(40, rng)
    path = tmp_path / "run.results.h5"
    left = write_results(path, tr, dt=0.01, binsize=3)
    assert len(left["Energy"]) == 40 % 3 and len(left["BackPropEstimator/Steps=4/OneRDM"]) == 20 % 3
    r = Results(path)
    names = r.observable_names()
    assert "Energy" in names and "BackPropEstimator/Steps=4/OneRDM" in names
    assert "ProjectionTime" in names and "Phase" in names
    e = np.asarray(tr.energy)
    np.testing.assert_allclose(r.bins("Energy"), e[:39].reshape(13, 3).mean(1), rtol=1e-15)
    np.testing.assert_allclose(r.bins("ProjectionTime"),
                               (np.arange(0, 80, 2)[:39] + 1).reshape(13, 3).mean(1) * 0.01)
    assert r.bins("EffectiveNumWalkers").dtype == np.float64
    rdm = r.bins("BackPropEstimator/Steps=4/OneRDM")
    assert rdm.shape == (6, 1, 3, 3) and rdm.dtype == np.complex128
    # a second write grows the datasets
    write_results(path, tr, dt=0.01, binsize=3)
    assert r.bins("Energy").shape == (26,)
    m, err = r.average("Energy", rebinsize=2)
    b, _ = stats.rebinning_analysis(r.bins("Energy"), rebinsize=2)
    np.testing.assert_allclose(m, b.mean(0))
    np.testing.assert_allclose(err, stats.standard_error(b))
    v, ve = r.evaluate(lambda a, c: a - c, "Energy", "OnebodyEnergy", rebinsize=2)
    a2, _ = stats.rebinning_analysis(r.bins("Energy") - r.bins("OnebodyEnergy"), rebinsize=2)
    np.testing.assert_allclose(v, a2.mean(0), rtol=1e-12)
    # shape mismatch on append is an error
    from safire_amd.results import _append
    with pytest.raises(h5inputs.H5Error):
        _append(path, "Stage0", "Energy", np.zeros((2, 3), np.complex128))
    if os.path.exists(H5DUMP):
        hdr = subprocess.run([H5DUMP, "-H", "-A", str(path)], capture_output=True, text=True).stdout
        assert 'ATTRIBUTE "format_version"' in hdr and 'ATTRIBUTE "__complex__"' in hdr
        assert 'GROUP "Measurements"' in hdr and 'GROUP "Stage0"' in hdr
        assert "H5S_UNLIMITED" in hdr
