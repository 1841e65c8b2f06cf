"""The network parameters committed for the default benchmark's
time-to-target field (bench.py, written by `bench.py --workload c5
--save-params`): a plain float32 array, one value per parameter of the
library's network, finite, and trained (not the seeded initialisation)."""
import os

import numpy as np

from ettention_amd import learned as L

PARAMS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "ettention_amd", "data",
                      "correction_net_512.npy")


def test_committed_parameters_fit_the_network():
    p = np.load(PARAMS, allow_pickle=False)
    assert p.dtype == np.float32
    assert p.shape == (L.parameter_count(),)
    assert np.all(np.isfinite(p))
    init = L.init_parameters(5000)
    assert np.linalg.norm(p - init) > 1e-3 * np.linalg.norm(init)
