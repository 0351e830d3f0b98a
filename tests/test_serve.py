"""Host side of the serving loop: the seeded request mix of the serving configuration."""
import numpy as np

from blochlib_amd import serve


def test_config5_generator_shape():
    reqs = serve.config5_requests(80)
    S = np.array([r.system.spins for r in reqs])
    assert S.min() >= 2 and S.max() <= 10
    assert all(int(round(r.sw / r.wr)) % 2 == 0 for r in reqs)
    rate = len(reqs) / reqs[-1].arrival
    assert 120 < rate < 300
