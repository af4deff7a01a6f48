"""Host logic of the surrogate's label generation (no GPU): the pipelined generator draws exactly the
densities of random_densities(n, seed), in order, and pairs each chunk's labels with its own densities."""
import numpy as np
import pytest

from beamz_amd.surrogate.data import generate, random_densities


class _Echo:
    """Stands in for TransmissionProblem: the 'spectrum' of a map is its mean density at every wavelength."""
    px, ref = 128, np.full(8, 2.0)
    sims = [None] * 4

    def submit(self, rho):
        assert 0 < len(rho) <= len(self.sims)
        self.rho = np.array(rho)

    def collect(self, m):
        return np.repeat(self.rho[:m].mean(axis=(1, 2))[:, None], 8, axis=1)


@pytest.mark.parametrize("n", [0, 1, 4, 10])
def test_generate_pairs_densities_with_their_labels(n):
    rho, T = generate(_Echo(), n, seed=3)
    assert np.array_equal(rho, random_densities(n, 3))
    assert T.shape == (n, 8) and T.dtype == np.float32
    assert np.allclose(T, rho.mean(axis=(1, 2))[:, None] / 2.0)
