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


def test_chunk_density_generator_equals_the_per_map_one():
    """binary_densities (one draw for the chunk, the blur on a thread pool) gives the maps of binary_density
    called once per map, bit for bit, and leaves the generator in the same state."""
    from beamz_amd.workloads import binary_densities, binary_density
    for m, n in ((0, 16), (1, 32), (37, 48)):
        r1, r2 = np.random.default_rng(5), np.random.default_rng(5)
        a = np.stack([binary_density(n, r1) for _ in range(m)]) if m else np.zeros((0, n, n), bool)
        b = binary_densities(m, n, r2)
        assert b.dtype == bool and b.shape == (m, n, n)
        assert np.array_equal(a, b)
        assert np.array_equal(r1.random(4), r2.random(4))
