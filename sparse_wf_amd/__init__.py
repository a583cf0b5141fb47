"""MI355X-native FiRE hot path (neighbour lists, embedding, orbitals, log-determinant,
forward-Laplacian local energy, low-rank MCMC) behind the reference model API."""
