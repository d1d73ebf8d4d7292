"""Deterministic synthetic inputs shared by fixture generation and tests.

Every array is drawn from numpy's PCG64 generator in a fixed order, so the
same seed gives the same limbs here and on any machine with this numpy.
"""
import numpy as np


def uniform_limbs(rng, primes, first_prime, limbs, n, upper_factor=1):
    """[limbs][n] int64 residues, limb l uniform in [0, factor*q[first+l])."""
    out = np.empty((limbs, n), np.int64)
    for l in range(limbs):
        q = int(primes[first_prime + l]) * upper_factor
        out[l] = rng.integers(0, q, n, dtype=np.uint64).astype(np.int64)
    return out


def hmult_inputs(primes, n, mid, chain, key_mid, key_polys, key_chain, seed):
    """Two rescaled INTT-form ciphertexts at `mid` (canonical residues) and a
    relinearization key at `key_mid` (NTT-form, lazy residues in [0, 2q)).

    Returns flat int64 arrays A, B ([2][chain][n]) and K
    ([key_polys][key_chain][n]).  The key's limbs follow the reference layout:
    the Q limbs from key_mid, then the special primes.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    primes = [int(q) for q in primes]
    A = np.stack([uniform_limbs(rng, primes, mid, chain, n) for _ in range(2)])
    B = np.stack([uniform_limbs(rng, primes, mid, chain, n) for _ in range(2)])
    K = np.stack([uniform_limbs(rng, primes, key_mid, key_chain, n, 2)
                  for _ in range(key_polys)])
    return A.ravel(), B.ravel(), K.ravel()
