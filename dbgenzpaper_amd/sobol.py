"""Sobol direction integers for the fused Sobol point generator.

The reference draws Sobol points from Sobol.jl 1.5.0 (``SobolSeq``,
src/qmc_generators.jl:57-65, 331-352), whose direction numbers are the
Joe & Kuo ``new-joe-kuo-6.21201`` table with the first dimension being the
van der Corput sequence (all m_k = 1).  Sobol.jl is not part of this
repository, so the same published table is read from the copy that SciPy ships
as plain data (``scipy/stats/_sobol_direction_numbers.npz``: the primitive
polynomial and initial m_k of each dimension).

Point ``idx`` of the sequence (the idx-th call of Sobol.jl ``next!``, starting
at idx = 1) is the XOR of the direction integers V[d, k] selected by the bits of
the Gray code idx ^ (idx >> 1); V[d, k] = m_{k+1} << (31 - k).
"""
from __future__ import annotations

import functools
import os

import numpy as np

MAX_BITS = 32


def _table_path() -> str:
    import scipy  # the table is data inside the SciPy installation

    return os.path.join(os.path.dirname(scipy.__file__), "stats", "_sobol_direction_numbers.npz")


@functools.lru_cache(maxsize=4)
def _table():
    with np.load(_table_path(), allow_pickle=False) as z:
        return np.array(z["poly"], np.int64), np.array(z["vinit"], np.int64)


def direction_integers(n_dims: int) -> np.ndarray:
    """(n_dims, 32) uint32 direction integers V[d, k] = m_{k+1} << (31 - k)."""
    poly, vinit = _table()
    if n_dims > poly.shape[0]:
        raise ValueError(f"Sobol directions exist for at most {poly.shape[0]} dimensions")
    V = np.zeros((max(n_dims, 0), MAX_BITS), np.uint32)
    for d in range(n_dims):
        m = np.zeros(MAX_BITS + 1, np.int64)  # m[1..32]
        if d == 0:
            m[1:] = 1
        else:
            p = int(poly[d])
            s = p.bit_length() - 1  # degree of the primitive polynomial
            for k in range(1, s + 1):
                m[k] = int(vinit[d, k - 1])
            for k in range(s + 1, MAX_BITS + 1):
                v = m[k - s] ^ (m[k - s] << s)
                for j in range(1, s):
                    if (p >> (s - j)) & 1:
                        v ^= m[k - j] << j
                m[k] = v
        for k in range(MAX_BITS):
            V[d, k] = np.uint32((int(m[k + 1]) << (MAX_BITS - 1 - k)) & 0xFFFFFFFF)
    return V


def skip_count(n: int) -> int:
    """Number of ``next!`` calls made by ``skip(SobolSeq(d), n)``: the largest
    2^k - 1 <= n (Sobol.jl's non-exact skip).  The reference skips n = 2m - 1
    (src/qmc_generators.jl:57-58, src/qmc.jl:178); for m a power of two this is
    exactly 2m - 1."""
    if n <= 0:
        return 0
    k = (n + 1).bit_length() - 1
    return (1 << k) - 1
