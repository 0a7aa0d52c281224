"""The oracle's GCMC / NVT result statistics, block_mean_sem (src/nvt.jl:126-140) and
fluctuation_qst (src/gcmc.jl:71-113), against independent computations: numpy's block means and
standard deviation where the blocks divide evenly, an explicit delete-one-block jackknife that
drops the samples themselves and recomputes cov/var with numpy, exact linear relations, and the
reference's argument errors."""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle as O
import pureadsorb_amd as pa


def test_block_mean_sem_even_blocks_match_numpy():
    x = np.random.default_rng(1).normal(3.0, 2.0, 120)
    m, se = O.block_mean_sem(x, 6)
    bm = x.reshape(6, 20).mean(axis=1)
    assert m == pytest.approx(bm.mean(), rel=1e-14)
    assert se == pytest.approx(bm.std(ddof=1) / math.sqrt(6), rel=1e-13)


def test_block_mean_sem_remainder_goes_to_the_last_block():
    """23 samples in 4 blocks: blocks of 5, 5, 5 and 8 (the last takes the remainder), and the
    mean is the mean of block means, not the pooled mean."""
    x = np.arange(23, dtype=np.float64) ** 1.5
    m, se = O.block_mean_sem(x, 4)
    bm = np.array([x[0:5].mean(), x[5:10].mean(), x[10:15].mean(), x[15:23].mean()])
    assert m == pytest.approx(bm.mean(), rel=1e-14)
    assert m != pytest.approx(x.mean(), rel=1e-6)
    assert se == pytest.approx(bm.std(ddof=1) / 2.0, rel=1e-13)


def test_block_mean_sem_constant_series_and_errors():
    assert O.block_mean_sem(np.full(40, 2.5), 10) == (2.5, 0.0)
    with pytest.raises(ValueError):
        O.block_mean_sem(np.ones(10), 1)
    with pytest.raises(ValueError):
        O.block_mean_sem(np.ones(19), 10)


def _jackknife_brute(U, N, kT, nb):
    """Delete-one-block jackknife with the block's samples removed from the arrays and the
    population covariance / variance recomputed by numpy on what is left."""
    ns = len(U)
    bl = ns // nb
    idx = [np.arange(b * bl, ns if b == nb - 1 else (b + 1) * bl) for b in range(nb)]

    def q_of(u, n):
        c = np.cov(np.vstack([u, n]), bias=True)
        return kT - c[0, 1] / c[1, 1]

    loo = np.array([q_of(np.delete(U, i), np.delete(N, i)) for i in idx])
    return q_of(U, N), math.sqrt((nb - 1) / nb * np.sum((loo - loo.mean()) ** 2)), np.corrcoef(U, N)[0, 1]


@pytest.mark.parametrize("ns,nb", [(20, 2), (37, 5), (200, 10)])
def test_fluctuation_qst_matches_an_explicit_jackknife(ns, nb):
    rng = np.random.default_rng(ns)
    N = rng.poisson(6.0, ns).astype(np.float64)
    U = -0.3 * N + rng.normal(0, 0.2, ns)
    kT = pa.KB * 298.15
    q, qe, c = O.fluctuation_qst(U, N, kT, nb)
    qb, qeb, cb = _jackknife_brute(U, N, kT, nb)
    assert q == pytest.approx(qb, rel=1e-11)
    assert qe == pytest.approx(qeb, rel=1e-8)
    assert c == pytest.approx(cb, rel=1e-11)


def test_fluctuation_qst_of_an_exact_linear_relation():
    """U = a + b N exactly: cov(U, N)/var(N) = b in every subsample, so q_st = kT - b, every
    leave-one-block-out estimate is the same (error 0) and corr(U, N) = sign(b)."""
    N = np.array([3, 5, 4, 7, 6, 2, 8, 5, 4, 6, 9, 3], dtype=np.float64)
    kT = 0.025
    for b in (-0.25, 0.5):
        U = 1.75 + b * N
        q, qe, c = O.fluctuation_qst(U, N, kT, 3)
        assert q == pytest.approx(kT - b, abs=1e-12)
        assert qe < 1e-10
        assert c == pytest.approx(math.copysign(1.0, b), abs=1e-12)


def test_fluctuation_qst_argument_errors():
    with pytest.raises(ValueError, match="same length"):
        O.fluctuation_qst(np.ones(10), np.ones(9), 0.025, 2)
    with pytest.raises(ValueError, match="at least two blocks"):
        O.fluctuation_qst(np.ones(10), np.ones(10), 0.025, 1)
    with pytest.raises(ValueError, match="production cycles"):
        O.fluctuation_qst(np.ones(9), np.ones(9), 0.025, 5)
