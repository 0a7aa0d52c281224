"""Pins the oracle's counter-based RNG and pose draws (src/state.jl:225-230, src/rng.jl:14-141,
src/moves.jl:154-161) against published values, exact integer arithmetic and mpmath."""
from __future__ import annotations

import random

import mpmath
import numpy as np
import pytest

import oracle as O

M64 = (1 << 64) - 1


def test_splitmix64_published_outputs():
    """splitmix64(seed, i) for i = 1, 2, 3 is the i-th output of Vigna's reference SplitMix64
    generator seeded with `seed` (the state advances by the golden gamma before mixing); its
    published outputs for seed 1234567 pin the constants and shifts."""
    assert [O.splitmix64(1234567, i) for i in (1, 2, 3)] == [
        6457827717110365317, 3203168211198807973, 9817491932198370423]


def test_mulhi64_matches_128bit_product():
    """test/rng_tests.jl:1-14"""
    rng = random.Random(1)
    for _ in range(20000):
        a, b = rng.getrandbits(64), rng.getrandbits(64)
        assert O.mulhi64(a, b) == (a * b) >> 64
    for a, b in ((0, 0), (M64, M64), (1, M64), (M64, 1)):
        assert O.mulhi64(a, b) == (a * b) >> 64


def test_chain_key_composition_and_dependence():
    """ChainRNG(seed, cycle, move) word i = splitmix64(splitmix64(splitmix64(seed, cycle), move), i)
    (src/rng.jl:43-61); every key component changes the stream (test/rng_tests.jl:16-31)."""
    w = O.chain_words(7, 3, 2, 4)
    key = O.splitmix64(O.splitmix64(7, 3), 2)
    assert [int(x) for x in w] == [O.splitmix64(key, i) for i in range(4)]
    b0 = int(w[0])
    assert int(O.chain_words(8, 3, 2, 1)[0]) != b0
    assert int(O.chain_words(7, 4, 2, 1)[0]) != b0
    assert int(O.chain_words(7, 3, 5, 1)[0]) != b0
    assert len(set(int(x) for x in w)) == 4
    # two's-complement reinterpretation of negative key components (src/rng.jl:48)
    assert int(O.chain_words(-1, 0, 0, 1)[0]) == int(O.chain_words(M64, 0, 0, 1)[0])


def test_unit_from_bits_exact():
    """(raw >> 11) * 2^-53 (src/rng.jl:70-71): exact multiples of 2^-53 in [0, 1)."""
    assert O.unit_f64(0) == 0.0
    assert O.unit_f64(M64) == 1.0 - 2.0**-53
    rng = random.Random(3)
    for _ in range(1000):
        r = rng.getrandbits(64)
        assert O.unit_f64(r) == (r >> 11) / 2.0**53
    assert O.unit_f32(M64) == np.float32(1.0 - 2.0**-24)


def test_rand_range_lemire():
    """1-based Lemire map mulhi64(raw, n) + 1 (src/rng.jl:93-97)."""
    rng = random.Random(4)
    for _ in range(2000):
        r, n = rng.getrandbits(64), rng.randint(1, 1000)
        v = O.rand_range(r, n)
        assert 1 <= v <= n and v == ((r * n) >> 64) + 1
    assert O.rand_range(M64, 1) == 1


def _ulp(x: float) -> float:
    return np.spacing(abs(x)) if x != 0 else 2.0**-1074


def test_sincospi_within_one_ulp_of_mpmath():
    """The build's sin(pi x)/cos(pi x) contract: <= 1 ulp from the exactly rounded value over
    [0, 2) (mpmath at 40 digits on the exact binary x), exact at the quadrant points."""
    mpmath.mp.dps = 40
    rng = np.random.default_rng(5)
    xs = list(rng.random(4000) * 2.0) + [0.0, 0.25, 0.5, 0.75, 1.0, 1.5, 1.0 - 2**-53, 2**-30, 0.125]
    worst = 0.0
    for x in xs:
        s, c = O.sincospi(float(x))
        es = float(mpmath.sinpi(mpmath.mpf(float(x))))
        ec = float(mpmath.cospi(mpmath.mpf(float(x))))
        worst = max(worst, abs(s - es) / _ulp(es), abs(c - ec) / _ulp(ec))
    assert worst <= 1.0, worst
    assert O.sincospi(0.5) == (1.0, -0.0) or O.sincospi(0.5)[0] == 1.0
    assert O.sincospi(0.0) == (0.0, 1.0)
    assert O.sincospi(1.0)[1] == -1.0 and O.sincospi(1.5)[0] == -1.0


def test_shoemake_unit_norm_and_convention():
    """Shoemake (x, y, z, w) = (sqrt(1-u1) cos 2pi u2, sqrt(u1) sin 2pi u3, sqrt(u1) cos 2pi u3,
    sqrt(1-u1) sin 2pi u2) (src/rng.jl:126-129): u = (0, 0, 0) is the unit x quaternion, and
    every draw has unit norm."""
    assert np.array_equal(O.shoemake(0.0, 0.0, 0.0), [1.0, 0.0, 0.0, 0.0])
    q = O.shoemake(1.0, 0.0, 0.25)
    assert np.allclose(q, [0.0, 1.0, 0.0, 0.0], atol=1e-16)
    rng = np.random.default_rng(6)
    for u in rng.random((2000, 3)):
        assert abs(np.linalg.norm(O.shoemake(*u)) - 1.0) < 4e-16


def test_pose_orientations_uniform_on_so3():
    """Rotation-invariance moments of rotated unit z under the Widom pose draws (the
    test/rng_tests.jl:125-189 check): mean 0 and second moments I/3 within 5 sigma."""
    A = np.diag([30.0, 30.0, 30.0])
    empty = O.OracleSystem(A, np.zeros((0, 3)), [], [], (1, 1, 1), [[3.0]], [[0.001]], 12.0, False,
                           [[0.0, 0.0, 0.0]], [0], [0.0], 12.0)
    n = 200_000
    _, poses = empty.energies(17, 0, 0, n, want_poses=True)
    q = poses[:, 3:]
    u, w = q[:, :3], q[:, 3:]
    v = np.array([0.0, 0.0, 1.0])
    t = np.cross(u, v) + w * v
    r = v + 2 * np.cross(u, t)
    assert np.allclose(np.linalg.norm(r, axis=1), 1.0, atol=1e-14)
    se1 = np.sqrt(1 / 3 / n)
    assert np.all(np.abs(r.mean(axis=0)) < 5 * se1)
    m2 = r.T @ r / n
    se2 = np.sqrt((1 / 5 - 1 / 9) / n)  # Var(x^2) for x uniform on the sphere's component
    assert np.all(np.abs(np.diag(m2) - 1 / 3) < 5 * se2)
    assert np.all(np.abs(m2[np.triu_indices(3, 1)]) < 5 * np.sqrt(1 / 15 / n))
    # positions are A f with f uniform in [0, 1)^3
    f = poses[:, :3] / 30.0
    assert np.all((f >= 0) & (f < 1))
    assert np.all(np.abs(f.mean(axis=0) - 0.5) < 5 * np.sqrt(1 / 12 / n))


def test_pose_key_layout():
    """Insertion j of system s draws f from words 0..2 and (u1, u2, u3) from words 3..5 of
    ChainRNG(splitmix64(seed, s+1), j+1, 0)."""
    A = np.array([[13.0, -6.5, 0.0], [0.0, 11.258330249197702, 0.0], [0.0, 0.0, 15.0]])
    for seed, s, j in ((42, 0, 0), (42, 3, 17), (2**63 + 5, 1, 123456)):
        w = O.chain_words(O.splitmix64(seed, s + 1), j + 1, 0, 6)
        u = [O.unit_f64(int(x)) for x in w]
        pos, q = O.pose(seed, s, j, A)
        f = np.array(u[:3])
        assert np.array_equal(pos, np.array([A[i, 0] * f[0] + A[i, 1] * f[1] + A[i, 2] * f[2] for i in range(3)]))
        assert np.array_equal(q, O.shoemake(*u[3:]))


@pytest.mark.parametrize("ninsert,nsys,run", [(37, 4, 3), (100, 7, 5), (11, 3, 2), (1, 1, 1), (500, 13, 1), (45, 2, 5)])
def test_system_counts_match_brute_force(ninsert, nsys, run):
    """src/widom.jl:34-53 against a brute-force count of sys_of_index (test/widom_tests.jl:209-220)."""
    counts = np.zeros(nsys, dtype=np.int64)
    for g in range(1, ninsert + 1):
        counts[((g - 1) // run) % nsys] += 1
    assert np.array_equal(O.system_counts(ninsert, nsys, run), counts)
    assert O.default_run(ninsert, nsys) == min(max((ninsert // nsys) // 4, 1), 256)
