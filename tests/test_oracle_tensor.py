"""Pins the oracle's Set::Matrix4 access and the cubic constructor.

* MajorMinor slot table against the reference's uid groups (Set/Matrix4_MajorMinor.H:327-375)
* isotropic / diagonal entries against dense tensors built from their definitions
* Euler rotation against scipy's independent intrinsic XZX rotation
* rotated cubic: R = I gives C11/C12/C44; a 90-degree cubic-group rotation leaves it unchanged;
  the stored slots are the last-written permutation of the rotated (non minor-symmetric) tensor
"""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from alamo_amd import problems as P

# uid groups of Set/Matrix4_MajorMinor.H:332-372, uid = i + 3j + 9k + 27l
UID_GROUPS = [[0], [27, 9, 3, 1], [54, 18, 6, 2], [36, 4], [63, 45, 7, 5], [72, 8], [30, 12, 28, 10],
              [57, 21, 33, 15, 55, 19, 29, 11], [39, 37, 31, 13], [66, 48, 64, 46, 34, 16, 32, 14],
              [75, 73, 35, 17], [60, 24, 56, 20], [42, 58, 22, 38], [69, 51, 61, 25, 65, 47, 59, 23],
              [78, 74, 62, 26], [40], [67, 49, 43, 41], [76, 44], [70, 52, 68, 50], [79, 77, 71, 53], [80]]


def test_mm_slot_table(ora):
    seen = set()
    for slot, group in enumerate(UID_GROUPS):
        for uid in group:
            i, j, k, l = uid % 3, (uid // 3) % 3, (uid // 9) % 3, uid // 27
            assert ora.mm_slot(i, j, k, l) == slot
            assert P.mm_slot(i, j, k, l) == slot
            seen.add(uid)
    assert seen == set(range(81))


def test_isotropic_entries(ora):
    lam, mu = 1.7, 0.6
    d = np.eye(3)
    dense = mu * (np.einsum("ik,jl->ijkl", d, d) + np.einsum("il,jk->ijkl", d, d)) + lam * np.einsum("ij,kl->ijkl", d, d)
    c = np.array([lam, mu])
    for idx in np.ndindex(3, 3, 3, 3):
        assert ora.c4(P.ISO, c, *idx) == pytest.approx(dense[idx], abs=0)


def test_diagonal_entries(ora):
    A = np.arange(1.0, 10.0)
    for i, j, k, l in np.ndindex(3, 3, 3, 3):
        want = A[3 * i + j] if (i == k and j == l) else 0.0
        assert ora.c4(P.DIAGONAL, A, i, j, k, l) == want


def test_euler_rotation_matches_scipy(ora):
    rng = np.random.default_rng(3)
    for _ in range(20):
        a = rng.uniform(0, 2 * np.pi, 3)
        ref = Rotation.from_euler("XZX", [a[2], a[1], a[0]]).as_matrix()
        assert np.abs(ora.euler_R(*a) - ref).max() < 1e-14
        assert np.abs(P.euler_R(*a) - ref).max() < 1e-14


def test_cubic_identity_rotation(ora):
    c = ora.cubic(168.3, 1.221, 0.757)
    assert c[P.mm_slot(0, 0, 0, 0)] == 168.3 and c[P.mm_slot(2, 2, 2, 2)] == 168.3
    assert c[P.mm_slot(0, 0, 1, 1)] == 1.221 and c[P.mm_slot(1, 1, 2, 2)] == 1.221
    assert c[P.mm_slot(0, 1, 0, 1)] == 0.757 and c[P.mm_slot(1, 2, 1, 2)] == 0.757
    assert c[P.mm_slot(0, 0, 0, 1)] == 0.0 and c[P.mm_slot(0, 1, 0, 2)] == 0.0


def test_cubic_group_invariance(ora):
    base = ora.cubic(1.68, 1.21, 0.75)
    for ax in ("x", "y", "z"):
        R = Rotation.from_euler(ax, 90, degrees=True).as_matrix()
        out = np.zeros(21)
        import ctypes
        ora.lib().ora_cubic_R(ctypes.c_double(1.68), ctypes.c_double(1.21), ctypes.c_double(0.75),
                              np.ascontiguousarray(R).ctypes.data_as(ora.D), out.ctypes.data_as(ora.D))
        assert np.abs(out - base).max() < 1e-12


def test_cubic_last_write_wins(ora):
    """The 21 slots hold the LAST permutation written in (p,q,s,t) loop order of the rotated
    tensor built from the reference's non minor-symmetric Ctmp."""
    C11, C12, C44 = 1.68, 1.21, 0.75
    Ct = np.zeros((3, 3, 3, 3))
    for i, j, k, l in np.ndindex(3, 3, 3, 3):
        if i == j == k == l:
            Ct[i, j, k, l] = C11
        elif i == k and j == l:
            Ct[i, j, k, l] = C44
        elif i == j and k == l:
            Ct[i, j, k, l] = C12
    a = (0.4, 1.3, 2.2)
    R = Rotation.from_euler("XZX", [a[2], a[1], a[0]]).as_matrix()
    full = np.einsum("pi,qj,sk,tl,ijkl->pqst", R, R, R, R, Ct)
    last = {}
    for p, q, s, t in np.ndindex(3, 3, 3, 3):
        last[P.mm_slot(p, q, s, t)] = full[p, q, s, t]
    got = ora.cubic(C11, C12, C44, *a)
    assert np.abs(got - np.array([last[k] for k in range(21)])).max() < 1e-13
    # and the Python generator used for the workloads agrees
    assert np.abs(P.cubic_mm(C11, C12, C44, R) - got).max() < 1e-13
    # the quirk is real: the stored tensor is not the rotated physical cubic tensor
    Cs = Ct.copy()
    for i, j in np.ndindex(3, 3):
        if i != j:
            Cs[i, j, j, i] = C44
    phys = np.einsum("pi,qj,sk,tl,ijkl->pqst", R, R, R, R, Cs)
    assert abs(phys[0, 1, 0, 1] - got[P.mm_slot(0, 1, 0, 1)]) > 1e-3


def test_splitmix64_known_answer():
    # first output of SplitMix64 seeded with 0 (Steele, Lea & Flood 2014 reference code)
    assert P.SplitMix64(0).next_u64() == 0xE220A8397B1DCDAF
    g = P.SplitMix64(12345)
    seq = [g.uniform() for _ in range(5)]
    assert np.array_equal(np.array(seq), P.splitmix_uniform_array(12345, 5))
