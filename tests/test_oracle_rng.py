"""Pins the oracle's (and the HIP walker's) random stream.

Philox4x32-10 against the Random123 known-answer vectors (Salmon et al.,
SC'11, kat_vectors) and against rocRAND's open-source host implementation
(/opt/rocm/include/rocrand/rocrand_philox4x32_10.h); the Box-Muller normals
against their distribution.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

KAT = [
    # counter (4 x u32), key (2 x u32), expected output (4 x u32)
    ((0x00000000, 0x00000000, 0x00000000, 0x00000000), (0x00000000, 0x00000000),
     (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff), (0xffffffff, 0xffffffff),
     (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,expected", KAT)
def test_philox_known_answers(oracle_lib, ctr, key, expected):
    out = oracle_lib.philox4x32_10(ctr, key)
    assert tuple(int(v) for v in out) == expected


@pytest.mark.skipif(not os.path.exists("/opt/rocm/bin/hipcc"), reason="needs hipcc for rocRAND")
def test_philox_matches_rocrand(oracle_lib, tmp_path):
    here = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")
    exe = os.path.join(here, "philox_xcheck")
    if not os.path.exists(exe):
        subprocess.check_call(["make", "-s", "-C", here, "philox_xcheck"])
    rng = np.random.default_rng(3)
    cases = rng.integers(0, 2**32, size=(200, 6), dtype=np.uint64)
    inp = "\n".join(" ".join(f"{int(v):x}" for v in row) for row in cases) + "\n"
    out = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True).stdout.split("\n")
    for row, line in zip(cases, out):
        ref = tuple(int(x, 16) for x in line.split())
        got = tuple(int(v) for v in oracle_lib.philox4x32_10(row[:4], row[4:]))
        assert got == ref


def test_normals_distribution(oracle_lib):
    n = oracle_lib.normals(seed=12345, stream=0, spin=7, step=0, draw0=0, n=400000).ravel()
    assert abs(n.mean()) < 5e-3
    assert abs(n.var() - 1) < 1e-2
    assert abs(np.mean(n ** 4) / np.mean(n ** 2) ** 2 - 3) < 5e-2
    # 23-bit uniforms cap |n| at sqrt(-2 ln 2^-24) ~ 5.77
    assert np.abs(n).max() < 5.8


def test_normals_are_keyed(oracle_lib):
    """Different (spin, step, draw, stream) coordinates give independent
    draws; the same coordinates give the same draw."""
    a = oracle_lib.normals(1, 0, 5, 9, 0, 1)
    b = oracle_lib.normals(1, 0, 5, 9, 0, 1)
    c = oracle_lib.normals(1, 0, 6, 9, 0, 1)
    d = oracle_lib.normals(1, 1, 5, 9, 0, 1)
    e = oracle_lib.normals(2, 0, 5, 9, 0, 1)
    assert np.array_equal(a, b)
    for x in (c, d, e):
        assert not np.allclose(a, x)


def test_uniform_grid_is_exact(oracle_lib):
    """The uniforms are ((x >> 9) + 1/2) 2^-23, exactly representable in
    binary32: the device and the oracle see the same values."""
    x = np.arange(0, 2**32, 2**20 + 12345, dtype=np.uint64)
    u = ((x >> 9).astype(np.float64) + 0.5) * 2.0 ** -23
    assert np.array_equal(u.astype(np.float32).astype(np.float64), u)
    assert u.min() > 0 and u.max() < 1


def test_walk_normals_bit_identical_to_product(oracle_lib):
    """The walker's normals (host build of the product's nz_rng.h) equal the
    oracle's restatement bit for bit (binary32, correctly rounded operations
    only), across seeds, spins, steps and draw indices."""
    from tests.cpu_build import lib as cpu_lib
    L = cpu_lib()
    rng = np.random.default_rng(3)
    for _ in range(40):
        seed = int(rng.integers(0, 2**63))
        spin, step, draw0 = (int(v) for v in rng.integers(0, 2**31, 3))
        n = 256
        got = np.zeros((n, 3), dtype=np.float32)
        L.nzcpu_normals(seed, 0, spin, step, draw0, n, got.ctypes.data)
        ref = oracle_lib.normals(seed, 0, spin, step, draw0, n)
        assert np.array_equal(got.astype(np.float64), ref)


def test_walk_normals_accuracy(oracle_lib):
    """The binary32 polynomial log and sin/cos against FP64 libm on the same
    uniforms: the normals are within a few float ulps of exact Box-Muller."""
    n = 20000
    out = oracle_lib.normals(12345, 0, 7, 9, 0, n)
    # regenerate the uniforms: Philox words -> 23-bit grid
    words = np.array([oracle_lib.philox4x32_10([k, 9, 7, 0], [12345 & 0xFFFFFFFF, 12345 >> 32])
                      for k in range(n)], dtype=np.uint64)
    u = ((words >> 9).astype(np.float64) + 0.5) / 2**23
    r0 = np.sqrt(-2 * np.log(u[:, 0]))
    r1 = np.sqrt(-2 * np.log(u[:, 2]))
    exact = np.stack([r0 * np.cos(2 * np.pi * u[:, 1]), r0 * np.sin(2 * np.pi * u[:, 1]),
                      r1 * np.cos(2 * np.pi * u[:, 3])], 1)
    err = np.abs(out - exact) / np.maximum(1.0, np.abs(exact))
    assert err.max() < 2e-6, err.max()
