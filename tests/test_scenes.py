"""Seeded workload generation (CPU)."""
import math

import numpy as np

from mpm_amd import scenes


def test_splitmix64_reference_sequence():
    # published output of SplitMix64 (Vigna, splitmix64.c) for seed 1234567
    assert scenes.splitmix64(1234567, 0, 5).tolist() == [
        6457827717110365317, 3203168211198807973, 9817491932198370423,
        4593380528125082431, 16408922859458223821]


def test_stream_is_counter_based():
    a = scenes.splitmix64(9, 0, 100)
    b = scenes.splitmix64(9, 40, 60)
    assert np.array_equal(a[40:], b)


def test_uniforms_are_exact_float32():
    u = scenes.uniform01(3, 0, 10000)
    assert u.dtype == np.float32 and u.min() >= 0 and u.max() < 1
    assert np.array_equal(u, (u * (1 << 24)).astype(np.int64) / np.float32(1 << 24))


def test_ball_sampling():
    s = scenes.Sampler(1)
    x = scenes.sample_ball(s, 20000, (0.5, 0.6, 0.5), 0.15).astype(np.float64)
    r = np.linalg.norm(x - (0.5, 0.6, 0.5), axis=1)
    assert r.max() <= 0.15 * (1 + 1e-6)
    # uniform by volume: P(r < R/2) = 1/8
    assert abs((r < 0.075).mean() - 0.125) < 0.01
    assert np.abs(x.mean(0) - (0.5, 0.6, 0.5)).max() < 0.003


def test_named_workloads_shapes():
    sc = scenes.elastic_sphere()
    assert sc.n == 10000 and sc.params.N == 16
    assert abs(sc.materials[0].particle_volume - 4 / 3 * math.pi * 0.15 ** 3 / 10000) < 1e-12
    sn = scenes.snow_spheres(n=50_000)
    n1, n2 = sn.extra["counts"]
    assert n1 + n2 == 50_000 and abs(n1 / n2 - (0.2 / 0.13) ** 3) < 0.01
    cube = scenes.elastic_cube(side=8)
    assert cube.n == 512 and cube.extra["logE"].shape == (512,)


def test_algorithmic_bytes():
    assert scenes.algorithmic_bytes(1_000_000, 128) == 254_000_000 + 64 * 128 ** 3
