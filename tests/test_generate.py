"""The seeded synthetic meshes of the benchmark configurations (host code, no GPU)."""
import numpy as np

from conftest import load_package

gen = load_package().generate


def test_splitmix64_known_values():
    # SplitMix64 of seed 0: the published first outputs of the generator (Vigna's reference code).
    r = gen.SplitMix64(0)
    assert [r.next_u64() for _ in range(3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_icosphere_counts_and_radius():
    for lvl in range(4):
        p, f = gen.icosphere(lvl)
        assert p.shape == (10 * 4**lvl + 2, 3)
        assert f.shape == (20 * 4**lvl, 3)
        assert np.allclose(np.linalg.norm(p, axis=1), 1.0, atol=1e-15)
        # every edge shared by exactly two faces (closed manifold)
        e = np.sort(np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]]), axis=1)
        _, cnt = np.unique(e, axis=0, return_counts=True)
        assert np.all(cnt == 2)


def test_torus_counts():
    p, f = gen.torus(128, 64, 1.0, 0.4)
    assert p.shape == (8192, 3) and f.shape == (16384, 3)


def test_configs_are_deterministic():
    a, fa = gen.config_icosphere_l3()
    b, fb = gen.config_icosphere_l3()
    assert np.array_equal(a, b) and np.array_equal(fa, fb)
    assert a.shape == (642, 3) and fa.shape == (1280, 3)
    d = np.linalg.norm(a, axis=1) - 1.0
    assert 0.003 < d.std() < 0.03
