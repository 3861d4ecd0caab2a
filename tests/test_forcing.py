"""Helical forcing parameters on the host: the oracle's restatement pinned by known answers and
the reference's own assertions (ref src/utils/helical_forcing.cc), then the library's C-ABI
generator (seed_rng, generateHelicalForcingParams, random_uniform_real_01,
loadForcingParamsToMeshInfo) against the oracle, bit for bit. Host code only: no GPU."""
import numpy as np
import pytest

import oracle.oracle as O
from astaroth_amd import astaroth as ac


def test_mt19937_known_answer():
    # The C++ standard fixes the 10000th draw of a default-seeded (5489) std::mt19937.
    O.seed_rng(5489)
    for _ in range(9999):
        O.mt19937_draw()
    assert O.mt19937_draw() == 4123659995


@pytest.mark.parametrize("kmin,kmax,count", [(4.5, 5.5, 350), (0.8, 1.2, 6), (1.0, 2.0, 32),
                                             (2.0, 3.0, 96)])
def test_shell_population(kmin, kmax, count):
    # 350 is the reference's own assertion (helical_forcing.cc gen_k); the others are the
    # integer points of the shells, counted here by brute force.
    brute = sum(1 for x in range(-6, 7) for y in range(-6, 7) for z in range(-6, 7)
                if kmin <= np.sqrt(x * x + y * y + z * z) <= kmax)
    assert brute == count
    assert O.shell_size(kmin, kmax) == count


@pytest.mark.parametrize("relhel", [0.0, 0.3, 1.0, -1.0])
def test_oracle_forcing_vectors(relhel):
    O.seed_rng(2024)
    for _ in range(50):
        p = O.helical_forcing_params(relhel, 4.5, 5.5)
        k, fre, fim = p["k"], p["f_re"], p["f_im"]
        kn = np.linalg.norm(k)
        assert 4.5 <= kn <= 5.5 and np.all(k == np.round(k))
        # |f|^2 = 1 and f perpendicular to k (solenoidal forcing)
        assert abs(fre @ fre + fim @ fim - 1.0) < 1e-12
        assert abs(fre @ k) < 1e-12 * kn and abs(fim @ k) < 1e-12 * kn
        if relhel == 0.0:
            assert np.all(fim == 0.0)
        if relhel == 1.0:
            # maximally helical: the reference's assertions
            np.testing.assert_allclose(-np.cross(k, fim), kn * fre, atol=1e-12)
            np.testing.assert_allclose(np.cross(k, fre), kn * fim, atol=1e-12)
        assert -np.pi <= p["phase"] < np.pi
        assert p["kaver"] == (5.5 - 4.5) / 2 + 4.5


def test_oracle_draws_are_uniform():
    # the sampled k cover the shell: over 3500 draws every one of the 350 vectors shows up
    # with the right frequency (chi-square loose bound)
    O.seed_rng(7)
    counts = {}
    for _ in range(3500):
        k = tuple(O.helical_forcing_params(0.0, 4.5, 5.5)["k"])
        counts[k] = counts.get(k, 0) + 1
    assert len(counts) == 350
    c = np.array(list(counts.values()), dtype=float)
    chi2 = float(((c - 10.0) ** 2 / 10.0).sum())
    assert chi2 < 500.0


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("relhel,kmin,kmax", [(0.0, 4.5, 5.5), (0.5, 4.5, 5.5), (1.0, 1.0, 3.0),
                                              (-0.7, 0.8, 1.2)])
def test_library_generator_matches_oracle(prec, relhel, kmin, kmax):
    lib = ac.load(prec)
    rt = lib.np_real
    for seed in (1, 312256655):
        lib.seed_rng(seed)
        O.seed_rng(seed)
        for _ in range(20):
            got = lib.generateHelicalForcingParams(rt(relhel), rt(0.07), rt(kmin), rt(kmax))
            # the library receives AcReal arguments and widens them to double, as the
            # reference's generator does
            want = O.helical_forcing_params(float(rt(relhel)), float(rt(kmin)), float(rt(kmax)))
            v = lambda s: np.array([s.x, s.y, s.z])
            assert np.array_equal(v(got.k_force), want["k"].astype(rt))
            assert np.array_equal(v(got.ff_hel_re), want["f_re"].astype(rt))
            assert np.array_equal(v(got.ff_hel_im), want["f_im"].astype(rt))
            assert got.phase == rt(want["phase"])
            assert got.kaver == (rt(kmax) - rt(kmin)) / rt(2) + rt(kmin)
            assert got.magnitude == rt(0.07)


def test_library_uniform01_and_meshinfo():
    lib = ac.load("fp64")
    lib.seed_rng(99)
    draws = [lib.random_uniform_real_01() for _ in range(1000)]
    assert all(0.0 <= d < 1.0 for d in draws)
    assert abs(np.mean(draws) - 0.5) < 0.05
    # one uniform draw per call from the shared generator: the forcing generator continues it
    lib.seed_rng(5)
    lib.random_uniform_real_01()
    a = lib.generateHelicalForcingParams(0.2, 1.0, 4.5, 5.5)
    O.seed_rng(5)
    O.mt19937_draw()
    O.mt19937_draw()  # a double draw of std::uniform_real_distribution takes two 32-bit words
    b = O.helical_forcing_params(0.2, 4.5, 5.5)
    assert a.phase == b["phase"]
    info = lib.config((16, 16, 16))
    assert lib.loadForcingParamsToMeshInfo(a, info) == 0
    assert info["AC_forcing_magnitude"] == a.magnitude and info["AC_forcing_phase"] == a.phase
    assert (info["AC_k_forcex"], info["AC_k_forcey"], info["AC_k_forcez"]) == \
        (a.k_force.x, a.k_force.y, a.k_force.z)
    assert (info["AC_ff_hel_rex"], info["AC_ff_hel_rey"], info["AC_ff_hel_rez"]) == \
        (a.ff_hel_re.x, a.ff_hel_re.y, a.ff_hel_re.z)
    assert (info["AC_ff_hel_imx"], info["AC_ff_hel_imy"], info["AC_ff_hel_imz"]) == \
        (a.ff_hel_im.x, a.ff_hel_im.y, a.ff_hel_im.z)
    assert info["AC_kaver"] == a.kaver


def test_forcing_config_defaults():
    # ref config/astaroth.conf:34-35,58-61; forcing off unless asked for
    lib = ac.load("fp64")
    info = lib.config((16, 16, 16))
    assert info["AC_forcing_period_steps"] == 1
    assert info["AC_forcing_magnitude"] == 1e-5 and info["AC_relhel"] == 0.0
    assert info["AC_kmin"] == 0.8 and info["AC_kmax"] == 1.2
    assert info["AC_lforcing"] == 0
