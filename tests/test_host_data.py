"""Host layer on the CPU: the conservation sums (oracle and product host path)
against closed forms, the G1/Zeno loaders on the reference's fixtures, and the
synthetic workload generators."""
import math

import numpy as np
import pytest

import oracle_lib
from conftest import REFDATA
from nbody_amd import nbody_data


def test_stats_two_body_circular_orbit():
    """Equal masses m at (+-a, 0, 0) on a circular orbit: v^2 = G m / (4 a)."""
    m, a = 0.5, 0.5
    v = math.sqrt(m / (4 * a))
    pos = np.array([[a, 0, 0], [-a, 0, 0]], dtype=float)
    vel = np.array([[0, v, 0], [0, -v, 0]], dtype=float)
    mass = np.array([m, m])
    for st in (oracle_lib.stats(pos, vel, mass, True), _product_stats(pos, vel, mass)):
        assert np.allclose(st[0:3], 0.0, atol=1e-17)                              # P
        assert st[5] == pytest.approx(2 * m * a * v, rel=1e-15) and st[3] == st[4] == 0   # L_z
        assert st[6] == pytest.approx(m * v * v, rel=1e-15)                      # KE
        assert st[7] == pytest.approx(-m * m / (2 * a), rel=1e-15)               # PE
        assert np.allclose(st[8:11], 0.0, atol=1e-17) and st[11] == 1.0          # CM, M


def _product_stats(pos, vel, mass):
    d = nbody_data()
    d.set_quiet(True)
    d.set_bodies(pos, vel, mass)
    d.set_check_list("PLEV")
    return d.statistics(None)


def test_product_host_stats_equal_oracle():
    d = nbody_data()
    d.set_quiet(True)
    d.generate("plummer", 300, 3)
    d.set_check_list("PLEV")
    pos, vel, m, _ = d.bodies()
    assert np.array_equal(d.statistics(None), oracle_lib.stats(pos, vel, m, True))


def test_pe_close_pairs_contribute_zero():
    pos = np.array([[0, 0, 0], [1e-5, 0, 0], [3, 4, 0]], dtype=float)
    vel = np.zeros((3, 3))
    m = np.array([1.0, 1.0, 2.0])
    st = oracle_lib.stats(pos, vel, m, True)
    pe = -(1 * 2) / 5.0 - (1 * 2) / math.dist(pos[1], pos[2])
    assert st[7] == pytest.approx(pe, rel=1e-15)


def test_load_initial_state_fixture():
    d = nbody_data()
    assert d.load(str(REFDATA / "initial_state.txt"))
    assert d.get_count() == 16
    pos, vel, m, r = d.bodies()
    assert pos[0, 1] == 50.0 and m[0] == 999.0 and r[0] == 33.0 and r[1] == 44.0


@pytest.mark.parametrize("name", ["inv_column.txt", "inv_columns_count.txt", "inv_radius.txt"])
def test_load_rejects_invalid(name):
    d = nbody_data()
    d.set_quiet(True)
    assert not d.load(str(REFDATA / name))


def test_load_zeno():
    d = nbody_data()
    assert d.load_initial(str(REFDATA / "zeno_ascii.txt"), "Zeno")
    assert d.get_count() > 0
    assert not d.load_initial(str(REFDATA / "zeno_ascii.txt"), "Nope")


def test_save_load_round_trip(tmp_path):
    d = nbody_data()
    d.set_quiet(True)
    d.generate("plummer", 100, 1)
    f = tmp_path / "s.txt"
    assert d.save(str(f))
    e = nbody_data()
    assert e.load(str(f))
    assert e.is_equal(d, 0.0)


def test_generators():
    d = nbody_data()
    d.set_quiet(True)
    d.generate("ball", 4096, 1)
    pos, vel, m, _ = d.bodies()
    assert np.all(np.linalg.norm(pos, axis=1) <= 1.0)
    assert abs(vel.std() - 0.3) < 0.01 and np.allclose(m, 1 / 4096)

    d.generate("plummer", 1 << 16, 2)
    pos, vel, m, _ = d.bodies()
    r = np.linalg.norm(pos, axis=1)
    assert r.max() <= 10.0 and m.sum() == pytest.approx(1.0)
    for rr in (0.5, 1.0, 2.0):       # Plummer cumulative mass M(r) = r^3 / (1 + r^2)^(3/2), truncated at 10
        frac = rr ** 3 / (1 + rr * rr) ** 1.5 / (1000 / 101 ** 1.5)
        assert abs(np.mean(r < rr) - frac) < 0.01
    # virial ratio of a Plummer sphere ~ 1 (2 KE / |PE| with PE = -3 pi / 32)
    ke = 0.5 * np.sum(m * np.sum(vel ** 2, axis=1))
    assert abs(2 * ke / (3 * math.pi / 32) - 1.0) < 0.05
    for ax in range(3):
        assert len(np.unique(pos[:, ax])) == len(pos)      # no tied coordinates

    d.generate("two_plummer", 4096, 3)
    pos, vel, m, _ = d.bodies()
    assert abs(np.median(pos[:2048, 0]) + 5) < 0.2 and abs(np.median(pos[2048:, 0]) - 5) < 0.2
    assert abs(vel[:2048, 0].mean() - 0.25) < 0.05 and abs(vel[2048:, 0].mean() + 0.25) < 0.05

    d.generate("cube", 4096, 5)
    pos, vel, m, _ = d.bodies()
    assert np.all(np.abs(pos) <= 1.0) and np.all(vel == 0)
    assert m.sum() == pytest.approx(1.0, rel=1e-12)
    assert abs(np.log(m).std() - 0.5) < 0.03

    e = nbody_data()
    e.generate("cube", 4096, 5)
    assert e.is_equal(d, 0.0)          # seeded: identical bodies
    with pytest.raises(ValueError):
        e.generate("spiral", 8, 1)
