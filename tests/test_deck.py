"""Deck generator, LS-DYNA/Control.dat text and geometry preprocessing."""
import os

import numpy as np

from taylor_amd import deck


def test_kfile_roundtrip(tmp_path):
    mesh = deck.taylor_bar_mesh(1.0e-3, length=4e-3, seed=1)
    path = os.path.join(tmp_path, "bar.k")
    deck.write_k(path, mesh)
    back = deck.read_k(path)
    assert np.array_equal(back.coo, mesh.coo)          # 16-wide fields carry every digit used
    assert np.array_equal(back.elcon, mesh.elcon)
    assert np.array_equal(back.sets[2], mesh.sets["WALL"])


def test_hex_volumes_sum_to_lattice_volume():
    mesh = deck.plate_mesh(1e-3, 6e-3, 5e-3, 4e-3, jitter=0.0)
    geo = deck.preprocess(mesh)
    assert abs(geo.vol.sum() - 6e-3 * 5e-3 * 4e-3) < 1e-15
    assert np.allclose(geo.win0, 1e-3)


def test_taylor_problem_sizes_and_bcs():
    prob = deck.taylor_problem(0.5e-3, seed=0)
    assert 9000 < prob.n < 12000                       # ~10k nodes at 0.5 mm
    wall = prob.sets["WALL"] - 1
    assert np.all(prob.ebc[wall, 2] == 1) and np.all(prob.vinit[wall, 2] == 0)
    assert np.all(prob.coo[wall, 2] == 0.0)
    assert abs(prob.dt - 0.3 * 0.5e-3 / deck.p_wave_speed(deck.copper_linear())) < 1e-20
    assert abs(deck.p_wave_speed(deck.copper_linear()) - 4585.6) < 1.0


def test_splitmix64_known_values():
    # first outputs of SplitMix64 seeded with 0 (Steele, Lea & Flood, reference C code)
    z = deck.splitmix64(0, 3)
    assert [int(v) for v in z] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
