"""The Fortran host (taylor_amd/host/mega_drive, bind(C) interfaces of
include/mega.h) drives the same device step as the Python host: identical
final positions and plastic strain on the same model."""
import os
import subprocess

import numpy as np
import pytest

from taylor_amd import deck, mega

pytestmark = pytest.mark.gpu
DRIVE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "taylor_amd", "host", "mega_drive")


@pytest.mark.skipif(not os.path.exists(DRIVE), reason="Fortran host not built")
@pytest.mark.parametrize("precision", [mega.MEGA_FP64, mega.MEGA_MIXED])
def test_fortran_host_equals_python_host(tmp_path, precision):
    prob = deck.taylor_problem(0.7e-3, length=6e-3, seed=0)
    model, result = str(tmp_path / "model.bin"), str(tmp_path / "result.bin")
    mega.write_model(model, prob, nsteps=30, chunk=10, precision=precision)
    out = subprocess.run([DRIVE, model, result], capture_output=True, text=True, timeout=120)
    print(out.stdout)
    assert out.returncode == 0, out.stderr
    r = np.fromfile(result, dtype=np.float64)
    x_f, eps_f = r[:3 * prob.n].reshape(prob.n, 3), r[3 * prob.n:]
    g = mega.Mega(prob, precision=precision)
    for _ in range(3):
        g.step(10)
    x, eps = g.download("x"), g.download("eps")
    g.close()
    assert np.array_equal(x, x_f) and np.array_equal(eps, eps_f)


def write_deck(tmp_path, prob, mesh):
    kfile = str(tmp_path / "bar.k")
    deck.write_k(kfile, mesh)
    deck.write_control(str(tmp_path / "Control.dat"), kfile, prob, list(mesh.sets.keys()), time_output=1.0)
    return str(tmp_path / "Control.dat")


def reference_step_count(dt, time_end):
    """MAIN.F90:644-656: TIME = TIME + DLT after each step, exit once TIME > TIME_END."""
    t, k = 0.0, 0
    while True:
        k += 1
        t = t + dt
        if t > time_end:
            return k


@pytest.mark.skipif(not os.path.exists(DRIVE), reason="Fortran host not built")
@pytest.mark.parametrize("lagrangian,precision,mat_type", [(1, mega.MEGA_FP64, 2), (1, mega.MEGA_MIXED, 2),
                                                           (0, mega.MEGA_FP64, 2), (1, mega.MEGA_FP64, 6),
                                                           (1, mega.MEGA_FP64, 3)])
def test_fortran_host_reads_the_reference_input(tmp_path, lagrangian, precision, mat_type):
    """The Fortran host reads the reference's own input -- Control.dat and the
    LS-DYNA model file it names -- preprocesses the geometry on the device and
    runs the reference's number of steps; the Python host builds the same model
    from the generator's arrays.  Final positions and plastic strain are equal
    bit for bit (J2, and the damage types 6 and 3, whose Drucker-Prager cone
    both hosts prepare from the deck's friction angle and cohesion)."""
    from damage_util import damage_material
    h, length = 0.7e-3, 6e-3
    mat = None if mat_type == 2 else damage_material(mat_type)
    prob = deck.taylor_problem(h, length=length, seed=0, lagrangian=lagrangian, pdsearch_steps=5, material=mat)
    prob.time_end = 40.5 * prob.dt
    mesh = deck.taylor_bar_mesh(h, length=length, seed=0)
    ctl = write_deck(tmp_path, prob, mesh)
    nsteps = reference_step_count(prob.dt, prob.time_end)
    result = str(tmp_path / "result.bin")
    out = subprocess.run([DRIVE, "--deck", ctl, result, str(precision), "0", "16"], capture_output=True,
                         text=True, timeout=120)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    assert f"steps {nsteps}" in out.stdout
    r = np.fromfile(result, dtype=np.float64)
    x_f, eps_f = r[:3 * prob.n].reshape(prob.n, 3), r[3 * prob.n:]
    g = mega.Mega(prob, precision=precision)
    done = 0
    while done < nsteps:
        k = min(16, nsteps - done)
        g.step(k)
        done += k
    x, eps = g.download("x"), g.download("eps")
    g.close()
    assert np.array_equal(x, x_f) and np.array_equal(eps, eps_f)


@pytest.mark.skipif(not os.path.exists(DRIVE), reason="Fortran host not built")
def test_fortran_host_reference_input_reaches_the_reference_final_state(tmp_path):
    """The deck of the CPU-runnable bar (spacing 0.5 mm, 80 us), written as the
    fixture's generator wrote it for the reference program
    (tests/golden/make_golden.py --final), read and run to the end by the
    Fortran host on the GPU in FP64: final length and mushroom radius against
    the reference program's own final state."""
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "taylor_h05_final.npz"))
    mesh = deck.taylor_bar_mesh(0.5e-3, seed=0)
    prob = deck.taylor_problem(0.5e-3, seed=0)
    nsteps = int(z["nsteps"][0])
    prob.time_end = prob.dt * (nsteps - 0.5)
    ctl = write_deck(tmp_path, prob, mesh)
    result = str(tmp_path / "result.bin")
    out = subprocess.run([DRIVE, "--deck", ctl, result, str(mega.MEGA_FP64), "0", "500"], capture_output=True,
                         text=True, timeout=300)
    print(out.stdout[-400:])
    assert out.returncode == 0, out.stdout + out.stderr
    assert f"steps {nsteps}" in out.stdout
    x = np.fromfile(result, dtype=np.float64)[:3 * prob.n].reshape(prob.n, 3)
    wall = prob.sets["WALL"] - 1
    L, L0 = x[:, 2].max(), z["x"][:, 2].max()
    R = np.sqrt(x[wall, 0] ** 2 + x[wall, 1] ** 2).max()
    R0 = np.sqrt(z["x"][wall, 0] ** 2 + z["x"][wall, 1] ** 2).max()
    print(f"length rel {abs(L - L0) / L0:.2e}, radius rel {abs(R - R0) / R0:.2e}")
    assert abs(L - L0) <= 1e-8 * L0 and abs(R - R0) <= 1e-8 * R0


def _numbers(line):
    try:
        return [float(t) for t in line.split()]
    except ValueError:
        return None


def compare_frames(ours, ref, rtol):
    """Line by line: text records equal; numeric records within rtol of the
    largest magnitude of their block (integers exactly).  Returns the largest
    relative difference seen."""
    lo, lr = ours.splitlines(), ref.splitlines()
    assert len(lo) == len(lr), (len(lo), len(lr))
    scale, block, blk_of = {}, "POINTS", []
    for line in lr:
        if line.startswith("SCALARS ") or line.startswith("POINTS ") or line.startswith("CELLS") or \
                line.startswith("CELL_TYPES"):
            block = line.split()[1] if line.startswith("SCALARS ") else line.split()[0]
        blk_of.append(block)
        v = _numbers(line)
        if v:
            scale[block] = max(scale.get(block, 0.0), max(abs(x) for x in v))
    worst = 0.0
    for a, b, blk in zip(lo, lr, blk_of):
        va, vb = _numbers(a), _numbers(b)
        if va is None or vb is None or not vb:
            assert a.rstrip() == b.rstrip(), (a, b)
            continue
        assert len(va) == len(vb), (a, b)
        if all(("." not in t and "E" not in t) for t in b.split()):
            assert va == vb, (blk, a, b)
            continue
        s = max(scale.get(blk, 0.0), 1e-300)
        d = max(abs(x - y) for x, y in zip(va, vb)) / s
        assert d <= rtol, (blk, a, b, d)
        worst = max(worst, d)
    return worst


@pytest.mark.skipif(not os.path.exists(DRIVE), reason="Fortran host not built")
def test_fortran_host_writes_the_reference_frames(tmp_path):
    """With 'unformatted' among the output variables the Fortran host writes
    the reference's VTK frames on the reference's schedule: against the
    reference program's own frames of the same deck
    (tests/golden/vtk_frames.npz), every text record equal and every value
    within 1e-9 of its field's largest magnitude (FP64)."""
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    from make_golden import vtk_deck
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "vtk_frames.npz"))
    prob = vtk_deck(str(tmp_path))
    out = subprocess.run([DRIVE, "--deck", str(tmp_path / "Control.dat"), str(tmp_path / "r.bin")],
                         capture_output=True, text=True, timeout=120)
    print(out.stdout[-600:])
    assert out.returncode == 0, out.stdout + out.stderr
    # the metrics line against the same quantities taken from the result file
    n = prob.vol.shape[0]
    res = np.fromfile(tmp_path / "r.bin", dtype=np.float64)
    x, eps = res[:3 * n].reshape(n, 3), res[3 * n:]
    fixed = prob.ebc[:, 2] == 1
    want = (x[:, 2].max() - x[:, 2].min(), np.hypot(x[fixed, 0], x[fixed, 1]).max(),
            np.sqrt((prob.vol * eps ** 2).sum() / prob.vol.sum()))
    line = [l for l in out.stdout.splitlines() if l.startswith("metrics: length")][0].split()
    got = (float(line[2]), float(line[4]), float(line[6]))
    print("metrics", got, want)
    assert np.allclose(got, want, rtol=1e-13, atol=0) and want[2] > 0
    for name, text in zip(z["names"], z["texts"]):
        ours = open(tmp_path / str(name)).read()
        worst = compare_frames(ours, str(text), 1e-9)
        print(name, "largest relative difference", worst)
    assert not os.path.exists(tmp_path / "MEGAOUT_0003.vtk")


@pytest.mark.skipif(not os.path.exists(DRIVE), reason="Fortran host not built")
def test_fortran_host_automatic_time_step_follows_the_reference_program(tmp_path):
    """A peridynamic deck with IAUTO_TS = 1 (DLT_FAC 0.3, the small mm bar of
    tests/golden/pd_auto.npz, TIME_END as the reference program ran it): the
    Fortran host sums TIME from the device's DLT of every step, stops after
    the reference program's 40 steps, and ends at its TIME and positions."""
    a = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "pd_auto.npz"))
    h, length = 0.7, 6.0   # mm
    prob = deck.taylor_problem(0.7e-3, length=6e-3, seed=0, pdsearch_steps=3, peridynamics=1, mm_units=True)
    prob.time_end = float(a["run_time_end"][0])
    mesh = deck.taylor_bar_mesh(h, radius=3.8, length=length, seed=0)
    assert np.array_equal(mesh.coo, prob.coo)
    kfile = str(tmp_path / "bar.k")
    deck.write_k(kfile, mesh)
    ctl = str(tmp_path / "Control.dat")
    deck.write_control(ctl, kfile, prob, ["BAR", "WALL"], time_output=1.0, dlt_fac=float(a["run_dlt_fac"][0]))
    result = str(tmp_path / "result.bin")
    out = subprocess.run([DRIVE, "--deck", ctl, result, "0", "0", "16"], capture_output=True, text=True, timeout=120)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    line = [ln for ln in out.stdout.splitlines() if ln.startswith("automatic time step:")][0]
    steps, t = int(line.split()[3]), float(line.split()[6].rstrip(","))
    r = np.fromfile(result, dtype=np.float64)
    x = r[:3 * prob.n].reshape(prob.n, 3)
    e_t = abs(t - a["run_time"][-1]) / a["run_time"][-1]
    e_x = float(np.abs(x - a["step40_x"]).max() / np.abs(a["step40_x"]).max())
    print(f"steps {steps}, TIME {t:.17e} (reference {a['run_time'][-1]:.17e}, {e_t:.1e}), positions {e_x:.1e}")
    assert steps == 40 == len(a["run_time"])
    assert e_t < 1e-14 and e_x < 1e-12

