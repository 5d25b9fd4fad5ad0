"""The Fortran host's reader of the reference's input (taylor_amd/host/
mega_deck.f90) refuses malformed decks with a message and exit status 2,
before any device work, as the reference's EXIT_PROGRAM(…, 2) does for user
errors (READ.F90, PRE_CONTROL.F90, ASSIGN_TO_NODE.f90)."""
import os
import subprocess

import pytest

from taylor_amd import deck

DRIVE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "taylor_amd", "host", "mega_drive")


def write_deck(tmp_path, edit=None):
    h, length = 0.7e-3, 3e-3
    prob = deck.taylor_problem(h, length=length, seed=0)
    mesh = deck.taylor_bar_mesh(h, length=length, seed=0)
    kfile = str(tmp_path / "bar.k")
    deck.write_k(kfile, mesh)
    ctl = str(tmp_path / "Control.dat")
    deck.write_control(ctl, kfile, prob, list(mesh.sets.keys()), time_output=1.0)
    if edit:
        txt = open(ctl).read()
        open(ctl, "w").write(edit(txt))
    return ctl


def run(ctl, tmp_path):
    return subprocess.run([DRIVE, "--deck", ctl, str(tmp_path / "r.bin")], capture_output=True, text=True,
                          timeout=60)


@pytest.mark.skipif(not os.path.exists(DRIVE), reason="Fortran host not built")
@pytest.mark.parametrize("edit,message", [
    (lambda t: t.replace("*TIME PARAMETERS", "*TIME PARAMS"), "card *TIME PARAMETERS not found"),
    (lambda t: t.replace("*MATERIALS", "*MATERIAL"), "material definitions were not found"),
    (lambda t: t.replace("#SET\nWALL\n#FIXITY", "#SET\nBAR\n#FIXITY"), None),
    (lambda t: t + "*BOUNDARY CONDITION\n#SET\nWALL\n#FIXITY\n1 0 0\n", "multiple values of fixity"),
    (lambda t: t.replace("#SET\nWALL\n#FIXITY", "#SET\nNOSUCHSET\n#FIXITY"), "set name not found"),
    (lambda t: t.replace("#IAUTO_TS\n  0", "#IAUTO_TS\n  1"), "IAUTO_TS = 1 needs *PERIDYNAMICS 1"),
    (lambda t: t.replace("#IAUTO_TS\n  0", "#IAUTO_TS\n  2"), "choose IAUTO_TS = (0,1)"),
    (lambda t: t.replace("#TYPE\n2\n", "#TYPE\nJ2\n"), "error reading MAT_TYPE"),
])
def test_reader_refuses_malformed_decks(tmp_path, edit, message):
    ctl = write_deck(tmp_path, edit)
    out = run(ctl, tmp_path)
    if message is None:
        # fixity 0 0 1 on the whole bar instead of the wall: a valid deck the reader
        # accepts; only the device step is left, which this container cannot run
        assert "deck error" not in out.stdout
        return
    assert out.returncode == 2, out.stdout + out.stderr
    assert message in out.stdout


@pytest.mark.skipif(not os.path.exists(DRIVE), reason="Fortran host not built")
def test_reader_refuses_missing_model_file(tmp_path):
    ctl = write_deck(tmp_path, lambda t: t.replace("bar.k", "nothere.k"))
    out = run(ctl, tmp_path)
    assert out.returncode == 2 and "cannot open the model file" in out.stdout
