"""The grid recipes against the reference's own example grids: each knot
file the reference ships (DBSR_HF/examples/Fe/Fe.knot, DBSR_MCHF/He/He.knot,
Rb_bound_example/dbsr_cc27/knot.dat, DBSR_HF/examples/Be_092m1.knot; copied
as data fixtures under tests/golden/knots) lists the knots its grid recipe
produced (grid_type 1: mkgrid_01 with a Fermi or point nucleus), printed
with 16 significant digits.  Rebuilding the grid from the file's
parameters, as def_grid does, must give those knots to the print's
rounding (1e-15 relative) and the same ns."""
import os

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN

FILES = ["Fe.knot", "He.knot", "Rb_cc27_knot.dat", "Be_092m1.knot"]


@pytest.mark.parametrize("name", FILES)
def test_knot_file_rebuilt_from_its_parameters(name):
    from dbsr3_amd.grid import knot_grid, read_knot
    path = os.path.join(GOLDEN, "knots", name)
    p = read_knot(path)
    t, ns, ks, ksp, ksq = knot_grid(path)
    assert (ns, ks) == (p["ns"], p["ks"])
    assert (ksp, ksq) == (p["ksp"], p["ksq"])
    listed = p["points"]
    assert np.all(np.abs(t - listed) <= 1e-15 * np.abs(listed)), np.max(np.abs(t - listed) / np.maximum(listed, 1e-300))
    # the oracle's recipe gives the same knots bit for bit
    to, no = O.mkgrid(1, ks, p["hi"], p["rmax"], p["atomic_number"], p["nuclear"], p["a_fermi"], p["r_uniform"],
                      p["ml"], p["he"], p["hmax"])
    assert no == ns and np.array_equal(to, t)


def test_fe_fixture_is_fe_knot():
    """tests/golden/fe_knot_t.txt (the knots every Fe-grid test uses) holds
    Fe.knot's listed points"""
    from dbsr3_amd.grid import read_knot
    p = read_knot(os.path.join(GOLDEN, "knots", "Fe.knot"))
    assert np.array_equal(np.loadtxt(os.path.join(GOLDEN, "fe_knot_t.txt")), p["points"])
