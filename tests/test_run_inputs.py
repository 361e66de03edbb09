"""The reference's run inputs (tg_host_read_run_inputs; Read_data,
datread.C:36-451): files written here in the token order of the reference's
reader and GUI writer (titan_gui.py:1183-1340), read back and
non-dimensionalised as Read_data does."""
import math

import pytest

from discrete_adjoint_amd import problems


def _write(d, scale=True, test_point=True, materials=(("basalt", 32.0, 22.0),
                                                     ("ash", 30.0, 15.0))):
    sim = ["1", "2",
           "30 1500 2200 150 100 30 2.5 45",      # pile: h x y major minor rot speed dir
           "10 900 800 60 60 0 0 0",
           "20",                                  # cells across the smallest pile
           "2000", "600", "10", "100",            # maxiter maxtime timeoutput timesave
           "1", "0", "1",                         # adapt, viz, order
           "/gis", "loc", "PERMANENT", "dem", "0",
           "-1",                                  # edge height: default
           ("5 1200 2100" if test_point else "-2 x y"),
           "1", "0 0 100 100"]                    # one discharge plane
    (d / "simulation.data").write_text("\n".join(sim) + "\n")
    if scale:
        (d / "scale.data").write_text("1000\n10\n9.8\n")
    lines = [str(len(materials))]
    for name, phi_int, phi_bed in materials:
        lines += [name, f"{phi_int} {phi_bed}"]
    (d / "frict.data").write_text("\n".join(lines) + "\n")


def test_read_and_scale(tmp_path):
    _write(tmp_path)
    f, piles = problems.read_run_inputs(tmp_path)
    L, H, G = 1000.0, 10.0, 9.8
    assert (f["srctype"], f["npiles"], f["nsources"]) == (1, 2, 0)
    assert (f["length_scale"], f["height_scale"], f["gravity_scale"]) == (L, H, G)
    assert f["epsilon"] == H / L and f["time_scale"] == math.sqrt(L / G)
    p = piles[0]
    assert p["height"] == 30 / H and p["xc"] == 1500 / L and p["yc"] == 2200 / L
    assert p["major"] == 150 / L and p["minor"] == 100 / L and p["rot_deg"] == 30.0
    assert p["cosrot"] == math.cos(30 * math.pi / 180) and p["sinrot"] == math.sin(30 * math.pi / 180)
    v = math.sqrt(L * G)
    assert p["vx"] == 2.5 * math.cos(45 * math.pi / 180) / v
    assert p["vy"] == 2.5 * math.sin(45 * math.pi / 180) / v
    # the smallest axis: twice the smallest radius of any pile
    assert f["smallest_axis"] == 2.0 * 60 / L and f["cells_across"] == 20
    assert (f["maxiter"], f["maxtime"], f["timeoutput"], f["timesave"]) == (2000, 600, 10, 100)
    assert f["nd_maxtime"] == 600 / math.sqrt(L / G)
    assert f["nd_nextoutput"] == 10 / math.sqrt(L / G)
    assert (f["adaptflag"], f["vizflag"], f["order"]) == (1, 0, 1)
    assert (f["gis_main"], f["gis_sub"], f["gis_mapset"], f["gis_map"]) == \
        ("/gis", "loc", "PERMANENT", "dem")
    assert f["edge_height"] == H / 10000.0 * 10.0 / H
    assert f["test_height"] == 5 / H and f["test_x"] == 1200 / L and f["test_y"] == 2100 / L
    assert f["nplanes"] == 1
    assert f["nmat"] == 2 and f["matnames"] == ["basalt", "ash"]
    assert f["intfrict"] == 32.0 * math.pi / 180.0
    assert f["bedfrict"] == [22.0 * math.pi / 180.0, 15.0 * math.pi / 180.0]


def test_defaults_and_errors(tmp_path):
    _write(tmp_path, scale=False, test_point=False)
    f, piles = problems.read_run_inputs(tmp_path)
    assert (f["length_scale"], f["height_scale"], f["gravity_scale"]) == (1.0, 1.0, 1.0)
    assert piles[1]["height"] == 10.0 and f["test_height"] == float("inf")
    (tmp_path / "frict.data").unlink()
    with pytest.raises(ValueError, match="-4"):
        problems.read_run_inputs(tmp_path)
    _write(tmp_path)
    text = (tmp_path / "simulation.data").read_text().split("\n")
    (tmp_path / "simulation.data").write_text("\n".join(text[:6]))   # cut short
    with pytest.raises(ValueError, match="-5"):
        problems.read_run_inputs(tmp_path)
