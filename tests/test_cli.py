"""CLI behaviour, re-expressed from the reference's tests/test_ccs_cli.py: XYZ
parsing, radii, validation (CPU), and end-to-end solves on the GPU."""
import argparse
import csv
import json
from pathlib import Path

import numpy as np
import pytest

from ehss_amd.ccs_cli import parser, prepare_frame, radius_table, read_xyz, validate_args

EXAMPLES = Path(__file__).resolve().parent / "golden" / "examples"


def options(tmp_path, *extra):
    return parser().parse_args(["unused.xyz", "--rays", "1024", "--directions", "8", "--grid-size", "128",
                                "--repeats", "2", "--output", str(tmp_path / "result"), *extra])


def test_xyz_multiframe_comment_and_fortran_coordinates(tmp_path):
    p = tmp_path / "unicode molecule.xyz"
    p.write_text("2\n\nc 1D0 2 3\nCl 4 5 6\n\n1\nsecond\nO 0 0 0\n", encoding="utf-8-sig")
    frames = read_xyz(p)
    assert len(frames) == 2 and frames[0].comment == ""
    assert frames[0].elements == ("C", "Cl")
    np.testing.assert_array_equal(frames[0].coordinates, [[1, 2, 3], [4, 5, 6]])


@pytest.mark.parametrize("text", ["", "0\ncomment\n", "2\ncomment\nC 0 0 0\n", "1\nC 0 0 0\n", "1\nc\nC nan 0 0\n",
                                  "1\nc\n6 0 0 0\n", "1\nc\nC 0 0 0 1\n", "1\nProperties=species:S:1:pos:R:3\nC 0 0 0\n",
                                  "1\nc\nC 0 0 0\ntrailing garbage\n"])
def test_bad_xyz_rejected(tmp_path, text):
    p = tmp_path / "bad.xyz"
    p.write_text(text)
    with pytest.raises(ValueError):
        read_xyz(p)


def test_units_selection_and_explicit_radii(tmp_path):
    p = tmp_path / "input.xyz"
    p.write_text("3\nmixed\nC 10 20 30\nH 10.1 20 30\nZn 10 20.2 30\n")
    frame = read_xyz(p)[0]
    with pytest.raises(ValueError, match="Zn"):
        prepare_frame(frame, radius_table(), 1.4)
    override = tmp_path / "radii.json"
    override.write_text('{"Zn":1.39,"H":1.2}')
    s, info = prepare_frame(frame, radius_table(override), 1.4, "nm", True)
    assert len(s.radii) == 2 and info["selected_atom_rows_1based"] == [1, 3]
    np.testing.assert_allclose(s.radii, [3.1, 2.79])
    np.testing.assert_allclose(s.centers[1] - s.centers[0], [0, 2, 0], atol=1e-12)
    p.write_text("2\ncoincident\nC 0 0 0\nO 0 0 0\n")
    with pytest.raises(ValueError, match="coincident"):
        prepare_frame(read_xyz(p)[0], radius_table(), 1.4)


@pytest.mark.parametrize("data", [[], {"H": -1}, {"H": True}, {"H": float("nan")}, {"cl": 1.7, "Cl": 1.8}])
def test_bad_radii_rejected(tmp_path, data):
    p = tmp_path / "r.json"
    p.write_text(json.dumps(data))
    with pytest.raises(ValueError):
        radius_table(p)


@pytest.mark.parametrize("flag,value", [("--rays", "1000"), ("--grid-size", "4"), ("--directions", "3"),
                                        ("--repeats", "0"), ("--probe-radius", "nan"), ("--probe-radius", "-1"),
                                        ("--seed", "-1"), ("--max-collisions", "100001")])
def test_invalid_options(tmp_path, flag, value):
    with pytest.raises(ValueError):
        validate_args(options(tmp_path, flag, value))


def test_example_radii_override_file():
    table = radius_table(EXAMPLES / "radii_override.json")
    assert table["H"] == 1.20 and table["C"] == 1.70


@pytest.mark.gpu
def test_single_sphere_analytic_and_repeat_reproducibility(tmp_path):
    """tests/test_ccs_cli.py:59-70 with BVH PA."""
    from ehss_amd.ccs_cli import solve_frame
    from ehss_amd.state import Spheres
    args = options(tmp_path, "--pa-method", "bvh")
    s = Spheres(np.zeros((1, 3)), np.array([3.1]))
    a, b = solve_frame(s, args, False), solve_frame(s, args, False)
    truth = np.pi * 3.1 ** 2
    assert a["summary"]["ehss_angstrom2"]["mean"] == pytest.approx(truth, rel=2e-6)
    assert a["summary"]["pa_angstrom2"]["mean"] == pytest.approx(truth, rel=.003)
    assert a["summary"] == b["summary"]
    for r in a["runs"]:
        assert sum(r["ehss"]["collision_histogram"]) == args.rays
        assert r["ehss"]["unresolved_rays"] == 0 and r["ehss"]["tail_bound_angstrom2"] == 0


@pytest.mark.gpu
def test_cli_against_oracle_direct_trace(tmp_path, oracle):
    """tests/test_ccs_cli.py:73-85: cap 1 with a forced unresolved tail."""
    from ehss_amd.ccs_cli import solve_frame
    from ehss_amd.state import Spheres
    rng = np.random.default_rng(667)
    s = Spheres(rng.normal(size=(23, 3)) * 2, np.full(23, 1.2))
    args = options(tmp_path, "--pa-method", "bvh", "--repeats", "1", "--max-collisions", "1")
    result = solve_frame(s, args, False)["runs"][0]
    src = oracle.external_source(s.centers, s.radii, 10, args.seed)
    ref = oracle.readout(src, oracle.trace(oracle.static_tree(s.centers, s.radii, 4), s.centers, s.radii,
                                           src["origins"], src["directions"], cap=1), 1)
    assert ref["tail_bound"] > 0
    assert result["ehss"]["ehss_angstrom2"] == pytest.approx(ref["omega"], abs=1e-9)
    assert result["ehss"]["tail_bound_angstrom2"] == pytest.approx(ref["tail_bound"])
    assert result["pa"]["pa_angstrom2"] == pytest.approx(ref["pa"])


@pytest.mark.gpu
def test_execute_writes_outputs_and_refuses_overwrite(tmp_path):
    """tests/test_ccs_cli.py:88-103 plus the example inputs."""
    from ehss_amd.ccs_cli import execute
    p = tmp_path / "two frames.xyz"
    text = "1\nfirst\nC 0 0 0\n1\nsecond\nC 10 20 30\n"
    p.write_text(text)
    args = options(tmp_path, "--method", "both", "--pa-method", "bvh", "--all-frames", "--repeats", "1")
    args.xyz = [str(p), str(EXAMPLES / "water.xyz"), str(EXAMPLES / "single_carbon.xyz")]
    payload = execute(args)
    saved = json.loads((args.output / "results.json").read_text(encoding="utf-8"))
    assert saved == payload and len(saved["records"]) == 4
    assert saved["records"][0]["summary"]["ehss_angstrom2"]["sd"] is None
    assert saved["records"][0]["summary"] == saved["records"][1]["summary"]
    carbon = saved["records"][3]["summary"]["ehss_angstrom2"]["mean"]
    assert carbon == pytest.approx(30.1907, rel=1e-4)  # docs/CCS_CLI.md:185-190
    with (args.output / "summary.csv").open(encoding="utf-8-sig", newline="") as f:
        rows = list(csv.DictReader(f))
    assert len(rows) == 4
    assert p.read_text() == text
    before = (args.output / "results.json").read_bytes()
    with pytest.raises(ValueError, match="already exists"):
        execute(args)
    assert (args.output / "results.json").read_bytes() == before


@pytest.mark.gpu
def test_default_adaptive_pa_single_sphere(tmp_path):
    """tests/test_ccs_cli.py:59-70 with the default --pa-method adaptive."""
    from ehss_amd.ccs_cli import solve_frame
    from ehss_amd.state import Spheres
    args = options(tmp_path)
    assert args.pa_method == "adaptive"
    s = Spheres(np.zeros((1, 3)), np.array([3.1]))
    a = solve_frame(s, args, False)
    truth = np.pi * 3.1 ** 2
    assert a["summary"]["pa_angstrom2"]["mean"] == pytest.approx(truth, rel=0.01)
    assert a["summary"]["ehss_angstrom2"]["mean"] == pytest.approx(truth, rel=2e-6)
    assert a["summary"] == solve_frame(s, args, False)["summary"]


def test_main_exits_2_on_invalid_input(tmp_path, capsys):
    from ehss_amd.ccs_cli import main
    with pytest.raises(SystemExit) as exc:
        main([str(tmp_path / "missing*.xyz"), "--output", str(tmp_path / "out")])
    assert exc.value.code == 2 and "No XYZ files match" in capsys.readouterr().err
    (tmp_path / "out").mkdir()
    (tmp_path / "a.xyz").write_text("1\n\nC 0 0 0\n")
    with pytest.raises(SystemExit) as exc:
        main([str(tmp_path / "a.xyz"), "--output", str(tmp_path / "out")])
    assert exc.value.code == 2 and "already exists" in capsys.readouterr().err


def test_csv_row_leaves_unrequested_methods_blank():
    from ehss_amd.ccs_cli import CSV_COLUMNS, _csv_row
    rec = dict(input_file="m.xyz", frame=2, selected_atoms=7, timings=dict(all_replicates_with_setup_s=0.5),
               summary=dict(ehss_angstrom2=dict(mean=10.0, sd=None), ehss_tail_bound_mean_angstrom2=0.0,
                            ehss_unresolved_rays_total=0))
    row = _csv_row(rec)
    assert tuple(row) == CSV_COLUMNS
    assert row["pa_mean_A2"] is None and row["pa_sd_A2"] is None and row["ehss_mean_A2"] == 10.0
    assert row["atoms"] == 7 and row["frame"] == 2


def test_signed_zero_coordinates_are_coincident(tmp_path):
    p = tmp_path / "z.xyz"
    p.write_text("3\n\nC 0 0 0\nO 1 1 1\nN -0.0 0 -0\n")
    with pytest.raises(ValueError, match="coincident"):
        prepare_frame(read_xyz(p)[0], radius_table(), 1.4)
