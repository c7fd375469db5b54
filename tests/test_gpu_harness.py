"""The C++ harness links the C ABI and prints the reference's verification
block; its numbers match the published LULESH 2.0 answer at 30^3."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HARNESS = os.path.join(ROOT, "lulesh-accelerate_amd", "lulesh_mi355x")
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def run(args, cwd=None):
    out = subprocess.run([HARNESS] + args, capture_output=True, text=True,
                         timeout=300, cwd=cwd)
    assert out.returncode == 0, out.stderr
    return out.stdout


def test_harness_30_cubed_2_0():
    import lulesh_oracle
    res = lulesh_oracle.parse_reference_output(run(["-s", "30"]))
    assert res["cycles"] == 932
    assert f"{res['origin_energy']:12.6e}" == f"{2.025075e+05:12.6e}"
    assert abs(res["max_rel_diff"] - 1.461e-12) <= 1e-12


def test_harness_1_0_and_cycle_limit():
    import lulesh_oracle
    res = lulesh_oracle.parse_reference_output(run(["-s", "10", "--ic", "1"]))
    assert res["cycles"] == 593
    assert f"{res['origin_energy']:12.6e}" == f"{5.174622e+05:12.6e}"
    res = lulesh_oracle.parse_reference_output(run(["-s", "20", "-i", "25"]))
    assert res["cycles"] == 25


def test_harness_rejects_bad_options():
    out = subprocess.run([HARNESS, "-s", "abc"], capture_output=True, text=True)
    assert out.returncode != 0


def test_harness_unstructured_mesh(tmp_path):
    import lulesh_oracle
    mesh = os.path.join(ROOT, "tests", "golden", "sedov15oct.lmesh")
    vtk = str(tmp_path / "end.vtk")
    res = lulesh_oracle.parse_reference_output(run(["-u", mesh, "--vtk", vtk]))
    o = lulesh_oracle.Oracle.from_mesh(lulesh_oracle.parse_lmesh(open(mesh).read()),
                                       cbrt_mode=1, threads=8)
    o.run()
    assert res["cycles"] == o.cycle
    e0 = lulesh_oracle.lib().lo_scalar(o.h, b"e_origin")
    assert f"{res['origin_energy']:12.6e}" == f"{e0:12.6e}"
    assert os.path.getsize(vtk) > 0


def test_harness_viz_dump(tmp_path):
    """-v writes the reference's dump name for the final cycle (-f is taken
    and one file written)."""
    run(["-s", "6", "-i", "12", "-v", "-f", "4", "-q"], cwd=str(tmp_path))
    assert os.path.getsize(tmp_path / "lulesh_plot_c12.vtk") > 0
