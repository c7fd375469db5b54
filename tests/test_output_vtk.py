"""writeDomain's replacement (lulesh_write_domain_arrays): a host-only
function of the library, checked on a state of the oracle by reading the
VTK file back."""
import numpy as np


def read_vtk(path):
    raw = open(path, "rb").read()
    pos = 0

    def line():
        nonlocal pos
        end = raw.index(b"\n", pos)
        s = raw[pos:end].decode()
        pos = end + 1
        return s

    def take(n, dtype):
        nonlocal pos
        a = np.frombuffer(raw, dtype=dtype, count=n, offset=pos)
        pos += a.nbytes
        return a

    out = {}
    assert line().startswith("# vtk DataFile")
    out["title"] = line()
    assert line() == "BINARY"
    assert line() == "DATASET UNSTRUCTURED_GRID"
    assert line() == "FIELD FieldData 2"
    assert line() == "TIME 1 1 double"
    out["time"] = float(take(1, ">f8")[0])
    line()
    assert line() == "CYCLE 1 1 int"
    out["cycle"] = int(take(1, ">i4")[0])
    line()
    tag, n, _ = line().split()
    out["points"] = take(3 * int(n), ">f8").reshape(-1, 3)
    line()
    tag, nc, size = line().split()
    out["cells"] = take(int(size), ">i4").reshape(int(nc), 9)
    line()
    assert line().split()[0] == "CELL_TYPES"
    out["types"] = take(int(nc), ">i4")
    line()
    assert line() == f"CELL_DATA {nc}"
    for name in ("e", "p", "v", "q"):
        assert line() == f"SCALARS {name} double 1"
        assert line() == "LOOKUP_TABLE default"
        out[name] = take(int(nc), ">f8")
        line()
    assert line() == f"POINT_DATA {n}"
    assert line() == "VECTORS velocity double"
    out["velocity"] = take(3 * int(n), ">f8").reshape(-1, 3)
    return out


def test_vtk_roundtrip(lulesh_gpu, oracle, tmp_path):
    o = oracle.Oracle(5, mode=2, threads=2)
    o.run(12)
    st = {k: o.get(k) for k in ("x", "y", "z", "xd", "yd", "zd", "e", "p", "v", "q")}
    path = str(tmp_path / "dom.vtk")
    lulesh_gpu.write_domain_arrays(path, 12, o.scalar("time"), st)
    got = read_vtk(path)
    assert got["cycle"] == 12 and got["time"] == o.scalar("time")
    assert np.array_equal(got["points"], np.stack([st["x"], st["y"], st["z"]], 1))
    assert np.array_equal(got["velocity"], np.stack([st["xd"], st["yd"], st["zd"]], 1))
    for k in ("e", "p", "v", "q"):
        assert np.array_equal(got[k], st[k]), k
    assert np.all(got["types"] == 12) and np.all(got["cells"][:, 0] == 8)
    # the connectivity is the oracle's element -> node list
    nodelist = o.get_int("nodelist").reshape(-1, 8)
    assert np.array_equal(got["cells"][:, 1:], nodelist)
