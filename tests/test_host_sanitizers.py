"""The host code (mesh builders, .lmesh reader) and the C oracle built with
AddressSanitizer and UndefinedBehaviorSanitizer and run on the fixtures,
malformed mesh files and short oracle runs (tests/host_sanitize_driver.cc).
CPU only: g++/gcc build the host sources directly, no GPU code is built."""
import gzip
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "lulesh-accelerate_amd", "csrc")
GOLDEN = os.path.join(ROOT, "tests", "golden")
SAN = ["-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
       "-fno-omit-frame-pointer"]


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None,
                    reason="needs g++ and gcc")
def test_host_code_under_sanitizers(tmp_path):
    oracle_o = tmp_path / "oracle.o"
    subprocess.run(["gcc"] + SAN + ["-fopenmp", "-ffp-contract=off", "-c",
                    os.path.join(ROOT, "oracle", "lulesh_oracle.c"), "-o", str(oracle_o)],
                   check=True, capture_output=True, timeout=300)
    exe = tmp_path / "driver"
    subprocess.run(["g++", "-std=c++17"] + SAN + [
        "-fopenmp", "-I" + CSRC, "-I" + os.path.join(ROOT, "oracle"),
        os.path.join(ROOT, "tests", "host_sanitize_driver.cc"),
        os.path.join(CSRC, "lulesh_mesh.cpp"), str(oracle_o), "-o", str(exe)],
        check=True, capture_output=True, timeout=300)
    big = tmp_path / "sedov21oct.lmesh"
    with gzip.open(os.path.join(GOLDEN, "sedov21oct.lmesh.gz"), "rb") as f, \
            open(big, "wb") as g:
        shutil.copyfileobj(f, g)
    good = open(os.path.join(GOLDEN, "sedov15oct.lmesh")).read()
    broken = {"truncated": good[: len(good) // 2],
              "empty": "",
              "garbage": "this is not a mesh\n",
              "negative": "-5 8\n" + good.split("\n", 1)[1]}
    # indices beyond int32 that wrap to a valid index when narrowed
    # (2^32 + 5 -> 5): the first node index, the first neighbour index and
    # the first entry of the x symmetry set
    tok = good.split()
    ne, nn = int(tok[0]), int(tok[1])
    first_node = 2 + 3 * nn
    first_nbr = first_node + 8 * ne
    first_set = first_nbr + 6 * ne + 1
    for name, pos in (("wrapped_node", first_node), ("wrapped_neighbour", first_nbr),
                      ("wrapped_set", first_set)):
        t = list(tok)
        t[pos] = str((1 << 32) + int(t[pos]))
        broken[name] = " ".join(t) + "\n"
    paths = [os.path.join(GOLDEN, "sedov15oct.lmesh"), str(big)]
    for name, text in broken.items():
        p = tmp_path / f"{name}.lmesh"
        p.write_text(text)
        paths.append(str(p))
    paths.append(str(tmp_path / "missing.lmesh"))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
               OMP_NUM_THREADS="2")
    out = subprocess.run([str(exe)] + paths, capture_output=True, text=True,
                         timeout=600, env=env)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-4000:]
    assert "host code clean" in out.stdout
    assert "sedov15oct.lmesh: ok 875 1126" in out.stdout
    assert "sedov21oct.lmesh: ok 31408 34446" in out.stdout
    assert "force blocks" in out.stdout and "FAILED" not in out.stdout
    for name in list(broken) + ["missing"]:
        assert f"{name}.lmesh: rejected" in out.stdout, name
