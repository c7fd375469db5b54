"""Writes the golden fixtures of tests/golden/ from binaries built from the
reference sources (oracle/Makefile -> oracle/_ref/).

Each fixture is the full final state of one reference run, dumped by
oracle/ref_dump*.cc: the reference's own output on the Sedov problem, so
the GPU tests can check against it on a machine without the reference tree.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import lulesh_oracle as lo  # noqa: E402

# (name, mode, nx, its, extra 2.0 options)
RUNS = [
    ("ref10_s5", 1, 5, None, ()),
    ("ref10_s10", 1, 10, None, ()),
    ("ref20_s5", 2, 5, None, ()),
    ("ref20_s10", 2, 10, None, ()),
    ("ref20_s12_i60_r7_b2_c3", 2, 12, 60, ("-r", "7", "-b", "2", "-c", "3")),
    ("ref20_s16_i40", 2, 16, 40, ()),
]


def main() -> None:
    lo.build()
    meta = {}
    with tempfile.TemporaryDirectory() as td:
        for name, mode, nx, its, extra in RUNS:
            dump = os.path.join(td, name + ".bin")
            out, st = lo.run_reference(mode, nx, its=its, threads=4,
                                       dump_path=dump, extra_args=extra)
            summary = lo.parse_reference_output(out)
            arrays = {k: v for k, v in st.fields.items()}
            arrays["regNumList"] = st.reg_num_list
            np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
            meta[name] = {
                "mode": mode, "nx": nx, "its": its, "extra": list(extra),
                "cycle": st.cycle, "time": st.time,
                "deltatime": st.deltatime, "dtcourant": st.dtcourant,
                "dthydro": st.dthydro,
                "printed": {k: v for k, v in summary.items() if k != "elapsed"},
            }
            print(name, st.cycle, summary.get("origin_energy"))
    with open(os.path.join(HERE, "reference_runs.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
