"""The DLTC driver equations table (dtcf_amd/data/dltc_drivers.json) against
the reference's driver sources where they are present, and its
self-consistency (labels match ranks, every statement is run once)."""
import os
import subprocess
import sys

import pytest

from dtcf_amd import dltc_driver

REF = "/root/reference/DLTC"


@pytest.mark.parametrize("name", dltc_driver.DRIVERS)
def test_table_consistent(name):
    t = dltc_driver.load(name)
    run = [(k, i) for layer in t["layers"] for k, i in layer]
    assert sorted(i for k, i in run if k == "MUL") == list(range(len(t["muls"])))
    assert sorted(i for k, i in run if k == "SUM") == list(range(len(t["sums"])))
    for st in t["muls"] + t["sums"]:
        for key in ("c", "a", "b"):
            if key in st:
                kinds = t["tensors"][st[key]]
                idx = st[key + "_idx"]
                assert len(kinds) == len(idx)
                for k, x in zip(kinds, idx):
                    assert (k == "HOL") == x.startswith("h"), (name, st)


def test_table_regenerates_from_reference(tmp_path):
    if not os.path.isdir(REF):
        pytest.skip("reference tree not present")
    here = os.path.join(os.path.dirname(dltc_driver.__file__), "data")
    script = os.path.join(here, "extract_dltc_drivers.py")
    shipped = open(os.path.join(here, "dltc_drivers.json")).read()
    out = str(tmp_path / "drivers.json")
    r = subprocess.run([sys.executable, script, REF, out], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert open(out).read() == shipped
