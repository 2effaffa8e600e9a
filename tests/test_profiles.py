"""The committed rocprofv3 summaries (profiles/) and the traffic table bench.py
reads (profiles/hbm_traffic.json) agree with each other and name only kernels
that exist in the sources."""
import glob
import json
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROF = os.path.join(ROOT, "profiles")


def _sources():
    files = glob.glob(os.path.join(ROOT, "pcr-penta_amd", "csrc", "*")) + [os.path.join(ROOT, "bench.py")]
    return "".join(open(f).read() for f in files)


def _base(name):
    return re.sub(r"<.*", "", name).split("::")[-1]


def test_traffic_table_matches_summaries():
    table = json.load(open(os.path.join(PROF, "hbm_traffic.json")))
    assert table
    for workload, e in table.items():
        md = os.path.join(PROF, e["profile"])
        assert os.path.exists(md), (workload, e["profile"])
        text = open(md).read()
        assert f"--workload {workload} " in text, workload
        dom = e["dominant_kernel"]
        assert f"`{dom['name']}`" in text, (workload, dom["name"])
        for key in ("bytes_per_step", "solver_us_per_step", "valu_insts_per_step",
                    "lds_bank_conflicts_per_step", "fetch_factor", "write_factor"):
            assert e.get(key) is not None, (workload, key)
        # calibrated bytes: FETCH_SIZE counts half a 128-byte line fill, WRITE_SIZE is exact
        assert 1.9 < e["fetch_factor"] < 2.1 and e["write_factor"] == 1.0, workload


def test_profiled_kernels_exist_in_sources():
    src = _sources()
    for md in glob.glob(os.path.join(PROF, "*.md")):
        for name in re.findall(r"`(pcrp::[^`]*?)`", open(md).read()):
            assert re.search(r"\b" + re.escape(_base(name)) + r"\b", src), (os.path.basename(md), name)
