"""The profile summaries under profiles/ are made by
hastyrenderer_amd.rocprof_summary from rocprofv3's CSV files: check its
arithmetic on tiny hand-written files in that format."""
import os

from hastyrenderer_amd import rocprof_summary as rs


def _write(path, header, rows):
    with open(path, "w") as f:
        f.write(",".join(header) + "\n")
        for r in rows:
            f.write(",".join(str(x) for x in r) + "\n")


def test_stats_table(tmp_path):
    st = os.path.join(tmp_path, "s.csv")
    tr = os.path.join(tmp_path, "t.csv")
    _write(st, ["Name", "Calls", "TotalDurationNs", "AverageNs", "Percentage"],
           [['"void hasty::(anonymous namespace)::kExtend<false, false>(hasty::JobSet, hasty::PathBuffers)"', 4,
             5000000, 1250000, 62.5],
            ["kAdvance(unsigned int*)", 4, 3000000, 750000, 37.5]])
    _write(tr, ["Kernel_Name", "VGPR_Count", "SGPR_Count", "LDS_Block_Size", "Scratch_Size"],
           [['"void hasty::(anonymous namespace)::kExtend<false, false>(hasty::JobSet, hasty::PathBuffers)"', 64, 80,
             16384, 0]])
    out = rs.stats(st, tr).splitlines()
    assert out[0].startswith("| kernel | calls | total ms | avg us |")
    assert out[2] == "| void kExtend<false, false> | 4 | 5.000 | 1250.0 | 62.5 | 64 | 80 | 16384 | 0 |"
    assert out[3] == "| kAdvance | 4 | 3.000 | 750.0 | 37.5 |  |  |  |  |"


def test_pmc_mean_per_dispatch(tmp_path):
    pc = os.path.join(tmp_path, "c.csv")
    hdr = ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]
    _write(pc, hdr, [[1, "kExtend", "FETCH_SIZE", 10.0], [2, "kExtend", "FETCH_SIZE", 30.0],
                     [3, "kShade", "FETCH_SIZE", 5.0], [3, "kShade", "WRITE_SIZE", 7.0]])
    out = rs.pmc(pc, "FETCH_SIZE", ["kExtend"])
    assert "| kExtend | 2 | 20.0 | 40.0 |" in out
    assert "kShade" not in out
