"""cubismup_3d_amd.regions files kernels under the roctx range that enclosed their launch call,
on a database with the layout of rocprofv3's (regions and kernels). CPU only."""
import sqlite3

from cubismup_3d_amd import regions


def test_kernels_are_filed_under_the_enclosing_range(tmp_path):
    db = str(tmp_path / "t.db")
    c = sqlite3.connect(db)
    c.execute("create table regions (category text, name text, stack_id int, parent_stack_id int, extdata text)")
    c.execute("create table kernels (name text, stack_id int, duration int, start int)")
    c.executemany("insert into regions values (?,?,?,?,?)", [
        ("MARKER_CORE_RANGE_API", "roctxThreadRangeA", 1, 0, '{"message":"AdvDiff Kernel"}'),
        ("MARKER_CORE_RANGE_API", "roctxThreadRangeA", 2, 0, '{"message":"GradP"}'),
        ("MARKER_CORE_RANGE_API", "roctxThreadRangeA", 5, 0, '{"message":"AdvDiff Kernel"}'),
        ("HIP_RUNTIME_API_EXT", "hipLaunchKernel", 10, 1, "{}"),
        ("HIP_RUNTIME_API_EXT", "hipLaunchKernel", 11, 2, "{}"),
        ("HIP_RUNTIME_API_EXT", "hipLaunchKernel", 12, 5, "{}"),
        ("HIP_RUNTIME_API_EXT", "hipLaunchKernel", 13, 0, "{}"),
    ])
    c.executemany("insert into kernels values (?,?,?,?)", [
        ("void cup::k_adv<float>(float*)", 10, 1000, 1), ("void cup::k_gradp<float>(float*)", 11, 3000, 2),
        ("void cup::k_adv<float>(float*)", 12, 2000, 3), ("void cup::k_max(double*)", 13, 500, 4),
        ("__amd_rocclr_copyBuffer", 99, 100, 5)])
    c.commit()
    per, opened = regions.summarize(db)
    assert per["AdvDiff Kernel"]["cup::k_adv<float>"] == (2, 3000)
    assert per["GradP"]["cup::k_gradp<float>"] == (1, 3000)
    assert per["(outside any region)"]["cup::k_max"] == (1, 500)
    assert per["(no launch record)"]["__amd_rocclr_copyBuffer"] == (1, 100)
    assert opened["AdvDiff Kernel"] == 2
    assert "| AdvDiff Kernel | 2 | `cup::k_adv<float>` | 2 | 1.5 | 0.003 |" in regions.table(db)
