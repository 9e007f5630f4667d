"""CPU test of the split form's placement guard (gbsw_kernel_args.hpp:
placement_consistent, split_rerun_needed): the host reruns a job unsplit
exactly when two workgroups of one XCD work counter ran on different XCCs,
whatever the partition's number of XCCs."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("placement") / "placement_check")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", os.path.join(HERE, "host", "placement_check.cpp"), "-o", exe],
                   check=True)
    return exe


@pytest.mark.parametrize("mode,grid,rerun", [
    ("rr8", 2048, 0),    # 8 XCCs, workgroup w on XCC w % 8: the expected placement
    ("rr4", 1024, 0),    # 4 XCCs per partition: counters x and x + 4 share an XCC, still one per counter
    ("one", 256, 0),     # a one-XCC partition: every workgroup shares the L2
    ("rr8", 8, 0),
    ("shift", 2048, 1),  # the mapping changes half way through the grid
    ("swap", 2048, 1),   # a single workgroup on another XCC
])
def test_rerun_decision(checker, mode, grid, rerun):
    out = subprocess.run([checker, mode, str(grid)], capture_output=True, text=True, check=True).stdout
    assert int(out.strip()) == rerun
