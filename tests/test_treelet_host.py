"""The BVH builder's treelet restructuring (filianore-core_amd/csrc/fl_treelet.h,
run on the device level by level by k_treelet_level_lanes and k_treelet_level in
fl_build.hip) checked on the host: the same routines, compiled for the CPU with
a plain-memory policy (tables contiguous and lane-interleaved, subsets dealt
out over 64 lanes as the wave does), restructure binary
radix trees over synthetic primitive boxes, and every tree invariant is
checked afterwards (tests/native/treelet_check.cpp). The role is that of the
Embree build behind BVH::initialize_scene_geometry/finalise (bvh/bvh.cpp:7-47):
a valid tree over every triangle; the GPU side of the same invariants is
tests/test_gpu_bvh_structure.py."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "filianore-core_amd", "csrc")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("treelet") / "treelet_check")
    subprocess.check_call(["/opt/rocm/bin/hipcc", "-O2", "-std=c++17", "-I", CSRC,
                           os.path.join(HERE, "native", "treelet_check.cpp"), "-o", exe],
                          stderr=subprocess.DEVNULL)
    return exe


def run(exe, n, rounds, mode):
    r = subprocess.run([exe, str(n), str(rounds), str(mode)], capture_output=True, text=True, timeout=120)
    out = dict(line.split(" ", 1) for line in r.stdout.strip().splitlines())
    before, after = (float(x) for x in out["cost"].split())
    return r.returncode, int(out["violations"]), before, after


@pytest.mark.parametrize("n,rounds,mode", [(2, 1, 0), (3, 2, 1), (7, 2, 0), (8, 2, 1), (64, 3, 0), (5000, 2, 0),
                                           (20000, 2, 1), (1000, 2, 3), (20000, 2, 4)])
def test_restructured_tree_is_valid_and_no_worse(checker, n, rounds, mode):
    rc, bad, before, after = run(checker, n, rounds, mode)
    assert bad == 0 and rc == 0
    assert after <= before * (1 + 1e-5)


def test_restructuring_lowers_the_cost_of_a_shell_with_huge_boxes(checker):
    # a few boxes spanning the scene sit deep in the Morton tree and inflate
    # every ancestor; treelets lift them towards the root
    rc, bad, before, after = run(checker, 20000, 2, 4)
    assert bad == 0
    assert after < 0.8 * before
