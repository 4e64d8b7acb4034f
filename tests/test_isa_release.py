"""The bottom-up refit (vierkant_amd/csrc/bvh.hip, k_refit) hands each node's box from one thread to another,
possibly on another XCD, through an acq_rel atomic on the node's arrival counter. The release half needs the
box stores and the L2 write-back (buffer_wbl2) to complete before the atomic is performed; hipcc leaves that
wait out when the store and the atomic meet across a loop back edge, so the kernel spells it out. This
compiles bvh.hip for gfx950 and checks the emitted code: between every box store and the next arrival
atomic there is a write-back followed by s_waitcnt vmcnt(0). A compiler that reorders or drops it fails here
instead of producing an intermittently stale parent box (a BVH that culls the closest hit)."""
import pathlib
import shutil
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent


def refit_isa(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not pathlib.Path(hipcc).exists():
        pytest.skip("hipcc is not available")
    out = tmp_path / "bvh.s"
    cmd = [hipcc, "-O3", "-std=c++20", "--offload-arch=gfx950", "-ffp-contract=off", "-Xclang", "-target-feature",
           "-Xclang", "-packed-fp32-ops", f"-I{ROOT / 'include'}", f"-I{ROOT / 'vierkant_amd' / 'include'}",
           f"-I{ROOT / 'vierkant_amd' / 'csrc'}", "-x", "hip", "--offload-device-only", "-S",
           str(ROOT / "vierkant_amd" / "csrc" / "bvh.hip"), "-o", str(out)]
    subprocess.run(cmd, check=True, capture_output=True)
    lines = out.read_text().split("\n")
    start = next(i for i, l in enumerate(lines) if l.startswith("_ZN4vkrt7k_refit"))
    end = next(i for i in range(start, len(lines)) if lines[i].startswith(".Lfunc_end"))
    return [l.strip() for l in lines[start:end] if l.startswith("\t") and not l.strip().startswith((".", ";"))]


def test_refit_release_waits_for_the_box_write_back(tmp_path):
    isa = refit_isa(tmp_path)
    atomics = [i for i, l in enumerate(isa) if l.startswith("global_atomic_add")]
    assert atomics and any(l.startswith("global_store") for l in isa)
    for a in atomics:
        # the straight run of instructions that leads to the atomic (back to the previous branch, memory
        # access or label): whichever block precedes it at run time, the stores before it are complete only
        # if this run holds a write-back followed by a vmcnt(0) wait
        k = a - 1
        while k >= 0 and not isa[k].startswith(("s_branch", "s_cbranch", "global_", "buffer_load", "flat_")):
            k -= 1
        run = isa[k + 1:a]
        wb = [i for i, l in enumerate(run) if l.startswith("buffer_wbl2")]
        assert wb, f"no L2 write-back before the arrival atomic: {run}"
        assert any(l.startswith("s_waitcnt") and "vmcnt(0)" in l for l in run[wb[0] + 1:]), \
            f"the write-back is not waited for before the arrival atomic: {run}"
        # the acquire half: the result is waited for and the caches invalidated before the child loads
        tail = isa[a + 1:a + 4]
        assert any("vmcnt(0)" in l for l in tail) and any(l.startswith("buffer_inv") for l in tail), tail
