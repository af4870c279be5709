"""Build-level checks of the gfx950 kernels (CPU only: hipcc cross-compiles).

Register spills are silent on the GPU but cost a large share of a kernel's
time: the radix-8 NTT pass once went from 20 to 652 bytes of scratch per lane
after a change that only moved its body into a helper, and ran 3x slower.
This test compiles kernels.hip for the device and bounds every kernel's
scratch use."""
import re
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
SRC = ROOT / "learn-zkp_amd" / "csrc" / "kernels.hip"
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

# bytes of scratch per lane allowed; every other kernel must use none.  The
# three-wave build of the radix-8 pass (at most 168 VGPRs) trades a spill of
# 180 bytes for occupancy and is used only where the autotuner measures it
# faster.  The product sumcheck's single-block rounds call their transcript
# step out of line (prod_transcript_call): the scratch is that call's frame,
# the registers it saves once per round; inlined, the step and the table
# work share one allocation and spill instead, at the same measured speed.
ALLOWED = {"k_ntt_multi3_w3": 192, "k_prod_tail": 224}


@pytest.fixture(scope="module")
def scratch(tmp_path_factory):
    if not Path(HIPCC).exists():
        pytest.skip("hipcc not available")
    out = tmp_path_factory.mktemp("kbuild") / "kernels.o"
    cmd = [HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", "--cuda-device-only", "-c", str(SRC), "-o", str(out),
           "-Rpass-analysis=kernel-resource-usage"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    names = re.findall(r"Function Name: (\S+)", r.stderr)
    sizes = [int(v) for v in re.findall(r"ScratchSize \[bytes/lane\]: (\d+)", r.stderr)]
    assert names and len(names) == len(sizes)
    return dict(zip(names, sizes))


def test_kernels_do_not_spill(scratch):
    bad = {}
    for name, size in scratch.items():
        limit = next((v for k, v in ALLOWED.items() if k in name), 0)
        if size > limit:
            bad[name] = (size, limit)
    assert not bad, f"kernels over their scratch bound (bytes/lane, bound): {bad}"


def test_hot_kernels_present(scratch):
    for k in ["k_nodes", "k_subtree", "k_leaves_elem", "k_leaves_pair", "k_ntt_multiILi3", "k_fold", "ks_subtree",
              "ks_ntt_multiILi3"]:
        assert any(k in n for n in scratch), k
