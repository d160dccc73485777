"""Register budgets of the hand-tuned gfx950 kernels, read from the assembly
hipcc writes for this build's sources (no GPU needed): the spectral products
kernels hold their accumulators and one item of prefetched operands in
registers, and a spill to scratch memory makes them several times slower;
the real FFTs of the transforms hold their 2L samples in registers."""
import re
import shutil
import subprocess
from pathlib import Path

import pytest

PKG = Path(__file__).resolve().parents[1] / "parallel3.0_omp_multidomin_1node_amd"
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


def _kernel_resources(src, tmp_path):
    if not Path(HIPCC).exists():
        pytest.skip("hipcc not available")
    out = tmp_path / "k.s"
    flags = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-Xclang",
             "-target-feature", "-Xclang", "-packed-fp32-ops", "-x", "hip", "--cuda-device-only", "-S"]
    subprocess.run([HIPCC, *flags, str(PKG / "csrc" / src), "-o", str(out)], check=True, capture_output=True,
                   timeout=600)
    text = out.read_text()
    res = {}
    for m in re.finditer(r"\.amdhsa_kernel (\S+)\n(.*?)\.end_amdhsa_kernel", text, re.S):
        name, body = m.group(1), m.group(2)
        scratch = int(re.search(r"\.amdhsa_private_segment_fixed_size (\d+)", body).group(1))
        vgpr = int(re.search(r"\.amdhsa_next_free_vgpr (\d+)", body).group(1))
        res[name] = (vgpr, scratch)
    return res


def test_spectral_products_kernels_do_not_spill(tmp_path):
    res = _kernel_resources("batched_fft.hip", tmp_path)
    prods = {k: v for k, v in res.items() if "k_conv_far_fft" in k}
    assert len(prods) == 6, res.keys()  # vector and matrix-core forms, sub-filter lengths 16, 32, 64
    for name, (vgpr, scratch) in prods.items():
        assert scratch == 0, (name, vgpr, scratch)
        assert vgpr <= 256, (name, vgpr)
    # the real FFTs of the default 16-lag level and of the 32-lag level run in registers
    ffts = {k: v for k, v in res.items()
            if any(f in k for f in ("k_fft_windows", "k_fft_inverse", "k_fft_filters")) and ("Li16E" in k or "Li32E" in k)}
    assert len(ffts) == 8, res.keys()
    for name, (vgpr, scratch) in ffts.items():
        assert scratch == 0, (name, vgpr, scratch)
