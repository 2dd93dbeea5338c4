"""Source-level guards of the HIP kernels (CPU only).

The doubled-K FP16 MFMAs of gfx950 (v_mfma_f32_16x16x32_f16 and its 32x32x16
sibling) produced wrong products when the compiler rewrote the upper half of
an A or B operand's registers a few instructions after issue: in the
four-wave fused training kernel every 16-record tile but the first came out
wrong, and the same build with each K=32 product issued as two K=16 MFMAs was
exact.  The kernels therefore build every FP16 product from K=16 MFMAs; this
test keeps the doubled-K builtins out of the sources."""

from __future__ import annotations

import re
from pathlib import Path

CSRC = Path(__file__).resolve().parents[1] / "nan_amd" / "csrc"
DOUBLED_K = re.compile(r"__builtin_amdgcn_mfma_f32_(16x16x32|32x32x16)_(f16|bf16)")


def test_no_doubled_k_fp16_mfma_in_kernel_sources():
    sources = sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.h"))
    assert sources
    offenders = [f"{p.name}:{i + 1}" for p in sources
                 for i, line in enumerate(p.read_text().splitlines()) if DOUBLED_K.search(line)]
    assert not offenders, f"doubled-K FP16 MFMA builtins in {offenders}"
