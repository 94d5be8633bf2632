"""CPU: the Fourier loader's row-broadcast multiply-adds are free of the DPP
read-after-VALU-write hazard.

`fma_row_bcast` (magnus-dp_amd/csrc/kernels_wbd8.hip) issues
`v_fmac_f64_dpp` from inline assembly.  A DPP instruction must not read a
VGPR that a VALU instruction wrote fewer than two wait states before it, and
the compiler's hazard pass does not look inside inline assembly, so the
invariant is checked on the code the compiler emits: the DPP source of every
`v_fmac_f64_dpp` (the table rows, written by vector loads) is not the
destination of a VALU instruction within the two preceding wait states.
Compiles the file for gfx950 to assembly (about half a minute) when hipcc
is present.
"""
from __future__ import annotations

import re
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
CSRC = ROOT / "magnus-dp_amd" / "csrc"


def _vgprs(op: str) -> set[int]:
    m = re.match(r"v\[(\d+):(\d+)\]", op)
    if m:
        return set(range(int(m.group(1)), int(m.group(2)) + 1))
    m = re.match(r"v(\d+)$", op)
    return {int(m.group(1))} if m else set()


def hazards(asm: str) -> tuple[int, list[str]]:
    """(count of v_fmac_f64_dpp, hazards found) in an assembly listing."""
    ins = [ln.strip() for ln in asm.splitlines() if ln.startswith("\t") and not ln.strip().startswith((".", ";"))]
    found, total = [], 0
    for idx, text in enumerate(ins):
        if not text.startswith("v_fmac_f64_dpp"):
            continue
        total += 1
        src0 = _vgprs(text.split(None, 1)[1].split(",")[1].strip())
        waits, j = 0, idx - 1
        while j >= 0 and waits < 2:
            prev = ins[j]
            mnem = prev.split()[0]
            if mnem == "s_nop":
                waits += int(prev.split()[1]) + 1
            else:
                if mnem.startswith("v_") and " " in prev:
                    dst = _vgprs(prev.split(None, 1)[1].split(",")[0].strip())
                    if dst & src0:
                        found.append(f"{prev} -> {text}")
                waits += 1
            j -= 1
    return total, found


def test_hazard_scan_flags_a_planted_write():
    planted = "\tv_mov_b64_e32 v[4:5], 0\n\tv_fmac_f64_dpp v[0:1], v[4:5], v[2:3] row_newbcast:3\n"
    clean = "\tglobal_load_dwordx2 v[4:5], v6, s[0:1]\n\ts_waitcnt vmcnt(0)\n\tv_fmac_f64_dpp v[0:1], v[4:5], v[2:3] row_newbcast:3\n"
    waited = "\tv_mov_b64_e32 v[4:5], 0\n\ts_nop 1\n\tv_fmac_f64_dpp v[0:1], v[4:5], v[2:3] row_newbcast:3\n"
    assert hazards(planted)[1] and not hazards(clean)[1] and not hazards(waited)[1]


@pytest.mark.skipif(shutil.which("hipcc") is None and not Path("/opt/rocm/bin/hipcc").exists(), reason="no hipcc")
def test_fourier_loader_row_broadcasts_have_no_dpp_hazard(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    out = tmp_path / "wbd8.s"
    subprocess.run([hipcc, "-O3", "-std=c++17", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
                    "-packed-fp32-ops", f"-I{CSRC}", f"-I{ROOT / 'include'}", "--cuda-device-only", "-S", "-o",
                    str(out), str(CSRC / "kernels_wbd8.hip")], check=True, capture_output=True, timeout=600)
    total, found = hazards(out.read_text())
    assert total > 0
    assert not found, found[:5]
