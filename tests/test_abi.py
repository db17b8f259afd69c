"""Object compatibility of the drop-in layout classes with the reference's
headers (variants/DrawingBase.h:13-48, ParHDE.h, WParHDE.h, PHDE.h,
PivotMDS.h): a probe compiled against each header set prints sizes, member
offsets and the vtable slot of run(); the outputs must be identical.  Runs
on the CPU where the reference's sources are present, skipped elsewhere."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
REF = Path("/root/reference/variants")
PROBE = Path(__file__).resolve().parent / "abi_probe.cpp"


@pytest.mark.skipif(not (REF / "DrawingBase.h").exists(), reason="reference headers not present")
@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_layout_objects_match_reference_abi(tmp_path):
    outs = []
    for i, inc in enumerate((REF, ROOT / "include")):
        exe = tmp_path / f"probe{i}"
        subprocess.run(["g++", "-std=c++17", "-w", "-I", str(inc), str(PROBE), "-o", str(exe)], check=True,
                       capture_output=True, text=True)
        outs.append(subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout)
    ref, ours = outs
    assert "run slot ParHDE 0" in ref
    assert ours == ref
