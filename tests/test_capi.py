"""The C ABI library loads and exports every entry point include/solidfem_hip.h declares."""
import re
import subprocess
from pathlib import Path

import solidfem_amd as sf
from solidfem_amd import _capi

ROOT = Path(__file__).resolve().parent.parent


def declared():
    text = (ROOT / "include" / "solidfem_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sf_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_entry_point_is_exported():
    names = declared()
    assert len(names) > 60
    out = subprocess.run(["nm", "-D", "--defined-only", str(_capi.LIBRARY)], capture_output=True, text=True, check=True).stdout
    exported = set(line.split()[-1] for line in out.splitlines() if line.strip())
    missing = [n for n in names if n not in exported]
    assert not missing, missing


def test_the_binding_covers_the_header():
    assert set(declared()) <= set(_capi.SIGNATURES) | {"sf_cancel_fn"}


def test_library_loads_and_reports_its_version():
    assert b"gfx950" in sf.library().sf_version()
