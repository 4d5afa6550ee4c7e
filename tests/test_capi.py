"""The C-ABI library loads and exports every function include/*.h declares
(no GPU calls)."""
import os
import re
import subprocess

from conftest import ROOT
from warpkernel_amd import _capi


def _declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(wk_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declarations_found():
    names = _declared_functions()
    assert {"wk_scan1", "wk_scan2", "wk_engine_create", "wk_run", "wk_run_x"} <= names


def test_library_exports_every_declared_symbol():
    lib = _capi.lib()
    for name in sorted(_declared_functions()):
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", _capi.LIB_PATH], capture_output=True, text=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    assert _declared_functions() <= exported


def test_python_binding_covers_every_symbol():
    assert _declared_functions() <= set(_capi.SIGNATURES)


def test_abi_version_and_error_string():
    lib = _capi.lib()
    assert lib.wk_abi_version() == 1
    assert isinstance(lib.wk_last_error(), bytes)
