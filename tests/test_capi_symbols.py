"""The C-ABI library loads on a CPU-only host and exports every function the
public header declares (no compute calls are made here)."""
import ctypes
import subprocess

from tests.helpers import header_functions
from xin26x_amd import capi


def test_header_parses():
    names = header_functions()
    assert "Xg266ComputeSad" in names and "xg_refine_frac" in names and "Xg266FuncInit" in names
    assert len(names) > 60


def test_library_exports_every_header_symbol():
    lib = ctypes.CDLL(capi.LIB_PATH)
    missing = [n for n in header_functions() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", capi.LIB_PATH], capture_output=True, text=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    assert set(header_functions()) <= exported


def test_bindings_cover_header():
    assert set(header_functions()) == set(capi.PROTOS)


def test_load_sets_prototypes():
    lib = capi.load()
    assert lib.xg_version().startswith(b"xg266")
    assert lib.Xg266ComputeSad.argtypes is not None


def test_library_has_gfx950_code_object():
    # the offload bundle of a library built here names its target
    data = open(capi.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in data


def test_no_unresolved_internal_symbols():
    # every symbol of the library's own namespace is defined inside it
    out = subprocess.run(["nm", "-D", "--undefined-only", capi.LIB_PATH], capture_output=True, text=True).stdout
    internal = [ln.split()[-1] for ln in out.splitlines() if "_ZN2xg" in ln or " xg_" in ln or " Xg2" in ln]
    assert not internal, internal
