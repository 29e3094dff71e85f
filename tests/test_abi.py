"""The C-ABI library loads and exports every symbol include/saige_gpu.h declares (no GPU
calls), and the Python mirror declares a signature for each of them.  CPU only."""
import ctypes
import os
import re
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "saige_gpu.h")
LIB = os.path.join(ROOT, "csaige_amd", "libsaige_gpu.so")


def header_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(saige_[a-z0-9_]+)\s*\(", text)))


def test_header_parses():
    syms = header_symbols()
    assert "saige_grm_mat_vec" in syms and "saige_fit_null" in syms and len(syms) >= 25


def test_library_exports_header_symbols():
    lib = ctypes.CDLL(LIB)
    missing = [s for s in header_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (saige_\w+)", out))
    assert set(header_symbols()) <= exported


def test_python_mirror_declares_every_symbol():
    from csaige_amd import _lib
    declared = {name for name, _, _ in _lib.SIGNATURES}
    assert set(header_symbols()) == declared


def test_error_path_without_device():
    # an argument error is reported through the status and saige_last_error, no GPU touched
    from csaige_amd import _lib
    lib = _lib.load()
    rc = lib.saige_xvx_inv_xv(None, 3, 1, None, None)
    assert rc == 1 and b"invalid argument" in lib.saige_last_error()


def test_header_compiles_as_c():
    src = "#include \"saige_gpu.h\"\nint main(void){return SAIGE_OK;}\n"
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", f"-I{ROOT}/include", "-x", "c", "-",
                        "-o", "/dev/null"], input=src, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
