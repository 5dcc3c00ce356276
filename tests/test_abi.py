"""The C-ABI library loads and exports every symbol include/gnx.h and
include/gnx_rng.h declare.
No GPU calls are made here."""
import ctypes
import os
import re
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "3d-gross-neveu-model_amd", "lib", "libgnx.so")
HDRS = [os.path.join(ROOT, "include", h) for h in ("gnx.h", "gnx_rng.h")]


def declared():
    txt = "".join(open(h).read() for h in HDRS)
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(gnx_[a-z0-9_]+)\s*\(", txt)))


def test_library_loads_and_exports():
    assert os.path.exists(LIB), "libgnx.so not built (run __graft_entry__.build())"
    lib = ctypes.CDLL(LIB)
    names = declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), f"{n} missing from libgnx.so"
    assert lib.gnx_abi_version() == 9


def test_library_does_not_link_oracle():
    out = subprocess.run(["ldd", LIB], capture_output=True, text=True).stdout
    assert "oracle" not in out
    syms = subprocess.run(["nm", "-D", LIB], capture_output=True, text=True).stdout
    assert "orc_" not in syms


def test_argument_errors_without_a_gpu():
    """Argument checks come before any HIP call: null handles and bad sizes
    return GNX_ERR_ARG (1) with a message naming the entry point."""
    lib = ctypes.CDLL(LIB)
    lib.gnx_last_error.restype = ctypes.c_char_p
    vp, dp = ctypes.c_void_p, ctypes.c_double

    def err(status, name):
        assert status == 1, (name, status)
        assert name in lib.gnx_last_error().decode(), lib.gnx_last_error()

    err(lib.gnx_dirac_batch(vp(None), 0, vp(None), vp(None), vp(None), 4), "gnx_dirac_batch")
    err(lib.gnx_multi_cg(vp(None), vp(None), vp(None), vp(None), dp(1e-10), ctypes.c_long(0), 1, vp(None)),
        "gnx_multi_cg")
    err(lib.gnx_hmc_ensemble_trajectory(vp(None), 1, vp(None), vp(None), vp(None), vp(None), dp(0.1), 2,
                                        dp(1e-10), dp(1e-3), ctypes.c_long(0), vp(None)),
        "gnx_hmc_ensemble_trajectory")
    err(lib.gnx_link_hmc_ensemble_trajectory(vp(None), 1, vp(None), vp(None), vp(None), vp(None), dp(0.02), 2,
                                             dp(1e-10), ctypes.c_long(0), vp(None)),
        "gnx_link_hmc_ensemble_trajectory")
    n = ctypes.c_int(0)
    err(lib.gnx_multi_create(0, ctypes.byref(n), ctypes.byref(n), 0, ctypes.byref(vp())), "gnx_multi_create")


def test_geometry_outside_exact_decode_is_refused():
    """Shapes whose lex indices overflow the kernels' 32-bit multiply-high
    division (accepted by the 2^26-site cap alone) are refused at creation
    with GNX_ERR_SIZE (4), before any HIP call; tests/host/stencil_check.hip
    shows on the CPU that these shapes would decode wrong."""
    lib = ctypes.CDLL(LIB)
    lib.gnx_last_error.restype = ctypes.c_char_p
    out = ctypes.c_void_p()
    for L, LT in ((128, 4000), (256, 1000), (220, 1364), (4094, 4), (2046, 16)):
        assert lib.gnx_create(L, LT, 0, ctypes.byref(out)) == 4, (L, LT)
        assert b"exact range" in lib.gnx_last_error()
        n, t = (ctypes.c_int * 2)(8, L), (ctypes.c_int * 2)(8, LT)
        assert lib.gnx_multi_create(2, n, t, 0, ctypes.byref(out)) == 4, (L, LT)
        assert b"system 1" in lib.gnx_last_error()
