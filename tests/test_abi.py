"""The C ABI (CPU only, no device calls): libkunpbfv.so loads, exports every
entry point include/kunpbfv.h declares, the Python binding covers all of them,
and the host-only parameter functions agree with the oracle."""
import ctypes
import os
import re

import numpy as np

import oracle_lib as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "kunpbfv.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(kbfv_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import kunpbfv_amd as kb
    names = declared()
    assert len(names) >= 40
    raw = ctypes.CDLL(kb.LIB_PATH)
    for n in names:
        assert hasattr(raw, n), n
    assert set(names) == set(kb.SIGNATURES), set(names) ^ set(kb.SIGNATURES)


def test_host_parameter_functions_match_oracle():
    import kunpbfv_amd as kb
    for qn, pn, tn in [(3, 1, 3), (28, 1, 4), (13, 1, 4)]:
        assert kb.chain_60bit(qn, pn, tn) == tuple(orc.chain_60bit(qn, pn, tn))
    qmin = 36028797001138177
    want = np.zeros(5, dtype=np.uint64)
    orc.lib.or_primes_below(1 << 15, qmin, 5, orc.P(want))
    assert kb.primes_below(1 << 15, qmin, 5) == [int(x) for x in want]
    d15 = kb.default_params(15)
    assert d15["gamma"] == 2 and d15["plain"] == 0xC0001 and len(d15["q"]) == 15
    d16 = kb.default_params(16)
    assert (len(d16["q"]), len(d16["p"]), len(d16["t"]), d16["plain"]) == (28, 1, 4, 65537)


def test_errors_are_reported_not_raised_across_the_abi():
    import kunpbfv_amd as kb
    lib = kb.lib
    # bad sizes are rejected with a status code and a message
    q = (ctypes.c_uint64 * 1)()
    rc = lib.kbfv_chain_60bit(0, 1, 1, q, q, q)
    assert rc == -1 and b"bad chain" in lib.kbfv_last_error()
    h = ctypes.c_void_p()
    rc = lib.kbfv_context_create(9, q, 1, q, 1, q, 1, 1, 65537, ctypes.byref(h))
    assert rc == -1 and b"logN" in lib.kbfv_last_error()


def test_one_hip_runtime_whatever_the_import_order():
    """torch bundles a libamdhip64.so with the library's soname: importing the
    package before torch must still leave exactly one HIP runtime mapped (two
    runtimes in one process leave torch without a device)."""
    import subprocess
    import sys
    code = ("import kunpbfv_amd, torch\n"
            "libs = {l.split()[-1] for l in open('/proc/self/maps') if 'libamdhip64' in l}\n"
            "print(len(libs), sorted(libs)[0])\n")
    out = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True, check=True)
    count, path = out.stdout.split()
    assert count == "1", out.stdout
    assert "torch" in path, path


def test_random_seeds_come_from_os_entropy():
    """Default keys are drawn from fresh OS entropy (the reference's
    std::random_device), never from a built-in constant: two seeds differ,
    and so do seeds from two processes."""
    import subprocess
    import sys
    import kunpbfv_amd as kb
    a, b = kb.random_seed(), kb.random_seed()
    assert len(a) == 32 and a != b and a != bytes(32)
    code = "import kunpbfv_amd as kb; print(kb.random_seed().hex())"
    outs = {subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True,
                           check=True).stdout.strip() for _ in range(2)}
    assert len(outs) == 2


def test_drop_in_default_seeds_are_not_constants():
    """The C++ drop-in's default SecretKey and BFVScheme seeds come from
    kbfv_random_seed (no fixed seed constant is left in the source)."""
    src = open(os.path.join(ROOT, "kunpbfv_amd", "csrc", "bfv_api.cpp")).read()
    hdr = open(os.path.join(ROOT, "include", "kunpbfv", "bfv.hpp")).read()
    assert src.count("kbfv_random_seed(") >= 2
    assert "0x5345435245544b59" not in src and "0x6b756e7042465600" not in hdr
