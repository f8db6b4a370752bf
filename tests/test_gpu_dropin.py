"""The C++ drop-in API (include/kunpbfv/bfv.hpp) driven exactly as the
reference's correct_mult.cu drives its own classes: expected m1*m2*m1 mod t in
every slot.  N = 16384 uses the reference's logN=14 set (two special primes),
N = 32768 the reference's logN=15 chain (with gamma = 2).  The SEAL-style
N = 4096 set (72-bit Q) has room for one product only, so it checks m1*m2."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "kunpbfv_amd", "examples", "correct_mult")

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,extra", [(4096, ["--no-const"]), (16384, []), (32768, [])])
def test_correct_mult_driver(n, extra):
    assert os.path.exists(EXE), "example driver not built (run __graft_entry__.build())"
    r = subprocess.run([EXE, str(n)] + extra, capture_output=True, text=True, timeout=240)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "slots_wrong=0" in r.stdout


OPS = os.path.join(ROOT, "kunpbfv_amd", "examples", "drop_in_ops")


@pytest.mark.parametrize("n", [4096, 16384])
def test_drop_in_ops_driver(n):
    """rotation, conjugation, negation, constant multiply/add, plaintext add and
    square through the reference's BFVScheme member names, ciphertext,
    plaintext and secret key through wire blobs, and two products through the
    split-IP seam (beforeIP, one batched external_product_T_swk_reuse, afterIP)"""
    assert os.path.exists(OPS), "example driver not built (run __graft_entry__.build())"
    r = subprocess.run([OPS, str(n)], capture_output=True, text=True, timeout=240)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "total_wrong=0" in r.stdout



MEMBERS = os.path.join(ROOT, "kunpbfv_amd", "examples", "batch_members")


@pytest.mark.parametrize("n", [4096, 16384])
def test_batch_and_scheme_members_driver(n):
    """multAndEqual_23_batch equals mult_23 on every pair (two calls, the
    second smaller, no device allocation after the first), encryptZero,
    compute_t_QInv_c0c1c2 with hps_decrypt_scale_and_round, addSWKey_23,
    addAutoKey_23 / automorphismAndAdd, the vector rotKey_vec_23, and (N =
    16384, L = 3) rescaleToAndEqual / rescaleAndEqual with decryption, addition
    across levels and HomMult below the full level"""
    assert os.path.exists(MEMBERS), "example driver not built (run __graft_entry__.build())"
    r = subprocess.run([MEMBERS, str(n)], capture_output=True, text=True, timeout=240)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "total_wrong=0" in r.stdout
