"""The C++ drop-in header (include/cufhe_gpu.hpp) exercised the way the
reference's own gate test is (reference test/test_gate_gpu.cc): a compiled
program, built in-tree by cufhe_amd/build.py, run on the GPU."""
import os
import subprocess

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu


def _run(name):
    exe = os.path.join(ROOT, "tests", "cpp", name)
    assert os.path.exists(exe), "build with `python cufhe_amd/build.py` first"
    p = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=240)
    print(p.stdout)
    return p


def test_cpp_gate_api_truth_tables():
    p = _run("test_gate_gpu.bin")
    assert p.returncode == 0 and "PASS" in p.stdout


def test_cpp_block_binary_api():
    """Block-binary keys + subset key switching through the C++ drop-in:
    lvl0 and lvl1 NAND/MUX truth tables and subset SampleExtractAndKeySwitch
    (reference test/test_block_binary_gpu.cc)."""
    p = _run("test_block_binary_gpu.bin")
    assert p.returncode == 0 and "ALL PASS" in p.stdout


def test_cpp_launcher_layer():
    """XxxBootstrap<brP, mu, iksP> / <iksP, brP, mu> (lvl01, lvl10, lvl02),
    Mux/NMux/Copy/Not launchers, Bootstrap(out, in, mu), SEIandKS,
    BootstrapTLWE2TRLWE and IdentityKeySwitch<lvl10/lvl20> called the way
    the reference's gate API calls them (src/cufhe_gates_gpu.cu:198-537),
    and a round trip through ctxtInitialize / ctxtDelete device copies."""
    p = _run("test_launchers_gpu.bin")
    assert p.returncode == 0 and "ALL PASS" in p.stdout


def test_cpp_many_stream_stress():
    """1000 streams x 10 chained plain-form gates each, driven by polling
    StreamQuery (reference test/test_intensive.cc): every Nand and Mux output
    decrypts correctly."""
    p = _run("test_intensive_gpu.bin")
    assert p.returncode == 0 and "ALL PASS" in p.stdout


def test_cpp_annihilate_api():
    """AnnihilateKeyPolynomialGen / ToDevice / AnnihilateKeySwitching /
    BatchWithWorkspace through include/annihilate_gpu.hpp at lvl1param and
    lvl2param, the reference test's three input cases (reference
    test/test_annihilate_gpu.cu)."""
    p = _run("test_annihilate_gpu.bin")
    assert p.returncode == 0 and "ALL ANNIHILATE TESTS PASSED" in p.stdout


def test_cpp_aes_ascon_api():
    """AESDec and ASCONXOF through the C++ drop-ins (include/aes_gpu.hpp,
    include/ascon_gpu.hpp): FIPS-197 Appendix B and C.1 blocks decrypted from
    two host threads at once, the published empty-message Ascon-Xof digest
    and the oracle's XOF of 00 .. 0f."""
    p = _run("test_aes_ascon_gpu.bin")
    assert p.returncode == 0 and "PASS" in p.stdout
