"""The engine's own kernel and host sources, compiled for the CPU against the
HIP emulation in tests/cpu_emulation (lockstep 64-lane waves, DPP and
shuffles through barriers, device memory pre-filled with garbage).

* emu_checks: input validation of the kernels under AddressSanitizer;
* the reference's SSW parity pairs through the emulated engine (C ABI),
  bit-exact against the frozen SSW results, with device memory filled with a
  pattern that reads as large positive integers.
"""
import ctypes
import json
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
EMU = os.path.join(HERE, "cpu_emulation")
CSRC = os.path.join(ROOT, "ipuma-lib_amd", "csrc")
CLANG = "/opt/rocm/llvm/bin/clang++"
BUILD = os.path.join(ROOT, "_build_emu")

pytestmark = pytest.mark.skipif(not os.path.exists(CLANG), reason="host clang++ not found")


def _compile(src, out, extra=(), lang="c++"):
    cmd = [CLANG, "-std=c++20", "-O1", "-g", "-fPIC", "-pthread", "-x", lang, *extra,
           "-I" + EMU, "-I" + CSRC, "-I" + os.path.join(ROOT, "include"), "-c", src, "-o", out]
    subprocess.run(cmd, check=True)


@pytest.fixture(scope="module")
def build_dir():
    os.makedirs(BUILD, exist_ok=True)
    yield BUILD


def test_kernel_input_validation_asan(build_dir):
    san = ["-fsanitize=address", "-fno-omit-frame-pointer"]
    objs = []
    for src, define in (("sw_kernels.hip", []), ("sw_tables.cpp", [])):
        o = os.path.join(build_dir, src + ".asan.o")
        _compile(os.path.join(CSRC, src), o, san + define)
        objs.append(o)
    main_o = os.path.join(build_dir, "emu_checks.o")
    _compile(os.path.join(EMU, "emu_checks.cpp"), main_o, san + ["-DEMU_DEFINE_GLOBALS"])
    exe = os.path.join(build_dir, "emu_checks")
    subprocess.run([CLANG, "-fsanitize=address", "-pthread", main_o, *objs, "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr


def test_ssw_parity_pairs_emulated(build_dir):
    objs = []
    for src, extra in (("sw_engine.cpp", ["-DEMU_DEFINE_GLOBALS"]), ("sw_kernels.hip", []),
                       ("sw_tables.cpp", [])):
        o = os.path.join(build_dir, src + ".o")
        _compile(os.path.join(CSRC, src), o, extra)
        objs.append(o)
    lib_path = os.path.join(build_dir, "libipuma_emu.so")
    subprocess.run([CLANG, "-shared", "-pthread", *objs, "-o", lib_path], check=True)
    lib = ctypes.CDLL(lib_path)
    ctypes.c_int.in_dll(lib, "_ZN3emu9fill_byteE").value = 0x3F
    import ipuma_amd as I
    I.lib.cache_clear()
    saved = I.LIB_PATH
    I.LIB_PATH = lib_path
    try:
        d = json.load(open(os.path.join(HERE, "golden", "ssw_fixtures.json")))
        case = d["cases"][0]  # 18 DNA pairs of ~100 bp, 8 lanes per pair
        c = case["config"]
        eng = I.SWAlgorithm(I.SWConfig(gapInit=c["gap_init"], gapExtend=c["gap_extend"]),
                            I.IPUAlgoConfig(tilesUsed=1 << 20, maxAB=1 << 16, maxBatches=1))
        s, ar, br = eng.compare([p[0] for p in case["pairs"]], [p[1] for p in case["pairs"]])
        qb, qe = I.unpack_range(ar)
        rb, re_ = I.unpack_range(br)
        for k, (score, eqb, eqe, erb, ere) in enumerate(case["expected"]):
            assert (int(s[k]), int(qb[k]), int(qe[k]), int(rb[k]), int(re_[k])) == \
                (score, eqb, eqe, erb, ere), k
        eng.close()
    finally:
        I.LIB_PATH = saved
        I.lib.cache_clear()
