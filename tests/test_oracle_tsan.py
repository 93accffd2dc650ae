"""ThreadSanitizer over the oracle's OpenMP restatement (CPU only).

The restatement shares flags and labels between threads through atomics on
purpose, so a data race reported here is an unintended one.  GCC's libgomp
is not visible to ThreadSanitizer (its barriers show up as races), so the
driver is built with LLVM's clang against LLVM's OpenMP runtime and run with
the Archer tool, which reports OpenMP's synchronisation to ThreadSanitizer.
Skipped where that toolchain is absent."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LLVM = "/opt/rocm/llvm"
CLANG = os.path.join(LLVM, "bin", "clang++")
ARCHER = os.path.join(LLVM, "lib", "libarcher.so")


@pytest.mark.skipif(not (os.path.exists(CLANG) and os.path.exists(ARCHER)), reason="clang/Archer not installed")
def test_restatement_has_no_data_race(tmp_path):
    exe = str(tmp_path / "tsan_check")
    src = [os.path.join(ROOT, "oracle", "tsan_check.cpp"), os.path.join(ROOT, "oracle", "scc_oracle.cpp")]
    subprocess.run([CLANG, "-O1", "-g", "-std=c++17", "-fopenmp", "-fsanitize=thread", *src, "-o", exe,
                    f"-L{LLVM}/lib", f"-Wl,-rpath,{LLVM}/lib"], check=True, capture_output=True, text=True)
    # the OpenMP runtime's own thread start-up and wake-up code is not
    # instrumented for the races it has by design
    supp = tmp_path / "tsan.supp"
    supp.write_text("called_from_lib:libomp.so\n")
    env = dict(os.environ, OMP_TOOL_LIBRARIES=ARCHER, ARCHER_OPTIONS="verbose=1",
               TSAN_OPTIONS=f"halt_on_error=0 exitcode=66 ignore_noninstrumented_modules=1 suppressions={supp}")
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=600)
    assert "Archer detected OpenMP application with TSan" in r.stderr + r.stdout
    assert "WARNING: ThreadSanitizer" not in r.stderr, r.stderr[-4000:]
    assert r.returncode == 0, (r.returncode, r.stdout[-2000:])
    assert r.stdout.count("0 labels differ") == 18
