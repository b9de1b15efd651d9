"""The host side of the drop-in under AddressSanitizer + UndefinedBehaviorSanitizer
and ThreadSanitizer (CPU only, tests/sanitize): the generator's threads, the
FASTQ chunker and reader, the Stage-0 sampler, low.bin packing and the
threaded file writers, each checked against a plain recomputation."""
from __future__ import annotations

import subprocess
import tempfile
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent / "sanitize"
LIB = Path(__file__).resolve().parents[1] / "kmertreecounter_amd" / "libkmc.so"


def sanitizers_available() -> bool:
    with tempfile.TemporaryDirectory() as td:
        src = Path(td) / "t.cpp"
        src.write_text("int main() { return 0; }\n")
        for flag in ("-fsanitize=address,undefined", "-fsanitize=thread"):
            r = subprocess.run(["g++", flag, str(src), "-o", str(Path(td) / "t")], capture_output=True)
            if r.returncode != 0:
                return False
    return True


@pytest.mark.skipif(not LIB.exists(), reason="libkmc.so is not built")
@pytest.mark.skipif(not sanitizers_available(), reason="g++ has no sanitizer runtimes here")
def test_host_code_under_asan_ubsan_and_tsan():
    r = subprocess.run(["make", "-C", str(HERE), "run"], capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert r.stdout.count("host driver ok") == 2
