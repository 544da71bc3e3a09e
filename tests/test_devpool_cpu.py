"""The device allocator's pool logic (block-dda_amd/csrc/devmem_pool.hpp) on
the host.  tests/cxx/devpool_check.cpp drives it over malloc with a capacity
limit, built with AddressSanitizer and UndefinedBehaviorSanitizer:

- best-fit reuse without a new raw allocation, split remainders handed out next;
- idle neighbours merged on free, wholly idle raw allocations trimmed;
- out of memory: trim and one retry, then the raw error code;
- random alloc/free sequences: fill patterns intact, no live overlap, live plus
  idle bytes equal to the held bytes;
- four threads allocating and freeing at once.
No reference counterpart (the reference keeps its blocks in host memory)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CXX_DIR = os.path.join(ROOT, "tests", "cxx")
EXE = os.path.join(CXX_DIR, "_build", "devpool_check")


def test_pool_logic_on_host():
    subprocess.check_call(["make", "-s", "-C", CXX_DIR, "devpool"])
    out = subprocess.run([EXE], capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stderr[-4000:]
    assert out.stdout.strip() == "ok"
