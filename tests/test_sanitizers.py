"""The host library objects and the kernel emulator under AddressSanitizer
and UndefinedBehaviorSanitizer (tests/sanitize): the init family, GF(2)[x]
service and inline genrand family on the host, and every kernel path the
emulator runs -- generation in each step layout with segment chains, split
jumps and split jump inputs, device init -- each checked against the host
reference functions.  A sanitizer report ends the program with an error."""
import os
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sanitize")


@pytest.fixture(scope="module")
def built():
    subprocess.check_call(["make", "-s", "-C", HERE, "-j2"])


@pytest.mark.parametrize("prog,marker", [("sanitize_host", "SANITIZE-HOST-OK"), ("sanitize_emu", "SANITIZE-EMU-OK")])
def test_under_asan_ubsan(built, prog, marker):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([os.path.join(HERE, prog)], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0 and marker in r.stdout, (r.stdout[-2000:] + r.stderr[-4000:])
