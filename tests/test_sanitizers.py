"""CPU AddressSanitizer / UndefinedBehaviorSanitizer builds of the host code
(no device): the oracle with a driver that calls every entry point
(oracle/san_driver.c), and the library's host code (the layout reader,
generators and writers of frw_layout.cpp) with the host build of the device
traversal code (tests/hostsim/san_main.cpp), run on the reference's fixtures
and on malformed input.  Any sanitizer report aborts the run (no recovery),
so a zero exit status and the driver's 'ok' line are the pass condition."""
import os
import subprocess

import pytest

from conftest import FIXTURES, ROOT, TESTS, layout_path

ORACLE_SAN = os.path.join(ROOT, "oracle", "_san", "san_driver")
HOST_SAN = os.path.join(TESTS, "hostsim", "san_main")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")


@pytest.fixture(scope="module")
def built():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "san"], check=True)
    subprocess.run(["make", "-s", "-C", os.path.join(TESTS, "hostsim"), "san"], check=True)
    return True


def run(args):
    out = subprocess.run(args, capture_output=True, text=True, timeout=300, env=ENV)
    assert "Sanitizer" not in out.stderr and "runtime error" not in out.stderr, out.stderr[-4000:]
    return out


@pytest.mark.parametrize("name,master", FIXTURES)
def test_oracle_under_sanitizers(built, name, master):
    out = run([ORACLE_SAN, layout_path(name), str(master), "300"])
    assert out.returncode == 0 and out.stdout.startswith("ok"), out.stdout + out.stderr[-2000:]


@pytest.mark.parametrize("name,master", FIXTURES)
def test_host_code_under_sanitizers(built, tmp_path, name, master):
    out = run([HOST_SAN, layout_path(name), str(master), "200", str(tmp_path)])
    assert out.returncode == 0 and out.stdout.startswith("ok"), out.stdout + out.stderr[-2000:]


def test_host_reader_on_malformed_input_under_sanitizers(built, tmp_path):
    bad = tmp_path / "bad"
    bad.write_text("rect 1 2 3\nrect a b c d\n" + "x" * 5000 + "\n<<\nrect 1 1 5 5\n")
    out = run([HOST_SAN, str(bad), "0", "10", str(tmp_path)])
    assert out.returncode != 0  # the reader refuses the file with a status code
    empty = tmp_path / "empty"
    empty.write_text("")
    out = run([HOST_SAN, str(empty), "0", "10", str(tmp_path)])
    assert out.returncode != 0
