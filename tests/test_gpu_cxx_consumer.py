"""The C++ test program (tests/cxx/ptree_consumer.cpp) drives the library through the
reference-shaped C++ API (include/spaqr_ptree.hpp over include/spaqr_gpu.h) with no Python in
its process: partition -> assemble -> factorize -> dist_gmres to 1e-12 on the 2D Poisson
problem, the Operation replay equal to solve(x), solve(b, x) and extract_x in the natural
ordering.

The program runs as a child of this test process, which has made cooperative launches of its
own and so holds the device's cooperative queue: the child's cooperative geqp3 launches are
refused with hipErrorIllegalState and must go through the ordinary launch within the
co-resident limit."""
import os
import subprocess

import pytest

from pspaqr_public_amd import build

pytestmark = pytest.mark.gpu


def test_cxx_consumer_runs_the_path():
    if not os.path.exists(build.CONSUMER):
        build.build_consumer()
    r = subprocess.run([build.CONSUMER], capture_output=True, text=True, timeout=240)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    line = [l for l in r.stdout.splitlines() if l.startswith("consumer:")][-1]
    fields = dict(kv.split("=") for kv in line.split()[1:])
    assert float(fields["relres"]) <= 1e-12
    assert float(fields["replay_diff"]) == 0.0 and fields["failed"] == "0"
    assert float(fields["ls_normal_rel"]) <= 1e-11 and int(fields["ls_m"]) > int(fields["ls_n"])
