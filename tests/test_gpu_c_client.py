"""The drop-in boundary exercised from C: tests/c_client/config1_client.c
includes only include/ssol.h and include/ssp.h, links -lssol, builds the
flat-mirror scene through the scene API and checks, from C, the closed
forms of ssol_solve's estimates, the receiver and per-primitive getters
(struct layouts as the header declares them), a mirror parameter rewritten
in its parameter buffer between two solves, and the RNG state round trip
through ssp_rng_write / ssp_rng_read (src/ssol.h:1107-1223,
src/ssol_estimator.c:368-418)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLIENT = os.path.join(ROOT, "tests", "c_client", "config1_client")


@pytest.mark.gpu
def test_c_client_config1(gpu):
    assert os.path.exists(CLIENT), "build the client with __graft_entry__.build()"
    p = subprocess.run([CLIENT, "1000000"], capture_output=True, text=True, timeout=120)
    print(p.stdout, p.stderr, flush=True)
    assert p.returncode == 0, p.stderr
    assert p.stdout.startswith("ok: 0 failed checks")
