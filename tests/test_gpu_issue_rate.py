"""The issue-rate probe (tests/native/issue_rate.hip) behind the vector and
scalar issue rates quoted with the VALU roofline: it runs, reports every
probe, and the rates keep their order (a 32-bit add issues at least as fast
as a 64-bit multiply-add; scalar instructions beside the vector ones do not
slow a multiply-add stream that issues no faster than the scalar unit)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROBE = os.path.join(ROOT, "tests", "native", "_build", "issue_rate")


@pytest.mark.gpu
def test_issue_rate_probe(gpu):
    assert os.path.exists(PROBE), "issue-rate probe not built (aoa_amd.build.build_issue_probe)"
    print("running", PROBE, flush=True)
    r = subprocess.run([PROBE], capture_output=True, text=True, timeout=120)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    rates = {}
    for line in r.stdout.strip().splitlines():
        name, ns, ms, t = line.split()
        rates[(name, int(ns))] = float(t)
        assert float(ms) > 0 and float(t) > 0, line
    assert len(rates) == 8, rates
    add, mad = rates[("v_add_u32", 0)], rates[("v_mad_u64_u32", 0)]
    assert add >= 0.95 * mad, rates
    assert rates[("v_mad_u64_u32+salu", 8)] >= 0.8 * mad, rates
