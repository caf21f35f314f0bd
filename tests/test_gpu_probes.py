"""The bandwidth probes bench.py reports as the roofline's ceilings: each
returns a finite positive rate, and arguments outside the documented sets
are refused before any launch."""
import math

import pytest

import bc_amd

pytestmark = pytest.mark.gpu

NBYTES = 64 << 20


def test_probes_return_rates(ctx):
    print("[gpu] copy, gather and line-gather probes on 64 MB", flush=True)
    rates = [ctx.probe_copy_bandwidth(NBYTES, 2)]
    rates += [ctx.probe_gather_bandwidth(p, NBYTES, 2) for p in (32, 64, 128, 256, 512)]
    rates += [ctx.probe_gather_lines(c, NBYTES, 2) for c in (1, 2, 4)]
    assert all(math.isfinite(r) and r > 0 for r in rates), rates


@pytest.mark.parametrize("call", [
    lambda c: c.probe_gather_lines(3, NBYTES, 2),
    lambda c: c.probe_gather_lines(1, 256, 2),
    lambda c: c.probe_gather_lines(1, NBYTES, 0),
    lambda c: c.probe_gather_bandwidth(96, NBYTES, 2),
])
def test_probes_refuse_bad_arguments(ctx, call):
    with pytest.raises(bc_amd.CentError) as e:
        call(ctx)
    assert e.value.status == bc_amd.CENT_E_INVAL
