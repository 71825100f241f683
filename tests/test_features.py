"""Feature matrix (tests/features.py): every case rendered by the HIP library must equal the CPU
oracle bit for bit -- colour, depth/stencil and statistics -- on the order-independent path
(no queries active) and on the in-order path (statistics queries active).  The CPU-only test
checks that each case renders something on the oracle, so a case cannot pass vacuously."""
import numpy as np
import pytest

import features as F

IDS = [c.name for c in F.CASES]


def _changed_fraction(lib, case, out):
    """Fraction of samples whose colour or depth/stencil bytes differ from the cleared target."""
    blank = F.render_case(lib, F.Case(case.name, [], samples=case.samples, fmt=case.fmt,
                                      clear_depth=case.clear_depth, clear_stencil=case.clear_stencil,
                                      viewport=case.viewport), stats=False)
    n = out["color"].shape[0] * out["color"].shape[1] * out["color"].shape[2]
    col = (out["color"] != blank["color"]).reshape(n, -1).any(1)
    ds = (out["ds"] != blank["ds"]).reshape(n, -1).any(1)
    return float(np.mean(col | ds))


@pytest.mark.parametrize("case", F.CASES, ids=IDS)
def test_case_renders_on_oracle(cpu_lib, case):
    out = F.render_case(cpu_lib, case, stats=True)
    frac = _changed_fraction(cpu_lib, case, out)
    assert frac >= case.min_covered, f"{case.name}: only {frac:.3f} of pixels drawn"
    if case.min_covered == 0.0:
        assert out["ps_invocations"] >= 0


@pytest.mark.gpu
@pytest.mark.parametrize("case", F.CASES, ids=IDS)
def test_case_hip_equals_oracle(hip_lib, cpu_lib, case):
    want = F.render_case(cpu_lib, case, stats=True)
    for stats in (False, True):
        got = F.render_case(hip_lib, case, stats=stats)
        assert np.array_equal(got["ds"], want["ds"]), f"{case.name} stats={stats}: depth/stencil differs " \
            f"({int((got['ds'] != want['ds']).sum())} bytes)"
        d = got["color"] != want["color"]
        assert not d.any(), f"{case.name} stats={stats}: {int(d.sum())} colour bytes differ"
        if stats:
            for k in ("ia_primitives", "cinvocations", "cprimitives", "ps_invocations", "backend_input_pixels"):
                assert got[k] == want[k], (case.name, k, got[k], want[k])
