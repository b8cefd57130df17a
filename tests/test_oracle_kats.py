"""Oracle vs the reference's own known-answer tests (CPU)."""
import numpy as np
import pytest

import oracle
from tests.golden import kats

CASES = [("box", c) for c in kats.BOX_KATS]


def _run(geo, kind, params, method, p, d):
    shape = getattr(geo, kind)(*params)
    pv = geo.place(geo.logical_volume(shape))
    pts = [np.array([p[k]], dtype=np.float64) for k in range(3)]
    dirs = None if d is None else [np.array([d[k]], dtype=np.float64) for k in range(3)]
    return geo.query(pv, method, pts, dirs)[0]


@pytest.mark.parametrize("kind,case", CASES, ids=[f"{k}-{i}-{c[1]}" for i, (k, c) in enumerate(CASES)])
def test_kat(kind, case):
    params, method, p, d, expect = case
    geo = oracle.Geometry()
    r = _run(geo, kind, params, method, p, d)
    assert kats.check(r, expect), (params, method, p, d, r, expect)
