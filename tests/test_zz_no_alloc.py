"""tp_field_deriv_noalloc (no allocation: its work area is the library's
per-GPU device array) against tp_field_deriv, through the C ABI: the same
results bit for bit on two calls in a row, the second on a smaller block
(its tables overwrite the first call's in the shared area), no change of
the free device memory across a call, and a clear error for a call larger
than the area."""
import numpy as np
import pytest

from caseutil import seeded_field
from tandemproject_amd import api

pytestmark = pytest.mark.gpu


def test_noalloc_equals_allocating_entry():
    """The device array is placed by the runtime when the library first
    looks it up on a device (tp_create, or the entry point's first call):
    a first tiny call does that, and the free device memory is then the same
    before and after each call."""
    n0 = 20 * 20 * 20
    tiny = api.DeviceArray(n0).upload(np.zeros(n0))
    out = api.DeviceArray(n0).upload(np.zeros(n0))
    api.field_deriv_noalloc((19, 19, 19), (10, 10), 1, tiny.address, out.address)
    api.device_synchronize()
    calls = [((159, 47, 33), 1, (10, 30)),     # first call: 48 x 34 lines of 160 points
             ((30, 24, 20), 3, (45, 45))]       # second, smaller: periodic 21-point lines
    for i, (dims, nn, ends) in enumerate(calls):
        print("no-alloc call", i, dims, "direction", nn, "ends", ends, flush=True)
        shape = (dims[2] + 1, dims[1] + 1, dims[0] + 1)
        f = seeded_field(shape, 70 + i)
        src = api.DeviceArray(f.size).upload(f)
        a = api.DeviceArray(f.size).upload(np.full(f.size, np.nan))
        b = api.DeviceArray(f.size).upload(np.full(f.size, np.nan))
        api.field_deriv(dims, ends, nn, src.address, a.address)
        api.device_synchronize()
        free0 = api.device_memory()[0]
        api.field_deriv_noalloc(dims, ends, nn, src.address, b.address)
        api.device_synchronize()
        free1 = api.device_memory()[0]
        assert free1 == free0
        ga, gb = a.download(), b.download()
        assert np.isfinite(ga).all()
        np.testing.assert_array_equal(gb, ga)


def test_noalloc_refuses_a_call_beyond_its_area():
    x = api.DeviceArray(8).upload(np.zeros(8))
    # 7-point xi lines, 2048 x 4096 of them: the work list alone needs 2 MiB
    # (refused on the host before any device access)
    with pytest.raises(api.TPError, match="work area"):
        api.field_deriv_noalloc((6, 2047, 4095), (10, 10), 1, x.address, x.address)
