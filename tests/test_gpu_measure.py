"""GPU checks of the measurement entry points: the step-kernel event timer and
the streaming state copy."""
import numpy as np
import pytest
import torch

from openhomelbm_amd import _native as N
from openhomelbm_amd.lattice import Lattice
from tests.helpers import random_state

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("lattice,shape,dtype,layout", [(27, (10, 12, 40), "f32", "rows"),
                                                        (27, (10, 12, 40), "f16", "soa"),
                                                        (9, (33, 70), "f64", "soa")])
def test_copy_state_duplicates_the_present_state(lattice, shape, dtype, layout):
    nz = shape[0] if lattice == 27 else 1
    lat = Lattice(lattice, shape[-1], shape[-2], nz, dtype=dtype, layout=layout)
    lat.set_state(random_state(lattice, shape, seed=3))
    lat.nu = 0.05
    lat.step(1)
    assert not torch.equal(lat.fields(0), lat.fields(1))
    lat.copy_state()
    torch.cuda.synchronize()
    assert torch.equal(lat.fields(0), lat.fields(1))


def test_kernel_timer_counts_each_step_launch():
    lat = Lattice(27, 64, 16, 12)
    lat.set_state(random_state(27, (12, 16, 64), seed=4))
    N.kernel_timer(8)
    try:
        lat.step(5)
        ms, n = N.kernel_timer_read()
        assert n == 5 and ms > 0.0
        lat.step(12)  # only the first 8 launches are recorded
        ms2, n2 = N.kernel_timer_read()
        assert n2 == 8 and ms2 > 0.0
        lat.step(2)
        assert N.kernel_timer_read()[1] == 2
        N.kernel_timer(100, every=3)
        lat.step(10)  # launches 0, 3, 6, 9
        assert N.kernel_timer_read()[1] == 4
    finally:
        N.kernel_timer(0)
    lat.step(3)
    assert N.kernel_timer_read() == (0.0, 0)
