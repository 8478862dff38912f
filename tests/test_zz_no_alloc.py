"""The allocation-free thermo entry point (tmd_thermo_noalloc) against
tmd_thermo through the C ABI: identical results, bit for bit, on two calls
in a row, the second on a smaller system; a call that would need more
scratch than the fixed array holds is refused with an error."""
import ctypes

import numpy as np
import pytest

from common import RC, SKIN, announce, perturbed_fcc

pytestmark = pytest.mark.gpu


def thermo_pair(tm, sim):
    L = tm.load()
    a = np.zeros(6)
    b = np.zeros(6)
    L.tmd_thermo(RC, 1.0, 1.0, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    tm.check()
    L.tmd_thermo_noalloc(RC, 1.0, 1.0, b.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    tm.check()
    return a, b


def test_thermo_noalloc_matches_thermo():
    import tinymd_amd as tm
    name, cus = tm.device_info()
    if cus == 0:
        pytest.fail("no HIP device visible to libtinymd_amd")
    for nc, seed in ((10, 11), (6, 12)):  # the second call on a smaller system
        announce(f"allocation-free thermo vs tmd_thermo, FCC {nc}^3")
        aabb, m, x, v = perturbed_fcc(nc, amp=0.1, seed=seed)
        sim = tm.Simulation(m, x, v, aabb, RC + SKIN)
        sim.copy_to_accelerator()
        sim.setup(False, RC, SKIN)
        sim.lennard_jones(False, RC)
        sim.run_lj(7, 5, 0.005, False, RC, SKIN)
        a, b = thermo_pair(tm, sim)
        assert a[5] > 0 and a[0] > 0
        assert np.array_equal(a, b), (a, b)
        # again after more steps (the scratch holds the previous call's sums)
        sim.run_lj(3, 5, 0.005, False, RC, SKIN)
        a, b = thermo_pair(tm, sim)
        assert np.array_equal(a, b), (a, b)
        sim.close()


def test_thermo_noalloc_refuses_unset_state():
    """Before the state is on the device the call sets an error (it would
    have to allocate) instead of allocating."""
    import tinymd_amd as tm
    aabb, m, x, v = perturbed_fcc(5, amp=0.1, seed=13)
    sim = tm.Simulation(m, x, v, aabb, RC + SKIN)
    out = np.zeros(6)
    tm.load().tmd_thermo_noalloc(RC, 1.0, 1.0, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    try:
        with pytest.raises(tm.TinyMDError, match="allocation-free"):
            tm.check()
    finally:
        tm.load().tmd_clear_error()
    sim.close()
