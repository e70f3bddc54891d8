"""The photon slot scheduler and the per-workgroup table layouts of the joint E-step,
checked on the host (no GPU): rsp_emc_schedule_check colours random lane x bank groups
(32 and 64 banks, uniform and bank-clustered entries) and fits random tile layouts.

* Every entry gets exactly one slot, no lane and no bank is used twice in a slot, and the
  slot count equals the maximum degree max(longest lane, busiest bank) -- the lower bound
  of Koenig's edge-colouring theorem, so no schedule has fewer slots.
* The broadcast-aware schedule (lanes reading the same entry may share its bank in a
  slot): every entry gets one slot, no lane is used twice in a slot, no bank holds two
  different entries in a slot, and it never takes more slots than the exact one.  On
  groups whose lanes share entries it takes fewer; on a half-wave whose 32 lanes hold the
  same five entries of bank 0 plus one own entry it reaches the optimum, 6 slots (the
  exact schedule needs 161: bank 0's degree).
* Every layout maps each 32-entry LDS row onto a permutation of its table row, sigma and
  its inverse agree, and each lane's 16-byte direct-to-LDS source offset stays inside its
  1 KB chunk and lands every entry where the photon slots address it.
"""
import numpy as np


def test_schedule_and_layouts_are_exact():
    from refspi_amd import _lib
    out = np.zeros(7, dtype='i8')
    _lib.call('rsp_emc_schedule_check', 3000, 12345, _lib.host_ptr(out, _lib.C.c_int64))
    groups, conflicts, over, layouts, faults, saved, shared6 = (int(v) for v in out)
    assert groups == 3000 and layouts == 6 * 6
    assert conflicts == 0
    assert over == 0
    assert faults == 0
    assert saved > 0
    assert shared6 == 6


def test_schedule_check_is_seeded():
    from refspi_amd import _lib
    a = np.zeros(7, dtype='i8')
    b = np.zeros(7, dtype='i8')
    _lib.call('rsp_emc_schedule_check', 200, 7, _lib.host_ptr(a, _lib.C.c_int64))
    _lib.call('rsp_emc_schedule_check', 200, 7, _lib.host_ptr(b, _lib.C.c_int64))
    assert np.array_equal(a, b) and a[0] == 200
