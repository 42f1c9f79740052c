"""CPU check of the lane algebra behind wave_sum (hipfoam_device.hpp).

wave_sum forms the xor butterfly v += v[lane ^ m], m = 32, 16, 8, 4, 2, 1,
with VALU lane moves: the swaps of the two 32-lane halves and of 16-lane rows
(v_permlane32_swap / v_permlane16_swap, each called with the same register as
both operands, the two results added), then DPP rotations by 8, 4, 2 and 1
within 16-lane rows.  The claim: every lane ends with exactly the bits the
xor butterfly gives it.  This emulates both lane by lane in float64 on inputs
whose magnitudes make the summation order visible, and checks them bit for
bit; a rotation applied before the xor steps it relies on is shown to differ,
so the check can fail.
"""

import numpy as np


def xor_butterfly(v):
    v = v.copy()
    lanes = np.arange(64)
    for m in (32, 16, 8, 4, 2, 1):
        v = v + v[lanes ^ m]
    return v


def swap_pair(v, half):
    """Results of a permlane swap called as swap(v, v): the first register
    keeps its even blocks and receives the other register's even blocks in
    its odd ones; the second receives the odd blocks (block = `half` lanes)."""
    lanes = np.arange(64)
    blk = lanes // half
    even = np.where(blk % 2 == 0, lanes, lanes - half)  # lane of the even block
    odd = even + half
    return v[even], v[odd]


def row_ror(v, k):
    lanes = np.arange(64)
    src = (lanes // 16) * 16 + (lanes % 16 + k) % 16
    return v[src]


def valu_wave_sum(v, order=(8, 4, 2, 1)):
    a, b = swap_pair(v, 32)
    v = a + b
    a, b = swap_pair(v, 16)
    v = a + b
    for k in order:
        v = v + row_ror(v, k)
    return v


def _inputs(rng, n):
    # mixed magnitudes: the association of the additions changes the bits
    for _ in range(n):
        yield rng.standard_normal(64) * np.exp2(rng.integers(-30, 30, 64))


def test_valu_sum_equals_xor_butterfly_bitwise():
    rng = np.random.default_rng(7)
    for v in _inputs(rng, 2000):
        np.testing.assert_array_equal(valu_wave_sum(v), xor_butterfly(v))


def test_swap_pairs_give_the_xor_partner():
    v = np.arange(64, dtype=np.float64)
    for half in (32, 16):
        a, b = swap_pair(v, half)
        lanes = np.arange(64)
        own_or_partner = np.sort(np.stack([a, b]), axis=0)
        expect = np.sort(np.stack([lanes, lanes ^ half]).astype(np.float64), axis=0)
        np.testing.assert_array_equal(own_or_partner, expect)


def test_rotations_need_the_xor_steps_before_them():
    """Rotating by 1 before 8 does not match the butterfly on these inputs:
    the equivalence rests on the symmetry the earlier steps create."""
    rng = np.random.default_rng(11)
    differs = any(not np.array_equal(valu_wave_sum(v, order=(1, 2, 4, 8)), xor_butterfly(v))
                  for v in _inputs(rng, 200))
    assert differs
