"""Readout errors on sampled bitstrings, after the reference's
test/emu_base/test_utils.py:120-235 (same scripted draws, same seeded
golden bag)."""

import random
from collections import Counter
from unittest.mock import call, patch

from emulators_amd.emu_base.utils import apply_measurement_errors, readout_with_error


def test_readout_with_error_thresholds():
    """test_utils.py:120-128: draws 0.6, 0.08, 0.4, 0.1, 0.04 against
    p_false_pos=0.1, p_false_neg=0.2."""
    with patch("emulators_amd.emu_base.utils.random.random") as rnd:
        rnd.side_effect = [0.6, 0.08, 0.4, 0.1, 0.04]
        ps = dict(p_false_pos=0.1, p_false_neg=0.2)
        assert readout_with_error("0", **ps) == "0"
        assert readout_with_error("0", **ps) == "1"
        assert readout_with_error("1", **ps) == "1"
        assert readout_with_error("1", **ps) == "0"
        assert readout_with_error("0", **ps) == "1"


def test_measurement_errors_scripted():
    """test_utils.py:131-211: zero probabilities leave the bag unchanged;
    with scripted per-bit readouts every bit of every shot is read once, in
    order, and the bag is rebuilt from the readouts."""
    bag = Counter({"1010": 3, "0101": 2, "1111": 1})
    assert apply_measurement_errors(bag, p_false_pos=0.0, p_false_neg=0.0) == bag
    ps = dict(p_false_pos=0.1, p_false_neg=0.2)
    reads = list("101100011010010100011111")
    with patch("emulators_amd.emu_base.utils.readout_with_error") as mock:
        mock.side_effect = reads
        out = apply_measurement_errors(bag, **ps)
        expected = ([call(b, **ps) for b in "1010"] * 3 + [call(b, **ps) for b in "0101"] * 2
                    + [call(b, **ps) for b in "1111"])
        mock.assert_has_calls(expected)
        assert mock.call_count == len(expected)
    assert out == Counter({"1011": 1, "0001": 2, "1010": 1, "0101": 1, "1111": 1})


def test_measurement_errors_seeded_golden():
    """test_utils.py:214-235: the reference's seeded golden bag (Python's
    Mersenne Twister, seed 0xDEADBEEF, one draw per bit)."""
    state = random.getstate()
    try:
        random.seed(0xDEADBEEF)
        bag = Counter({"101010001": 39845, "010110001": 2, "111100001": 1})
        out = apply_measurement_errors(bag, p_false_pos=0.000001, p_false_neg=0.0000085)
    finally:
        random.setstate(state)
    assert out == Counter({"101010001": 39843, "101010000": 1, "111010001": 1,
                           "010110001": 2, "111100001": 1})
    assert sum(out.values()) == sum(bag.values())
