"""Outcome counting (count_outcomes, simulator.rs:187-238) on the host: keys
past 64 classical bits, the OR of repeated measures into one classical bit
up to 128 bits and the last write above, classical bit 0 on the right and
string order (a BTreeMap of bitstrings)."""
import numpy as np

from qevedo.simulator import _count


def test_small_keys_and_bit_order():
    o = np.array([0b101, 0b101, 0b011, 0b000], dtype=np.uint64)
    assert _count(o, [(0, 0), (1, 1), (2, 2)], 3) == {"000": 1, "011": 1, "101": 2}
    # qubit 0 into classical bit 2 and qubit 2 into classical bit 0
    assert _count(o, [(0, 2), (2, 0)], 3) == {"000": 1, "100": 1, "101": 2}


def test_classical_bits_beyond_64_are_kept():
    o = np.array([1, 1 << 33, (1 << 33) | 1], dtype=np.uint64)
    c = _count(o, [(0, 64), (33, 127)], 128)
    one, top = "0" * 63 + "1" + "0" * 64, "1" + "0" * 127
    both = "1" + "0" * 62 + "1" + "0" * 64
    assert c == {one: 1, top: 1, both: 1}
    assert all(len(k) == 128 for k in c)


def test_repeated_classical_bit_is_an_or_up_to_128_bits():
    o = np.array([0b01, 0b10, 0b00], dtype=np.uint64)
    assert _count(o, [(0, 5), (1, 5)], 6) == {"000000": 1, "100000": 2}
    assert _count(o, [(0, 100), (1, 100)], 101) == {"0" * 101: 1, "1" + "0" * 100: 2}


def test_above_128_bits_the_last_write_wins():
    o = np.array([0b01, 0b10, 0b00], dtype=np.uint64)
    c = _count(o, [(0, 129), (1, 129)], 130)
    assert c == {"0" * 130: 2, "1" + "0" * 129: 1}


def test_counts_are_in_string_order():
    r = np.random.Generator(np.random.PCG64(5))
    o = r.integers(0, 1 << 20, size=4000, dtype=np.uint64)
    measures = [(q, 70 - 3 * q) for q in range(20)]
    c = _count(o, measures, 71)
    assert list(c) == sorted(c)
    assert sum(c.values()) == o.size
