"""The host forms of state_counts and cohens_kappa against the values the
reference's own tests hold (test/test_extras.jl:63, test/test_evaluation.jl:8-14),
and the byte-equality test the device counts use, checked exhaustively."""
import numpy as np
import pytest

from cahip import measures as MS
from cahip import model as M
from cahip.backend import count_dict


def test_state_counts_reference_vector():
    # state_counts([:a :b; :a :a]) == Dict(:a => 3, :b => 1), with :a -> 0, :b -> 1
    assert MS.state_counts(np.array([[0, 1], [0, 0]], np.int8)) == {0: 3, 1: 1}
    assert MS.state_counts(np.array([[2, 1], [2, 2]], np.int8))[2] == 3


def test_cohens_kappa_reference_vectors():
    a, b = 0, 1
    assert MS.cohens_kappa([[a, b], [b, a]], [[a, b], [b, a]]) == 1.0
    assert MS.cohens_kappa([[1, 0], [1, 0]], [[1, 1], [0, 0]]) == 0.0
    p = np.array([[1, 1, 0], [0, 1, 0]])
    o = np.array([[1, 0, 0], [1, 1, 0]])
    assert MS.cohens_kappa(p, o) == pytest.approx(1 / 3)  # p_o = 4/6, p_e = 1/2
    assert MS.cohens_kappa(np.full((3, 3), 5), np.full((3, 3), 5)) == 1.0
    assert MS.cohens_kappa(np.full((3, 3), 5), np.full((3, 3), 6)) == 0.0
    with pytest.raises(M.ArgumentError):
        MS.cohens_kappa(np.zeros(0, int), np.zeros(0, int))
    with pytest.raises(M.ArgumentError):
        MS.cohens_kappa(np.zeros((2, 3)), np.zeros((3, 2)))


def test_kappa_from_counts_is_the_brute_force_value():
    # p_e from the category products, against a direct double loop over cell pairs
    rng = np.random.default_rng(4)
    p = rng.integers(0, 3, 50)
    o = rng.integers(0, 3, 50)
    n = p.size
    pe = sum(1 for i in range(n) for j in range(n) if p[i] == o[j]) / n ** 2
    po = np.mean(p == o)
    assert MS.cohens_kappa(p, o) == pytest.approx((po - pe) / (1 - pe), abs=1e-15)


def test_count_dict_maps_bytes_to_int8_values():
    c = np.zeros(256, np.int64)
    c[0], c[1], c[255], c[128] = 4, 2, 1, 3
    assert count_dict(c) == {0: 4, 1: 2, -1: 1, -128: 3}


def zero_byte_tops(x):
    x = np.asarray(x, np.uint32)
    return ~(((x & np.uint32(0x7F7F7F7F)) + np.uint32(0x7F7F7F7F)) | x | np.uint32(0x7F7F7F7F))


def test_byte_equality_word_test_is_exact():
    # every (cell byte, value) pair in every byte position, with random neighbours:
    # the top bit of a byte is set exactly when the cell equals the value
    rng = np.random.default_rng(5)
    cells = np.repeat(np.arange(256, dtype=np.uint32), 256)
    vals = np.tile(np.arange(256, dtype=np.uint32), 256)
    for pos in range(4):
        other = rng.integers(0, 2 ** 32, cells.size, dtype=np.uint64).astype(np.uint32)
        w = (other & ~np.uint32(0xFF << (8 * pos))) | (cells << np.uint32(8 * pos))
        t = zero_byte_tops(w ^ (vals * np.uint32(0x01010101)))
        hit = (t >> np.uint32(8 * pos + 7)) & np.uint32(1)
        assert np.array_equal(hit.astype(bool), cells == vals)
        # and nothing but top bits is ever set
        assert not np.any(t & np.uint32(0x7F7F7F7F))
