"""Host logic of the mixed ensemble (no GPU): the issue order and side-stream
layout of the bins (largest bubble count first and alone, the others round-robin
over the remaining side streams)."""
import pytest

from coupledbubble_control.ensemble import stream_slots


def test_largest_bin_first_and_alone():
    # bins given as N = 2, 3, 4, 8 on the three side streams
    assert stream_slots([2, 3, 4, 8], 3) == [(3, 0), (2, 1), (1, 2), (0, 1)]
    # the N=8 bin's stream carries nothing else
    slots = dict(stream_slots([8, 2, 4, 3], 3))
    assert [i for i, s in slots.items() if s == 0] == [0]


def test_layouts_for_fewer_bins_and_streams():
    assert stream_slots([16, 2], 2) == [(0, 0), (1, 1)]
    assert stream_slots([2, 16], 2) == [(1, 0), (0, 1)]
    assert stream_slots([4], 1) == [(0, 0)]
    assert stream_slots([2, 3, 4, 8], 1) == [(3, 0), (2, 0), (1, 0), (0, 0)]
    assert stream_slots([2, 3, 4, 8], 2) == [(3, 0), (2, 1), (1, 1), (0, 1)]
    with pytest.raises(ValueError):
        stream_slots([2], 0)

