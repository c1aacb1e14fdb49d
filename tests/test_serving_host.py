"""Host-side serving logic (no GPU): the admission footprint and the
bookkeeping of the increment buffers under two-step rounds."""
import math

import pytest

from mun_amd.serving import corr_chunk, request_footprint


def test_footprint_counts_states_and_traces():
    NX = NZ = 161
    E, nt, pitch, S = 161 * 192, 2000, 200, 100
    K = math.ceil(nt / S)
    base = request_footprint(NX, NZ, E, nt, pitch, S)
    states = 2 + 2 * K + (S + 1)
    nodal = E * 4 * (1 + 1 + states + (corr_chunk(S) + 1) + 1) + E * 8
    traces = 2 * nt * pitch * 4 + nt * pitch * 8 + nt * 4  # seis, res; obs in double; wavelet
    assert base == nodal + traces
    full = request_footprint(NX, NZ, E, nt, pitch, S, full_history=True)
    assert full - base == E * 4 * ((nt + 1) - states)


@pytest.mark.parametrize("full", [False, True])
def test_two_step_rounds_add_the_second_increment_buffers(full):
    """Two steps per launch keep w and omega in two buffers each."""
    args = (161, 161, 161 * 192, 1500, 64, 100)
    one = request_footprint(*args, full_history=full)
    two = request_footprint(*args, full_history=full, two_step=True)
    assert two - one == 2 * 161 * 192 * 4


@pytest.mark.parametrize("nt", [1, 2, 99, 100, 101, 150, 301, 412])
def test_segment_launch_parity(nt):
    """A segment of ln steps takes ceil(ln / 2) two-step launches (an odd last
    step alone); each launch moves the increment to the other buffer, so it
    ends in the second buffer after an odd number of launches.  Walking the
    segments of a solve the way the server does keeps the flag equal to the
    parity of the total launch count."""
    S = 100
    flip, launches = False, 0
    for k in range(math.ceil(nt / S)):
        ln = min(S, nt - k * S)
        n_launch = (ln + 1) // 2
        launches += n_launch
        if n_launch % 2 == 1:
            flip = not flip
    assert flip == (launches % 2 == 1)
    assert launches == sum((min(S, nt - k * S) + 1) // 2 for k in range(math.ceil(nt / S)))
