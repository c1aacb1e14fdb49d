"""Host-side serving logic (no GPU): the admission footprint, the
bookkeeping of the increment buffers under two-step rounds, the request
arena and the split of size classes into lanes."""
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


def test_request_layout_is_aligned_disjoint_and_covers_the_footprint():
    """Every buffer of a request block starts on the alignment, none overlaps
    the next, and the block holds at least the footprint's bytes."""
    from mun_amd.serving import ALIGN, request_layout
    E, nt, pitch, S = 161 * 192 + 17, 1999, 37, 100
    for full in (False, True):
        for two in (False, True):
            lay, size = request_layout(E, nt, pitch, S, corr_chunk(S), E + 9 * 40 * 41, node_elems=161 * 161,
                                       full_history=full, two_step=two)
            end = 0
            for name, shape, dt, zero, off in lay:
                assert off % ALIGN == 0 and off >= end, name
                end = off + math.prod(shape) * (4 if str(dt).endswith("32") else 8)
            assert end <= size < end + ALIGN
            fp = request_footprint(161, 161, E, nt, pitch, S, grad_elems=E + 9 * 40 * 41, full_history=full,
                                   two_step=two)
            assert fp + 161 * 161 * 8 <= size <= fp + 161 * 161 * 8 + ALIGN * len(lay)
            names = [x[0] for x in lay]
            assert ("hist" in names) == full and ("seg" in names) != full and ("w2" in names) == two


def test_block_allocator_first_fit_merges_and_rejects_double_release():
    from mun_amd.serving import ALIGN, BlockAllocator
    a = BlockAllocator(10 * ALIGN + 5)
    assert a.size == 10 * ALIGN
    x = a.alloc(3 * ALIGN)
    y = a.alloc(2 * ALIGN + 1)          # rounds up to 3 blocks
    z = a.alloc(4 * ALIGN)
    assert (x, y, z) == (0, 3 * ALIGN, 6 * ALIGN) and a.alloc(1) is None
    a.release(y, 2 * ALIGN + 1)
    assert a.alloc(4 * ALIGN) is None    # the hole is 3 blocks
    assert a.alloc(2 * ALIGN) == 3 * ALIGN
    assert a.free == [(5 * ALIGN, ALIGN)]
    a.release(x, 3 * ALIGN)
    a.release(z, 4 * ALIGN)              # merges with the one-block tail of the hole
    assert a.free == [(0, 3 * ALIGN), (5 * ALIGN, 5 * ALIGN)]
    assert a.alloc(6 * ALIGN) is None
    a.release(3 * ALIGN, 2 * ALIGN)      # fills the gap: everything merges
    assert a.free == [(0, 10 * ALIGN)] and a.free_bytes() == a.size
    with pytest.raises(ValueError):
        a.release(0, ALIGN)


def test_serving_correlation_chunk_rule():
    """Chunks of at most 50 steps, even unless a chunk is the whole (odd)
    segment, and the last chunk of a segment never longer than the others."""
    for S in range(1, 300):
        c = corr_chunk(S)
        assert 1 <= c <= min(S, 50)
        assert c % 2 == 0 or c == S
        assert S - (S // c) * c <= c


def test_size_classes_split_into_two_lanes_where_one_solve_fills_the_gpu():
    """Two lanes where one solve's two-step launch (19 x 13 tile, 16 x 10
    owned elements) has a workgroup per compute unit: 200x200 (13 x 20 = 260
    tiles) on 256 compute units, not 160x160 (160 tiles) or the smaller
    served meshes."""
    from mun_amd.serving import lanes_for_mesh
    assert lanes_for_mesh(200, 200, 256) == 2
    assert lanes_for_mesh(120, 120, 256) == 1
    assert [lanes_for_mesh(n, n, 256) for n in (20, 40, 80, 160)] == [1, 1, 1, 1]
    assert lanes_for_mesh(160, 160, 160) == 2 and lanes_for_mesh(160, 160, 161) == 1
