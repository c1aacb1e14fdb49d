"""Index arithmetic of the two-step kernel (step2_kernel in
mun_amd/csrc/sem_kernels.h), emulated on the CPU with the tile constants
read from that header: over ragged meshes every in-domain node is updated by
exactly one (tile, thread) per launch (no node written twice, none left
out), every global read of an element block stays inside the solve's padded
array, the LDS node arrays are indexed in bounds, every LDS cell step two
reads was written by step one of the same tile, and under step two's own
thread mapping every edge contribution an owned element assembles was
published by exactly one thread and every cross-lane fifth-node read finds
the right neighbour's block in lane + 1.  No GPU needed."""
import os
import re

import numpy as np
import pytest

P = 4
HDR = os.path.join(os.path.dirname(__file__), "..", "mun_amd", "csrc", "sem_kernels.h")


def _shapes():
    text = open(HDR).read()
    ex = [int(v) for v in re.search(r"TILE2_EX\[NTILE2\] = \{([^}]*)\}", text).group(1).split(",")]
    ez = [int(v) for v in re.search(r"TILE2_EZ\[NTILE2\] = \{([^}]*)\}", text).group(1).split(",")]
    return list(zip(ex, ez))


SHAPES = _shapes()


class Tile:
    def __init__(self, ex2, ez2):
        self.EX2, self.EZ2 = ex2, ez2
        self.TX2, self.TZ2 = ex2 - 3, ez2 - 3
        self.T2 = (ex2 * ez2 + 63) // 64 * 64
        self.U1_W = (4 * (ex2 - 1) + 1 + 3) // 4 * 4
        self.U1_H = 4 * (ez2 - 1) + 1
        self.W1_W = (4 * self.TX2 + 1 + 3) // 4 * 4
        self.W1_H = 4 * self.TZ2 + 1


def _nodes(last_x, last_z):
    """An element's owned nodes: its lower-left 4 x 4, plus the right column /
    top row where it ends the domain."""
    nodes = [(j, i) for j in range(P) for i in range(P)]
    nodes += [(j, i) for j in range(P + 1) for i in range(P + 1)
              if (i == P and last_x and (j < P or last_z)) or (j == P and last_z and i < P)]
    return nodes


def _emulate(nex, nez, tile):
    EX2, EZ2, TX2, TZ2, T2 = tile.EX2, tile.EZ2, tile.TX2, tile.TZ2, tile.T2
    U1_W, U1_H, W1_W, W1_H = tile.U1_W, tile.U1_H, tile.W1_W, tile.W1_H
    NX, NZ = nex * P + 1, nez * P + 1
    pitch = (NX + 31) // 32 * 32
    sb = (NZ * pitch + 63) // 64 * 64
    ntx, ntz = -(-nex // TX2), -(-nez // TZ2)
    writes = np.zeros((NZ, NX), int)
    XW, ZW = EX2 - 2, EZ2 - 2
    NS2 = XW * ZW
    assert NS2 <= T2
    for tz in range(ntz):
        for tx in range(ntx):
            u1 = np.zeros((U1_H, U1_W), bool)
            w1 = np.zeros((W1_H, W1_W), bool)
            # step one: thread t -> element (t % EX2, t // EX2) of the whole block
            for t in range(T2):
                exl, ezl = t % EX2, t // EX2
                ex, ez = tx * TX2 - 2 + exl, tz * TZ2 - 2 + ezl
                inb = t < EX2 * EZ2
                valid = inb and 0 <= ex < nex and 0 <= ez < nez
                n1 = valid and exl >= 1 and ezl >= 1
                own = valid and 2 <= exl <= EX2 - 2 and 2 <= ezl <= EZ2 - 2
                last_x, last_z = ex == nex - 1, ez == nez - 1
                exc, ezc = min(max(ex, 0), nex - 1), min(max(ez, 0), nez - 1)
                base_c = P * ezc * pitch + P * exc
                assert 0 <= base_c and base_c + P * pitch + P < sb  # 5 x 5 block read (clamped)
                base = P * ez * pitch + P * ex
                if n1:
                    x0, z0 = P * (exl - 1), P * (ezl - 1)
                    x0w, z0w = P * (exl - 2), P * (ezl - 2)
                    for j, i in _nodes(last_x, last_z):
                        assert 0 <= base + j * pitch + i < sb  # w^n read, u^(n+1) written
                        assert z0 + j < U1_H and x0 + i < U1_W
                        u1[z0 + j, x0 + i] = True
                        if own:
                            assert 0 <= z0w + j < W1_H and 0 <= x0w + i < W1_W
                            w1[z0w + j, x0w + i] = True
            # step two: thread t -> element (t % XW, t // XW) of the step-two
            # block; waves past it skip the element math
            published = {}
            reads_u1, reads_w1, reads_pub = [], [], []
            loaded, right_reads = {}, []
            for t in range(T2):
                if (t & ~63) >= NS2:
                    assert t >= NS2  # a skipping wave holds no step-two element
                    continue
                in2 = t < NS2
                xl, zl = t % XW, (t // XW if in2 else ZW - 1)
                exl, ezl = xl + 1, zl + 1
                ex, ez = tx * TX2 - 1 + xl, tz * TZ2 - 1 + zl
                valid = in2 and 0 <= ex < nex and 0 <= ez < nez
                own = valid and xl >= 1 and zl >= 1
                last_x, last_z = ex == nex - 1, ez == nez - 1
                exc, ezc = min(max(ex, 0), nex - 1), min(max(ez, 0), nez - 1)
                base_c = P * ezc * pitch + P * exc
                assert 0 <= base_c and base_c + P * pitch + P < sb  # c^2 block read (clamped)
                assert P * zl + P < U1_H and P * xl + P < U1_W      # u^(n+1) block read from LDS
                loaded[t] = (exc, ezc)
                from_right = xl < XW - 1 and (t & 63) != 63 and ex + 1 < nex
                if valid and from_right:  # fifth node from lane t + 1's first node
                    right_reads.append((t + 1, (ex + 1, ez)))
                if valid:
                    reads_u1 += [(P * zl + j, P * xl + i) for j in range(P + 1) for i in range(P + 1)]
                if in2:
                    assert (ezl, exl) not in published and ezl < EZ2 and exl < EX2
                    published[(ezl, exl)] = t
                if own:
                    reads_pub += [(ezl, exl - 1), (ezl - 1, exl), (ezl - 1, exl - 1)]
                    base = P * ez * pitch + P * ex
                    x0w, z0w = P * (exl - 2), P * (ezl - 2)
                    for j, i in _nodes(last_x, last_z):
                        assert 0 <= base + j * pitch + i < sb  # u^(n+2), w^(n+2) written
                        assert 0 <= z0w + j < W1_H and 0 <= x0w + i < W1_W
                        reads_w1.append((z0w + j, x0w + i))
                        writes[P * ez + j, P * ex + i] += 1
            assert all(u1[z, x] for z, x in reads_u1), (nex, nez, tx, tz)
            assert all(w1[z, x] for z, x in reads_w1), (nex, nez, tx, tz)
            assert all(c in published for c in reads_pub), (nex, nez, tx, tz)
            assert all(loaded.get(lane) == blk for lane, blk in right_reads), (nex, nez, tx, tz)
    return writes


@pytest.mark.parametrize("shape", range(len(SHAPES)))
@pytest.mark.parametrize("nex,nez", [(1, 1), (2, 3), (7, 5), (16, 23), (17, 24), (33, 47), (37, 50), (40, 40),
                                     (20, 80), (160, 20)])
def test_two_step_tiles_partition_the_nodes(nex, nez, shape):
    writes = _emulate(nex, nez, Tile(*SHAPES[shape]))
    assert (writes == 1).all(), np.argwhere(writes != 1)[:5]


@pytest.mark.parametrize("shape", range(len(SHAPES)))
def test_two_step_tile_fits_the_cu(shape):
    t = Tile(*SHAPES[shape])
    lds_fp32 = 4 * (t.EZ2 * t.EX2 * (P + 1 + P) + t.U1_H * t.U1_W + t.W1_H * t.W1_W)
    assert t.T2 <= 1024 and lds_fp32 <= 160 * 1024      # one workgroup in a CU's 160 KB
    if shape != 3:
        assert 2 * lds_fp32 <= 160 * 1024                 # FP64 (twice the bytes) fits too
    assert t.TX2 * P * 4 % 64 == 0                        # owned rows start on 64-byte boundaries (FP32)
    assert shape != 0 or 2 * lds_fp32 <= 160 * 1024       # the default shape: two workgroups per CU
