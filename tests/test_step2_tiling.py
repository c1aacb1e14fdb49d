"""Index arithmetic of the two-step kernel (step2_kernel in
mun_amd/csrc/sem_kernels.h), emulated on the CPU with the tile constants
read from that header: over ragged meshes every in-domain node is updated by
exactly one (tile, thread) per launch (no node written twice, none left
out), every global read of an element block stays inside the solve's padded
array, the LDS node arrays are indexed in bounds, and every LDS cell step two
reads was written by step one of the same tile.  No GPU needed."""
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


def _emulate(nex, nez, tile):
    EX2, EZ2, TX2, TZ2, T2 = tile.EX2, tile.EZ2, tile.TX2, tile.TZ2, tile.T2
    U1_W, U1_H, W1_W, W1_H = tile.U1_W, tile.U1_H, tile.W1_W, tile.W1_H
    NX, NZ = nex * P + 1, nez * P + 1
    pitch = (NX + 31) // 32 * 32
    sb = (NZ * pitch + 63) // 64 * 64
    ntx, ntz = -(-nex // TX2), -(-nez // TZ2)
    writes = np.zeros((NZ, NX), int)
    for tz in range(ntz):
        for tx in range(ntx):
            u1 = np.zeros((U1_H, U1_W), bool)
            w1 = np.zeros((W1_H, W1_W), bool)
            reads_u1, reads_w1 = [], []
            for t in range(T2):
                exl, ezl = t % EX2, t // EX2
                ex, ez = tx * TX2 - 2 + exl, tz * TZ2 - 2 + ezl
                inb = t < EX2 * EZ2
                valid = inb and 0 <= ex < nex and 0 <= ez < nez
                n1 = valid and exl >= 1 and ezl >= 1
                r2 = valid and 1 <= exl <= EX2 - 2 and 1 <= ezl <= EZ2 - 2
                own = valid and 2 <= exl <= EX2 - 2 and 2 <= ezl <= EZ2 - 2
                last_x, last_z = ex == nex - 1, ez == nez - 1
                exc, ezc = min(max(ex, 0), nex - 1), min(max(ez, 0), nez - 1)
                base_c = P * ezc * pitch + P * exc
                assert 0 <= base_c and base_c + P * pitch + P < sb  # 5 x 5 block read (clamped)
                base = P * ez * pitch + P * ex
                if n1:
                    x0, z0 = P * (exl - 1), P * (ezl - 1)
                    x0w, z0w = P * (exl - 2), P * (ezl - 2)
                    nodes = [(j, i) for j in range(P) for i in range(P)]
                    nodes += [(j, i) for j in range(P + 1) for i in range(P + 1)
                              if (i == P and last_x and (j < P or last_z)) or (j == P and last_z and i < P)]
                    for j, i in nodes:
                        assert 0 <= base + j * pitch + i < sb  # w^n read, u / w written
                        assert z0 + j < U1_H and x0 + i < U1_W
                        u1[z0 + j, x0 + i] = True
                        if own:
                            assert 0 <= z0w + j < W1_H and 0 <= x0w + i < W1_W
                            w1[z0w + j, x0w + i] = True
                            reads_w1.append((z0w + j, x0w + i))
                            writes[P * ez + j, P * ex + i] += 1
                xl = 0 if exl < 1 else (EX2 - 2 if exl > EX2 - 2 else exl - 1)
                zl = 0 if ezl < 1 else (EZ2 - 2 if ezl > EZ2 - 2 else ezl - 1)
                assert P * zl + P < U1_H and P * xl + P < U1_W  # clamped step-two block
                if r2:
                    reads_u1 += [(P * zl + j, P * xl + i) for j in range(P + 1) for i in range(P + 1)]
            assert all(u1[z, x] for z, x in reads_u1), (nex, nez, tx, tz)
            assert all(w1[z, x] for z, x in reads_w1), (nex, nez, tx, tz)
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
    assert t.TX2 * P * 4 % 256 == 0                       # owned rows start on 256-byte boundaries (FP32)
    assert shape != 0 or 2 * lds_fp32 <= 160 * 1024       # the default shape: two workgroups per CU
