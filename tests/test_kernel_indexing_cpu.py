"""Host emulation of the fused per-pixel kernels' index arithmetic (no GPU).

k_meanfield_fused and k_adjoint_fused (pydensecrf_amd/csrc/filter.hip) map threads to
lane groups of S lanes holding R labels each (LaneGroup, dcrf_device.h), blocks to
pixels through the XCD-contiguous renumbering xcd_block, and stage each wave's slice
pairs in LDS before the groups read their own.  This file restates that arithmetic for
every thread of the launched grid and checks, for the shapes the tests and benchmarks
use (the one-label fixtures included: one lane per pixel, 64 pixels per wave):

* xcd_block is a bijection on the grid's blocks, and the active groups cover every
  pixel exactly once;
* every global read (unary / Q row, slice pairs) and write (Q, messages, b, dU) stays
  inside its buffer, for idle lanes and whole idle waves too (they read clamped,
  in-bounds addresses and write nothing);
* every LDS access stays inside the wave's slice of the dynamic LDS the launcher asks
  for (fused_lds_bytes), the Matrix message rows included.
"""
import numpy as np
import pytest

WAVE_BLOCK, XCDS = 256, 8


def labels_per_lane(L):
    for R in (4, 3, 2):
        if L % R == 0:
            return R
    return 1


def xcd_block(b, nb):
    x, i = b % XCDS, b // XCDS
    q, r = nb // XCDS, nb % XCDS
    return x * q + np.minimum(x, r) + i


def emulate(N, L, DS, NT, matrix_terms=0, ldq=None):
    """Check every index of one fused launch over N pixels, L labels, DS slice pairs per
    pixel (NT terms, matrix_terms of them Matrix with an LDS weight copy)."""
    R = labels_per_lane(L)
    S = (L + R - 1) // R
    assert S <= 64, "the fused kernels run in wave mode only"
    per = 64 // S
    items = (WAVE_BLOCK // 64) * per
    grid = (N + items - 1) // items
    ldq = L if ldq is None else ldq
    blocks = np.arange(grid)
    xb = xcd_block(blocks, grid)
    assert np.array_equal(np.sort(xb), blocks), "xcd_block is not a bijection"
    t = np.arange(WAVE_BLOCK)
    lane = t & 63
    g = lane // S
    l = lane - g * S
    on = g < per
    g = np.where(on, g, 0)
    l = np.where(on, l, np.where(l < S, l, 0))
    k0 = l * R
    wave = t >> 6
    p = (xb[:, None] * (WAVE_BLOCK // 64) + wave[None, :]) * per + g[None, :]  # (grid, 256)
    active = on[None, :] & (p < N)
    pc = np.where(active, p, 0)
    # unary / Q rows (unconditional loads at pc), stores at p when active
    assert k0.max() + R - 1 < L
    assert (pc * L + k0[None, :] + R - 1).max() < N * L
    assert (p[active] * ldq + np.broadcast_to(k0, p.shape)[active] + R - 1).max() < N * ldq
    # coverage: each pixel written by exactly one group (its k0 = 0 lane)
    first = active & (k0[None, :] == 0)
    assert np.array_equal(np.sort(p[first]), np.arange(N)), "pixels not covered exactly once"
    # slice-pair staging: the wave's first pixel, uniform over its lanes
    pw0 = p - g[None, :]
    pw = pw0.reshape(grid, WAVE_BLOCK // 64, 64)
    assert np.all(pw == pw[:, :, :1]), "lanes of a wave disagree on its first pixel"
    pw0w = pw[:, :, 0]  # (grid, waves)
    npairs = np.maximum(0, np.minimum(per, N - pw0w)) * DS
    base = pw0w * DS
    last = np.maximum(npairs - 1, 0)
    for i0 in range(0, per * DS, 128):
        run = i0 < npairs  # waves that execute this round
        for off in (0, 64):
            idx = base[..., None] + np.minimum(i0 + off + np.arange(64), last[..., None])
            assert idx[run].min(initial=0) >= 0 and idx[run].max(initial=0) < N * DS
            slot = np.broadcast_to(i0 + off + np.arange(64), idx.shape)
            wr = slot < npairs[..., None]  # LDS writes of this round
            assert np.all(slot[wr & run[..., None]] < per * DS)
    # each active group reads its own DS pairs from the wave's LDS slice
    own = np.broadcast_to(g[None, :] * DS, active.shape)[active]
    assert own.max(initial=0) + DS - 1 < per * DS
    # LDS: pairs of all waves, the Matrix weight copies, the message rows
    ldm = (L + 3) & ~3
    cwf = matrix_terms * L * ldm
    lds_bytes = (WAVE_BLOCK // 64) * per * DS * 8 + cwf * 4 + (
        (WAVE_BLOCK // 64) * NT * per * ldm * 4 if cwf else 0)
    assert lds_bytes <= 160 * 1024, "more LDS than a workgroup has"
    if cwf:
        msg0 = (WAVE_BLOCK // 64) * per * DS * 2 + cwf  # floats before the message rows
        rows = [msg0 + wave[on] * NT * per * ldm + (k * per + g[on]) * ldm + k0[on] + R - 1 for k in range(NT)]
        assert max(r.max() for r in rows) < lds_bytes // 4
    return grid, per


# (pixels, labels) of the one-label fixtures (tests/golden/crf_small.npz case 4,
# kgrad_small.npz case 7), of other fixtures and of the workloads, and edge sizes
SHAPES = [(17 * 9, 1), (12 * 7, 1), (10 * 10, 2), (13 * 7, 3), (17 * 9, 5), (24 * 20, 21), (320 * 213, 21),
          (1024 * 768, 21), (512 * 512, 21), (1, 1), (63, 4), (64, 4), (65, 4), (257, 1), (1000, 64), (999, 256)]


@pytest.mark.parametrize("N,L", SHAPES)
@pytest.mark.parametrize("DS,NT", [(9, 2), (3, 1), (6, 1)])
def test_fused_kernel_indices(N, L, DS, NT):
    emulate(N, L, DS, NT)


@pytest.mark.parametrize("N,L", [(17 * 9, 1), (13 * 8, 4), (24 * 20, 21), (512 * 512, 21), (700, 64)])
def test_fused_kernel_indices_matrix_rows(N, L):
    emulate(N, L, 9, 2, matrix_terms=2)


def test_padded_q_stride():
    """The inference's intermediate Q with a padded row stride (ldq > L)."""
    emulate(1024 * 768, 21, 9, 2, ldq=24)


def test_xcd_block_bijection_all_grids():
    for nb in range(1, 3000):
        b = np.arange(nb)
        assert np.array_equal(np.sort(xcd_block(b, nb)), b), nb
