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
  for (fused_lds_bytes), the Matrix message rows and the adjoint's statistic slots
  included.

It also restates launch_sum_rows's first kernel (adjoint.hip), which adds the adjoint's
per-block statistic rows and the matrix-core statistic's tiles: every entry is read
exactly once and every write stays inside its buffer; and the one-pass kernel gradient
k_kg_grad (kgrad.hip): its capped grid with several rounds per block takes every point
exactly once, each parameter lane adds every point's product exactly once from the LDS
row its group's leader wrote, and its partial rows fit the buffer the backward reserves;
and k_kg_relayout's row map from the plan's paired layout to its x layout.  Last, the
chunking of the staged host transfers (capi.hip), which is host arithmetic.
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


def emulate(N, L, DS, NT, matrix_terms=0, ldq=None, nstat=0):
    """Check every index of one fused launch over N pixels, L labels, DS slice pairs per
    pixel (NT terms, matrix_terms of them Matrix with an LDS weight copy; the adjoint
    with nstat terms writing statistic rows)."""
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
    # LDS: pairs of all waves, the Matrix weight copies, the message rows, the
    # adjoint's statistic slots
    ldm = (L + 3) & ~3
    cwf = matrix_terms * L * ldm
    waves = WAVE_BLOCK // 64
    stat_floats = nstat * WAVE_BLOCK * R + nstat * waves * L  # stat_lds_floats (wave mode)
    lds_bytes = waves * per * DS * 8 + cwf * 4 + (waves * NT * per * ldm * 4 if cwf else 0) + stat_floats * 4
    assert lds_bytes <= 160 * 1024, "more LDS than a workgroup has"
    msg0 = waves * per * DS * 2 + cwf  # floats before the message rows
    if cwf:
        rows = [msg0 + wave[on] * NT * per * ldm + (k * per + g[on]) * ldm + k0[on] + R - 1 for k in range(NT)]
        assert max(r.max() for r in rows) < lds_bytes // 4
    if nstat:
        sst = msg0 + (waves * NT * per * ldm if cwf else 0)
        assert sst + stat_floats == lds_bytes // 4
        # stat_put: thread t's R products of slot k at (k * 256 + t) * R
        put = (np.arange(nstat)[:, None, None] * WAVE_BLOCK + t[None, :, None]) * R + np.arange(R)
        assert put.max() < nstat * WAVE_BLOCK * R
        # stat_finish, wave sums: thread j = (slot, w, lab) reads lane w*64 + gi*S + ln,
        # register rr, for gi < per; that lane (group gi) holds label ln*R + rr = lab
        j = np.arange(nstat * waves * L)
        slot, rem = j // (waves * L), j % (waves * L)
        w, lab = rem // L, rem % L
        ln, rr = lab // R, lab % R
        gi = np.arange(per)
        lane_read = w[:, None] * 64 + gi[None, :] * S + ln[:, None]
        src = slot[:, None] * WAVE_BLOCK * R + lane_read * R + rr[:, None]
        assert src.min() >= 0 and src.max() < nstat * WAVE_BLOCK * R
        assert np.all((lane_read % 64) // S == gi[None, :]) and np.all((lane_read % 64) % S * R + rr[:, None] == lab[:, None])
        # the wave rows (written at j, after the slots) and the final sums' reads
        assert j.max() < nstat * waves * L
        fin = np.arange(nstat * L)
        rd = (fin // L)[:, None] * waves * L + np.arange(waves)[None, :] * L + (fin % L)[:, None]
        assert rd.max() < nstat * waves * L
        # statistic rows: block b writes gstat[b * L + lab], one row per block of the launch
        assert (grid - 1) * L + L - 1 < grid * L
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


@pytest.mark.parametrize("N,L", [(17 * 9, 1), (13 * 8, 4), (24 * 20, 21), (1024 * 768, 21), (700, 64)])
@pytest.mark.parametrize("matrix_terms,nstat", [(0, 2), (1, 1), (0, 1)])
def test_adjoint_statistic_slots(N, L, matrix_terms, nstat):
    """The fused adjoint's statistic slots (after the message rows) and the rows it
    writes per block."""
    emulate(N, L, 9, 2, matrix_terms=matrix_terms, nstat=nstat)


def test_padded_q_stride():
    """The inference's intermediate Q with a padded row stride (ldq > L)."""
    emulate(1024 * 768, 21, 9, 2, ldq=24)


def test_xcd_block_bijection_all_grids():
    for nb in range(1, 3000):
        b = np.arange(nb)
        assert np.array_equal(np.sort(xcd_block(b, nb)), b), nb


def sum_rows_reads(rows, cols, row_blocks=256):
    """(row, column) pairs read by launch_sum_rows's first kernel (k_rows_stage1, adjoint.hip),
    and the mid entries it writes; asserts every read and write is inside its buffer."""
    nb = min(row_blocks, rows)
    rpb = (rows + nb - 1) // nb
    nbu = (rows + rpb - 1) // rpb
    cpb = cols if cols < 256 else 256
    tpc = 256 // cpb
    t = np.arange(256)
    sub = t // cpb
    reads = np.zeros((rows, cols), np.int64)
    mids = set()
    for blk in range(nbu):
        r0, r1 = blk * rpb, min(rows, blk * rpb + rpb)
        for c0 in range(0, cols, cpb):
            c = c0 + t % cpb
            ok = (sub < tpc) & (c < cols)
            for tt in t[ok]:
                rr = np.arange(r0 + sub[tt], r1, tpc)
                assert rr.size == 0 or (rr.max() * cols + c[tt] < rows * cols)
                reads[rr, c[tt]] += 1
            for tt in range(min(cpb, cols - c0)):
                mids.add(blk * cols + c0 + tt)
    assert max(mids) < row_blocks * cols  # inside rows_mid_doubles(cols)
    return reads, mids, nbu


@pytest.mark.parametrize("rows,cols", [(1, 1), (7, 3), (2000, 21), (4095, 441), (3, 300), (1000, 256),
                                       (999, 257), (30000, 42)])
def test_sum_rows_reads_each_entry_once(rows, cols):
    reads, mids, nbu = sum_rows_reads(rows, cols)
    assert np.all(reads == 1)  # every entry once: no row or column dropped or doubled
    assert mids == set(range(nbu * cols))  # the second kernel reads exactly these


def one_launch_combine_list(counts, per, split=16):
    """The layout's one-launch recombination list (capi.hip ensure_layout, wave-mode
    shapes): (v, first, count, pad) items, light vertices first, padded to whole waves,
    then one wave per vertex with more than `split` partial rows."""
    light, heavy, first = [], [], 0
    for v, c in enumerate(counts):
        if c <= split:
            light.append((v, first, c, 0))
        else:
            pieces = min(per, (c + split - 1) // split)
            off = 0
            for j in range(per):
                n = (c - off) // (pieces - j) if j < pieces else 0
                heavy.append((v, first + off, n, pieces))
                off += n
        first += c
    if heavy:
        while len(light) % per:
            light.append((0, 0, 0, 0))
    return light + heavy, first


@pytest.mark.parametrize("L", [21, 4, 1, 64])
def test_one_launch_combine_waves(L):
    """k_combine_waves (filter.hip) reads cb[wave's first item] as the wave's head and
    lets a heavy vertex's pieces meet in the wave's LDS rows: every heavy vertex must
    fill exactly one whole wave (pieces <= groups per wave), light waves hold no heavy
    item, each vertex's pieces cover its partial rows exactly once, and no piece is
    longer than ceil(count / pieces) rows."""
    R = labels_per_lane(L)
    S = (L + R - 1) // R
    per = 64 // S
    rng = np.random.default_rng(L)
    counts = np.concatenate([rng.integers(2, 17, 500), [17, 31, 160, 245, 1000, 7844 // 32],
                             rng.integers(17, 300, 40)])
    rng.shuffle(counts)
    items, total = one_launch_combine_list(counts.tolist(), per)
    assert len(items) % per == 0 or all(it[3] == 0 for it in items)
    covered = np.zeros(total, np.int64)
    for w0 in range(0, len(items), per):
        wave = items[w0:w0 + per]
        pads = {it[3] for it in wave}
        if wave[0][3] == 0:
            assert pads == {0}, "a heavy piece in a light wave"
        else:
            assert len(wave) == per and len({it[0] for it in wave}) == 1 and pads == {wave[0][3]}
            pieces = wave[0][3]
            assert 2 <= pieces <= per
            assert all(it[2] == 0 for it in wave[pieces:])  # fillers after the pieces
            c = sum(it[2] for it in wave)
            assert max(it[2] for it in wave) <= -(-c // pieces)
        for v, f, n, _ in wave:
            covered[f:f + n] += 1
    assert np.all(covered == 1)
    # LDS: per rows of L floats per wave, 4 waves per block
    assert 4 * per * L * 4 <= 160 * 1024


def splat_wave_runs(groups, per):
    """The layout's in-wave runs (capi.hip ensure_layout, wave-mode global splat):
    groups = [[(v, n_entries), ...] per (tile, term) group], empty segments (n = 0)
    anywhere.  Returns the padded segment list [(v, out_marker)], with out_marker 'row'
    (a run head: final or partial row) or -1 (continues the segment before it)."""
    segs = []
    for g in groups:
        segs.extend((v, n) for v, n in g if n > 0)
        while len(segs) % per:
            segs.append((-1, 0))
    segs.extend((v, n) for g in groups for v, n in g if n == 0)  # empty segments at the end
    while len(segs) % per:
        segs.append((-1, 0))
    out = []
    for c, (v, _) in enumerate(segs):
        head = v >= 0 and (c % per == 0 or segs[c - 1][0] != v)
        out.append((v, "row" if head else -1))
    return out


@pytest.mark.parametrize("per", [9, 64, 3])
def test_splat_wave_runs(per):
    """k_splat (filter.hip) folds a group's sum into the group before it when the next
    segment (lane + S) has out < 0 and the same vertex, in the same wave: a run never
    crosses a wave, every wave holds segments of one (tile, term) group only (so the
    runs, and the bits, of an image are the same inside a batch), and a vertex's runs
    are its only outputs."""
    rng = np.random.default_rng(per)
    groups, inst = [], []
    for gi in range(40):
        g = []
        for v in sorted(rng.choice(200, rng.integers(0, 12), replace=False)):
            nseg = int(rng.integers(1, 25))
            g.extend([(int(v) + 1000 * gi, 32)] * nseg)
        if gi % 7 == 0:
            g.insert(0, (999999 - gi, 0))  # an empty segment (padding-only vertex)
        groups.append(g)
    segs = splat_wave_runs(groups, per)
    assert len(segs) % per == 0
    # every wave: one group (vertex ids of a group lie in [1000 gi, 1000 gi + 200))
    for w0 in range(0, len(segs), per):
        vs = {v // 1000 for v, _ in segs[w0:w0 + per] if 0 <= v < 900000}
        assert len(vs) <= 1
    # runs: a continuation follows its vertex in the same wave
    for c, (v, o) in enumerate(segs):
        if o == -1 and v >= 0:
            assert c % per != 0 and segs[c - 1][0] == v
    # the same group gives the same runs wherever it sits (alone vs after other groups)
    alone = splat_wave_runs([groups[5]], per)
    first = [s for s in alone if s[0] >= 0]
    inlist = [s for s in segs if 5000 <= s[0] < 5200]
    assert first == inlist


KG_GRAD_BLOCKS = 2048  # kgrad.hip kKgGradBlocks


def kg_grad_launch(N, L, d, full, kparam_blocks=512):
    """k_kg_grad (kgrad.hip): the grid and rounds launch_kg_grad picks, the points each
    lane group takes over its rounds (its leader leaves fg in the wave's LDS rows), and
    the points lane q < np of each wave reads back for its parameter products, with the
    feature-row and partial-row indices.  Returns, per point, how many groups took it
    and how many times each parameter lane added its product."""
    R = labels_per_lane(L)
    S = (L + R - 1) // R
    assert S <= 64
    per = 64 // S
    items = (WAVE_BLOCK // 64) * per
    need = (N + items - 1) // items
    rounds = (need + KG_GRAD_BLOCKS - 1) // KG_GRAD_BLOCKS
    nb = (need + rounds - 1) // rounds
    assert nb <= KG_GRAD_BLOCKS and nb * rounds >= need
    np_ = d * d if full else d
    assert np_ <= 64  # one lane per parameter
    # partial rows: blockIdx.x * np + q inside kparam_partial_doubles(dmax) for d <= dmax
    assert nb * np_ <= max(kparam_blocks, KG_GRAD_BLOCKS) * d * d
    t = np.arange(WAVE_BLOCK)
    w = t >> 6
    lane = t & 63
    g = lane // S
    l = lane - g * S
    on = g < per
    g = np.where(on, g, 0)
    l = np.where(on, l, np.where(l < S, l, 0))
    assert np.all(g < 64)  # fgl[w][g] rows
    taken = np.zeros(N, np.int64)
    added = np.zeros((np_, N), np.int64)
    for b in range(nb):
        b0 = xcd_block(b, nb) * rounds
        for i in range(rounds):
            p = ((b0 + i) * (WAVE_BLOCK // 64) + w) * per + g
            active = on & (p < N)
            assert np.all(np.where(active, p, 0) < max(N, 1))  # vtx / rank / P / x rows
            lead = active & (l == 0)
            np.add.at(taken, p[lead], 1)
            written = {(int(ww), int(gg)): int(pt) for ww, gg, pt in zip(w[lead], g[lead], p[lead])}
            for q in range(np_):
                for ww in range(WAVE_BLOCK // 64):
                    tq = ww * 64 + q
                    for gi in range(per):
                        pg = int(p[tq]) - int(g[tq]) + gi
                        if pg >= N:
                            break
                        assert written[(ww, gi)] == pg  # the row lane q reads holds pg's fg
                        added[q, pg] += 1
    return taken, added


@pytest.mark.parametrize("N,L,d,full", [(262144, 21, 5, False), (262144, 21, 2, True), (3072, 21, 5, True),
                                        (100, 1, 2, False), (4097, 4, 6, True), (100003, 21, 5, False),
                                        (12, 64, 3, False), (61, 256, 1, True)])
def test_kg_grad_points_once(N, L, d, full):
    taken, added = kg_grad_launch(N, L, d, full)
    assert np.all(taken == 1) and np.all(added == 1)


def kg_map_row(F, X, Z2, Z1, r):
    """kgrad.hip kg_map_row: a row of the kernel-gradient plan's paired layout in its x
    layout."""
    if r > Z2:
        return Z1 + (r - Z2)
    i = 0
    for k in range(1, len(F)):
        if r >= F[k]:
            i = k
    return X[i] + (r - F[i])


@pytest.mark.parametrize("V,partials", [([36254, 101712], [900, 5000]), ([5], [0]), ([7, 3, 11, 2], [1, 0, 4, 2])])
def test_kg_relayout_rows(V, partials):
    """The x layout's rows of every vertex row and partial row of the paired layout
    (forward rows F_i, reverse rows R_i = F_i + V_i side by side per term, the zero row,
    then the partial rows): vertex rows land on [X_i, X_i + V_i), partial rows after the
    x layout's zero row, one to one."""
    F, X, row, xrow = [], [], 0, 0
    for v in V:
        F.append(row)
        X.append(xrow)
        row += 2 * v
        xrow += v
    Z2, Z1 = row, xrow
    got = []
    for i, v in enumerate(V):
        rows = [kg_map_row(F, X, Z2, Z1, F[i] + j) for j in range(v)]
        assert rows == list(range(X[i], X[i] + v))
        got += rows
    P = sum(partials)
    prow = [kg_map_row(F, X, Z2, Z1, Z2 + 1 + q) for q in range(P)]
    assert prow == list(range(Z1 + 1, Z1 + 1 + P))
    assert len(set(got + prow)) == len(got) + len(prow) and max(got + prow, default=0) < Z1 + 1 + P


def chunk_off(n, T, t):
    """capi.hip chunk_off: chunk t of n bytes split T ways, 4 KiB aligned."""
    return n if t >= T else (n // T * t) & ~4095


def window_chunks(n, T):
    """capi.hip window_chunks: ~2 MiB chunks, between T and 8 T of them."""
    return min(max(n >> 21, T), 8 * T)


@pytest.mark.parametrize("n", [8 << 20, 22020096, 36488452, (8 << 20) + 3, 66060288, 123456789, 256 << 20])
def test_staged_copy_chunks(n):
    """The staged host transfers' chunks (capi.hip copy_to_host / copy_to_device /
    host_copy): for every thread and chunk count they tile [0, n) in order, without
    overlap, and all but the last start on a 4 KiB boundary; in a staged window
    (n >= 8 MiB) every chunk is non-empty and thread t takes chunks t, t + T, ..., so
    each chunk has exactly one thread."""
    for T in range(1, 9):
        for C in (T, window_chunks(n, T)):
            offs = [chunk_off(n, C, c) for c in range(C + 1)]
            assert offs[0] == 0 and offs[-1] == n
            assert all(a < b for a, b in zip(offs, offs[1:]))
            assert all(o % 4096 == 0 for o in offs[:-1])
        C = window_chunks(n, T)
        assert T <= C <= 8 * T
        owned = sorted(c for t in range(T) for c in range(t, C, T))
        assert owned == list(range(C))
