"""Pins of the oracle's H.264 intra prediction (clause 8.3; CPU, no GPU).

The oracle restates src/h264/intra.rs (Intra_4x4 / Intra_8x8 with the
reference filter, Intra_16x16, 4:2:0 chroma) and the availability and
decoding-order rules of src/h264/recon.rs:611-828, 936-1128. These tests
hold it to facts fixed by the standard rather than to its own formulas:

- a flat neighbourhood predicts itself in every mode of every kind;
- vertical / horizontal modes copy the row above / the column to the left;
- DC is the rounded mean of the available edges (128 with none);
- plane prediction reproduces an affine ramp exactly for small slopes
  (slopes -3..4: the 5/64 and 34/64 gradient scales round to 32 per unit);
- the diagonal modes are constant along their directions, horizontal-down
  is vertical-right transposed, horizontal-up saturates at the last left
  sample, the top-right substitution repeats p[N-1, -1];
- the [1 2 1] reference filter keeps constant and linear edges;
- block availability: the 4x4 blocks whose top-right is never available
  inside a macroblock are luma4x4BlkIdx 3, 7, 11, 13 and 15; the 8x8 block 3
  has none; constrained intra prediction hides inter neighbours;
- whole pictures: all-DC pictures without neighbours are 128; vertical and
  horizontal Intra_4x4 macroblocks propagate the edge they copy through the
  macroblock; a residual is added before the next block of the macroblock is
  predicted (decoding order).
"""
import ctypes as C

import numpy as np
import pytest

import oracle
import rivet_hip as rh
from rivet_hip import _abi
from rivet_hip import configs as cf

TOP, LEFT, TL, TR = 1, 2, 4, 8
ALL = TOP | LEFT | TL | TR
KIND_NXN, KIND_16, KIND_C = 0, 1, 2


def predict(kind, n, mode, av, plane, x0, y0):
    """Run the oracle predictor on a copy of `plane` for the block at
    (x0, y0); returns (rc, the block)."""
    p = np.ascontiguousarray(plane.copy())
    off = y0 * p.shape[1] + x0
    rc = oracle.lib.orc_h264_intra_pred(kind, n, mode, av, p.ctypes.data + off, p.shape[1])
    return rc, p[y0:y0 + n, x0:x0 + n].astype(np.int64)


def neighbourhood(n, top, left, corner, fill=0):
    """A plane with the block at (1, 1): row 0 holds the corner then 2n
    top samples, column 0 the left samples."""
    p = np.full((n + 2, 2 * n + 2), fill, np.uint8)
    p[0, 0] = corner
    p[0, 1:1 + len(top)] = top
    p[1:1 + n, 0] = left
    return p


SIZES = [(KIND_NXN, 4, range(9)), (KIND_NXN, 8, range(9)), (KIND_16, 16, range(4)), (KIND_C, 8, range(4))]


@pytest.mark.parametrize("kind,n,modes", SIZES)
@pytest.mark.parametrize("v", [0, 37, 128, 255])
def test_flat_neighbourhood_predicts_itself(kind, n, modes, v):
    p = neighbourhood(n, [v] * (2 * n), [v] * n, v)
    for mode in modes:
        rc, blk = predict(kind, n, mode, ALL, p, 1, 1)
        assert rc == 0 and np.all(blk == v), (kind, n, mode)


def _rand(seed, n):
    r = np.random.default_rng(seed)
    return r.integers(0, 256, 2 * n), r.integers(0, 256, n), int(r.integers(0, 256))


@pytest.mark.parametrize("kind,n,vmode,hmode", [(KIND_NXN, 4, 0, 1), (KIND_NXN, 8, 0, 1), (KIND_16, 16, 0, 1),
                                                (KIND_C, 8, 2, 1)])
def test_vertical_and_horizontal_copy_their_edge(kind, n, vmode, hmode):
    for seed in range(5):
        top, left, c = _rand(seed, n)
        p = neighbourhood(n, top, left, c)
        rc, blk = predict(kind, n, vmode, ALL, p, 1, 1)
        assert rc == 0
        want_v = top[:n] if kind != KIND_NXN or n != 8 else None
        if kind == KIND_NXN and n == 8:
            # Intra_8x8 copies the filtered edge: check against the filter.
            f = filtered8(top, left, c, ALL)
            want_v = f[:8]
        assert np.all(blk == np.asarray(want_v)[None, :]), "vertical"
        rc, blk = predict(kind, n, hmode, ALL, p, 1, 1)
        want_h = left if not (kind == KIND_NXN and n == 8) else filtered8(top, left, c, ALL)[16:24]
        assert rc == 0 and np.all(blk == np.asarray(want_h)[:, None]), "horizontal"


def filtered8(top, left, corner, av):
    i = (C.c_int32 * 25)(*[int(v) for v in top], *[int(v) for v in left], int(corner))
    o = (C.c_int32 * 25)()
    oracle.lib.orc_h264_filter8(i, av, o)
    return np.array(list(o), np.int64)


def test_reference_filter_keeps_constant_and_linear_edges():
    f = filtered8([90] * 16, [90] * 8, 90, ALL)
    assert np.all(f == 90)
    top = [10 + 3 * k for k in range(16)]
    left = [20 + 5 * k for k in range(8)]
    f = filtered8(top, left, 7, ALL)
    # [1 2 1] / 4 keeps a linear edge away from its ends.
    assert np.array_equal(f[1:15], top[1:15]) and np.array_equal(f[16 + 1:16 + 7], left[1:7])
    # The last top sample mirrors: (t14 + 3 t15 + 2) / 4 of a ramp of slope 3.
    assert f[15] == (top[14] + 3 * top[15] + 2) // 4


def test_reference_filter_is_121_over_the_l_shaped_edge():
    # 8.3.2.2.1 read as one filter: the edge p[-1, 7] .. p[-1, 0], p[-1, -1],
    # p[0, -1] .. p[15, -1] as a sequence, each sample replaced by
    # (prev + 2 cur + next + 2) / 4, the sequence's two ends replicated.
    for seed in range(20):
        top, left, c = _rand(500 + seed, 8)  # 16 top samples
        e = [int(v) for v in left[::-1]] + [c] + [int(v) for v in top]
        ext = [e[0]] + e + [e[-1]]
        want = [(ext[i - 1] + 2 * ext[i] + ext[i + 1] + 2) >> 2 for i in range(1, len(ext) - 1)]
        f = filtered8(top, left, c, ALL)
        got = [int(v) for v in f[16:24][::-1]] + [int(f[24])] + [int(v) for v in f[:16]]
        assert got == want


def rounded_mean(vals):
    vals = [int(v) for v in vals]
    return (sum(vals) + len(vals) // 2) // len(vals)


@pytest.mark.parametrize("kind,n,mode", [(KIND_NXN, 4, 2), (KIND_NXN, 8, 2), (KIND_16, 16, 2)])
def test_dc_is_the_rounded_mean_of_available_edges(kind, n, mode):
    for seed in range(6):
        top, left, c = _rand(100 + seed, n)
        p = neighbourhood(n, top, left, c)
        if kind == KIND_NXN and n == 8:
            f = filtered8(top, left, c, ALL)
            t_e, l_e = f[:8], f[16:24]
        else:
            t_e, l_e = top[:n], left
        for av, want in [(ALL, rounded_mean(list(t_e) + list(l_e))), (TOP | TR, None), (LEFT, None), (0, 128)]:
            if av == TOP | TR:
                want = rounded_mean(filtered8(top, left, c, av)[:8] if (kind == KIND_NXN and n == 8) else top[:n])
            if av == LEFT:
                want = rounded_mean(filtered8(top, left, c, av)[16:24] if (kind == KIND_NXN and n == 8) else left)
            rc, blk = predict(kind, n, mode, av, p, 1, 1)
            assert rc == 0 and np.all(blk == want), (av, want, blk[0, 0])


@pytest.mark.parametrize("kind,n,mode,count", [(KIND_NXN, 4, 2, 8), (KIND_16, 16, 2, 32), (KIND_C, 8, 0, 8)])
def test_dc_rounds_half_up(kind, n, mode, count):
    # Edge sums at exactly half a step: the mean must round up.
    half = count // 2
    top = [0] * (2 * n)
    left = [1] * half + [0] * (n - half)
    p = neighbourhood(n, top, left, 0)
    rc, blk = predict(kind, n, mode, ALL, p, 1, 1)
    assert rc == 0 and blk[0, 0] == 1
    left = [0] * n
    top = [1] * (half - 1) + [0] * (2 * n - half + 1)
    rc, blk = predict(kind, n, mode, ALL, neighbourhood(n, top, left, 0), 1, 1)
    assert rc == 0 and blk[0, 0] == 0


def test_chroma_dc_per_quadrant():
    top, left, c = _rand(7, 8)
    p = neighbourhood(8, top, left, c)
    rc, blk = predict(KIND_C, 8, 0, ALL, p, 1, 1)
    assert rc == 0
    # 8.3.4.1-3: the top-left and bottom-right quadrants average both edges;
    # the top-right takes its top samples, the bottom-left its left ones.
    assert np.all(blk[:4, :4] == rounded_mean(list(top[:4]) + list(left[:4])))
    assert np.all(blk[4:, 4:] == rounded_mean(list(top[4:8]) + list(left[4:8])))
    assert np.all(blk[:4, 4:] == rounded_mean(top[4:8]))
    assert np.all(blk[4:, :4] == rounded_mean(left[4:8]))


@pytest.mark.parametrize("kind,n,mode", [(KIND_16, 16, 3), (KIND_C, 8, 3)])
def test_plane_reproduces_affine_ramps(kind, n, mode):
    for a, b, c in [(100, 1, 2), (60, 4, -3), (200, -3, 1), (128, 0, 0), (50, 2, 4)]:
        f = lambda x, y: a + b * x + c * y  # noqa: E731
        top = [f(x, -1) for x in range(2 * n)]
        left = [f(-1, y) for y in range(n)]
        p = neighbourhood(n, np.clip(top, 0, 255), np.clip(left, 0, 255), f(-1, -1))
        rc, blk = predict(kind, n, mode, ALL, p, 1, 1)
        want = np.clip([[f(x, y) for x in range(n)] for y in range(n)], 0, 255)
        assert rc == 0 and np.array_equal(blk, want), (a, b, c)


@pytest.mark.parametrize("n", [4, 8])
def test_directional_modes_structure(n):
    for seed in range(8):
        top, left, c = _rand(300 + seed, n)
        # The top-right half repeats p[N-1, -1], as the edge swapped in
        # below must (the left column has no continuation).
        top = np.concatenate([top[:n], top[n - 1:n].repeat(n)])
        p = neighbourhood(n, top, left, c)
        pt = neighbourhood(n, np.concatenate([left, left[-1:].repeat(n)]), top[:n], c)  # edges swapped
        rc, ddl = predict(KIND_NXN, n, 3, ALL, p, 1, 1)
        assert rc == 0
        for s in range(2 * n - 2):  # constant along anti-diagonals (not the corner)
            vals = {int(ddl[y, s - y]) for y in range(n) if 0 <= s - y < n}
            assert len(vals) == 1
        rc, ddr = predict(KIND_NXN, n, 4, ALL, p, 1, 1)
        for d in range(-n + 1, n):
            assert len({int(ddr[y, y + d]) for y in range(n) if 0 <= y + d < n}) == 1
        _, ddr_t = predict(KIND_NXN, n, 4, ALL, pt, 1, 1)
        assert np.array_equal(ddr, ddr_t.T), "diagonal-down-right is symmetric in its edges"
        _, vr_t = predict(KIND_NXN, n, 5, ALL, pt, 1, 1)
        _, hd = predict(KIND_NXN, n, 6, ALL, p, 1, 1)
        assert np.array_equal(hd, vr_t.T), "horizontal-down is vertical-right transposed"
        _, hu = predict(KIND_NXN, n, 8, ALL, p, 1, 1)
        lastl = left[-1] if n == 4 else filtered8(top, left, c, ALL)[16 + 7]
        for y in range(n):
            for x in range(n):
                if x + 2 * y > 2 * n - 3:
                    assert hu[y, x] == lastl


def test_linear_top_closed_forms():
    # Slope 4 top row t[k] = 20 + 4k: diagonal-down-left gives t at x + y + 1
    # ((t[k] + 2 t[k+1] + t[k+2] + 2) / 4 of a line), vertical-left's even
    # rows the midpoints t[x + y/2] + 2.
    top = [20 + 4 * k for k in range(8)]
    p = neighbourhood(4, top, [0] * 4, 0)
    rc, ddl = predict(KIND_NXN, 4, 3, TOP | TR, p, 1, 1)
    for y in range(4):
        for x in range(4):
            if (x, y) != (3, 3):
                assert ddl[y, x] == 20 + 4 * (x + y + 1)
    rc, vl = predict(KIND_NXN, 4, 7, TOP | TR, p, 1, 1)
    for y in (0, 2):
        for x in range(4):
            assert vl[y, x] == 20 + 4 * (x + y // 2) + 2


def test_top_right_substitution():
    top, left, c = _rand(9, 4)
    p = neighbourhood(4, top, left, c)
    sub = np.concatenate([top[:4], [top[3]] * 4])
    q = neighbourhood(4, sub, left, c)
    for mode in (3, 7):
        _, a = predict(KIND_NXN, 4, mode, TOP | LEFT | TL, p, 1, 1)  # top-right unavailable
        _, b = predict(KIND_NXN, 4, mode, ALL, q, 1, 1)
        assert np.array_equal(a, b), mode


@pytest.mark.parametrize("kind,n,mode,av", [(KIND_NXN, 4, 0, LEFT), (KIND_NXN, 4, 1, TOP), (KIND_NXN, 8, 4, TOP | LEFT),
                                            (KIND_NXN, 4, 5, TOP | TL), (KIND_16, 16, 3, TOP | LEFT),
                                            (KIND_16, 16, 0, LEFT), (KIND_C, 8, 1, TOP), (KIND_C, 8, 3, LEFT | TL),
                                            (KIND_NXN, 4, 9, ALL), (KIND_16, 16, 4, ALL)])
def test_modes_without_their_neighbours_are_refused(kind, n, mode, av):
    p = neighbourhood(n, [1] * (2 * n), [1] * n, 1)
    rc, _ = predict(kind, n, mode, av, p, 1, 1)
    assert rc == _abi.RV_E_INVALID


# ---------------------------------------------------------------- availability

def intra_picture(w, h, cip, seed=5, pct=100):
    cfg = dict(cf.H264_C4, width=w, height=h, intra_percent=pct, intra_device=1, constrained_intra_pred=cip,
               b_picture=0, weighting=0, poc=1, ref_poc=[0, 0])
    return rh.GenPicture(rh.RV_CODEC_H264, cfg, seed)


def blk_idx(bx, by):
    """luma4x4BlkIdx of raster (bx, by) (6.4.3)."""
    return 8 * (by // 2) + 4 * (bx // 2) + 2 * (by % 2) + (bx % 2)


def test_block_availability_inside_a_macroblock():
    g = intra_picture(48, 48, 0)
    centre = 4  # every neighbour present
    for by in range(4):
        for bx in range(4):
            av = oracle.lib.orc_h264_block_avail(g.pic, centre, bx, by, 4)
            assert av & TOP and av & LEFT and av & TL
            never = blk_idx(bx, by) in (3, 7, 11, 13, 15)
            assert bool(av & TR) == (not never), (bx, by)
    for k in range(4):
        av = oracle.lib.orc_h264_block_avail(g.pic, centre, (k & 1) * 2, (k >> 1) * 2, 8)
        assert bool(av & TR) == (k != 3)
    # The picture's top-left macroblock: nothing above or to the left.
    assert oracle.lib.orc_h264_block_avail(g.pic, 0, 0, 0, 16) == 0
    # Top-right of the first row's last 4x4 block comes from the macroblock
    # above-right: absent at the picture's right edge.
    assert not oracle.lib.orc_h264_block_avail(g.pic, 5, 3, 0, 4) & TR
    assert oracle.lib.orc_h264_block_avail(g.pic, 4, 3, 0, 4) & TR


def test_constrained_intra_hides_inter_neighbours():
    g = intra_picture(48, 48, 1)
    mbs = g.pic.contents.mbs
    mbs[3].intra = 0  # left of the centre becomes inter
    mbs[1].intra = 0  # above
    av = oracle.lib.orc_h264_block_avail(g.pic, 4, 0, 0, 16)
    assert not av & TOP and not av & LEFT and av & TL
    g.pic.contents.constrained_intra_pred = 0
    assert oracle.lib.orc_h264_block_avail(g.pic, 4, 0, 0, 16) == TOP | LEFT | TL


def test_top_left_of_edge_blocks_comes_from_the_bordering_macroblock():
    # p[-1, -1] of a block in the macroblock's left column (below the first
    # row) lies in the left macroblock; of a block in its top row (right of
    # the first column) in the macroblock above.
    g = intra_picture(48, 48, 1)
    mbs = g.pic.contents.mbs
    mbs[3].intra = 0  # only the left neighbour is inter
    assert not oracle.lib.orc_h264_block_avail(g.pic, 4, 0, 1, 4) & TL
    assert oracle.lib.orc_h264_block_avail(g.pic, 4, 1, 0, 4) & TL
    assert not oracle.lib.orc_h264_block_avail(g.pic, 4, 0, 2, 8) & TL
    mbs[3].intra, mbs[1].intra = rh_dev(), 0  # only the one above is inter
    assert oracle.lib.orc_h264_block_avail(g.pic, 4, 0, 1, 4) & TL
    assert not oracle.lib.orc_h264_block_avail(g.pic, 4, 1, 0, 4) & TL
    assert not oracle.lib.orc_h264_block_avail(g.pic, 4, 2, 0, 8) & TL


def rh_dev():
    return _abi.RV_INTRA_DEVICE


# ---------------------------------------------------------------- whole pictures

def reconstruct(g, refs=None):
    c = g.pic.contents
    if refs is None:
        refs = [rh.gen_planes(c.width, c.height, 8, cf.ref_seed(77, k)) for k in range(2)]
    return oracle.reconstruct(rh.RV_CODEC_H264, g.pic, list(refs) + [None], 0), refs


def test_all_dc_picture_without_neighbours_is_mid_grey():
    g = intra_picture(64, 48, 1)
    c = g.pic.contents
    for a in range(c.n_mbs):
        m = c.mbs[a]
        m.intra_kind, m.chroma_mode = 2, 0
        m.modes[0] = 2
    c.n_blks = 0
    c.deblock = 0
    out, _ = reconstruct(g)
    for k in range(3):
        assert np.all(out[k] == 128)


def _inter_row_then_intra(mode, chroma_mode, residual=None, first=3):
    """MB row 0 (and MBs of row 1 before `first`) inter with whole-sample
    zero vectors and no residual; MBs first..5 of row 1 device Intra_4x4
    with every block in `mode`. Their partitions stay listed: the intra
    prediction replaces what motion compensation wrote there."""
    cfg = dict(cf.H264_C4, width=48, height=32, intra_percent=0, b_picture=0, weighting=0, poc=1, ref_poc=[0, 0],
               deblock=0, intra_device=1)
    g = rh.GenPicture(rh.RV_CODEC_H264, cfg, 11)
    c = g.pic.contents
    for i in range(c.n_parts):
        c.parts[i].mv[0].x = c.parts[i].mv[0].y = 0
    for a in range(first, 6):
        m = c.mbs[a]
        m.intra, m.intra_kind, m.transform_8x8, m.chroma_mode = _abi.RV_INTRA_DEVICE, 0, 0, chroma_mode
        for k in range(16):
            m.modes[k] = mode
    c.n_blks = 0
    if residual is not None:
        bx, by, dc = residual
        blk = c.blks[0]
        blk.x, blk.y, blk.comp, blk.size, blk.has_dc, blk.flags, blk.dc = 16 + bx * 4, 16 + by * 4, 0, 4, 1, 0, dc
        for i in range(16):
            c.coefs[blk.coef_off + i] = 0
        c.n_blks = 1
    return g


def test_vertical_intra4x4_propagates_the_row_above():
    g = _inter_row_then_intra(0, 2)
    out, refs = reconstruct(g)
    y, cb, cr = (np.asarray(p, np.int64) for p in out)
    assert np.array_equal(y[:16], refs[0][0][:16])  # the inter row is the reference crop
    assert np.all(y[16:32] == y[15][None, :])
    assert np.all(cb[8:16] == cb[7][None, :]) and np.all(cr[8:16] == cr[7][None, :])


def test_horizontal_intra4x4_propagates_the_left_column():
    g = _inter_row_then_intra(1, 1, first=4)
    out, refs = reconstruct(g)
    y, cb = np.asarray(out[0], np.int64), np.asarray(out[1], np.int64)
    assert np.array_equal(y[16:32, :16], refs[0][0][16:32, :16])  # MB 3 inter
    for row in range(16, 32):
        assert np.all(y[row, 16:48] == y[row, 15])
    for row in range(8, 16):
        assert np.all(cb[row, 8:24] == cb[row, 7])


def test_residual_is_added_before_the_next_block_is_predicted():
    dc = 64 * 7  # (dc + 32) >> 6 = 7 on every sample of the block
    g = _inter_row_then_intra(0, 2, residual=(0, 0, dc))
    out, _ = reconstruct(g)
    y = np.asarray(out[0], np.int64)
    base = y[15, 16:20]
    # Block (0, 0) of MB 4 gets +7; the blocks below it copy its last row.
    want = np.clip(base + 7, 0, 255)
    assert np.all(y[16:32, 16:20] == want[None, :])
    assert np.all(y[16:32, 20:32] == y[15, 20:32][None, :])
