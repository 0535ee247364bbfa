"""Slices, tiles and loop-filter exemptions in the oracle, pinned by
properties the rules fix rather than by a second copy of them:
boundary_strengths' slice and tile checks and per-slice
deblocking_disabled (src/hevc/deblock.rs:105-165), the slice beta / tc
offsets (:186-201, 264-294), the nDp / nDq exemptions (:274-275, 440-491),
SAO's neighbour rules (Neighbours::of, src/hevc/sao.rs:193-238) and its
exempt samples (:140-151, 280-283). Pictures come from the seeded
generator; its slice, tile and exemption streams leave every other stream
as it is, so a sliced picture and its one-slice twin hold the same CUs,
PUs and TUs."""
import ctypes as C

import numpy as np
import pytest

import oracle
import rivet_hip as rh
from rivet_hip import _abi
from rivet_hip import configs as cf

BASE = dict(cf.HEVC_C1, intra_percent=20, nz_permille=200, qp_min=30, qp_max=45)
TILED = dict(n_slices=4, tile_cols=3, tile_rows=2)


def gen(seed, **kw):
    return rh.GenPicture(rh.RV_CODEC_HEVC, dict(BASE, **kw), seed)


def run(g, seed, stop=0):
    refs = [rh.gen_planes(BASE["width"], BASE["height"], 8, cf.ref_seed(seed, k)) for k in range(2)]
    return oracle.reconstruct(rh.RV_CODEC_HEVC, g.pic, refs + [None], stop)


def tables(g):
    p = g.pic.contents
    ctb = 1 << p.log2_ctb
    wc, hc = (p.width + ctb - 1) // ctb, (p.height + ctb - 1) // ctb
    a = lambda ptr: np.ctypeslib.as_array(ptr, (hc, wc)).copy()  # noqa: E731
    return a(p.ctb_slice), a(p.ctb_slice_addr), a(p.ctb_tile), ctb


def edge_ctbs(g, vertical):
    """Per 4x4 block: whether its left (vertical) / top edge is a CTB
    boundary between different slices (by SliceAddrRs) or tiles, and
    whether its CTB's slice has deblocking disabled."""
    sl, addr, tile, ctb = tables(g)
    p = g.pic.contents
    h4, w4 = p.height // 4, p.width // 4
    ys, xs = np.mgrid[0:h4, 0:w4] * 4
    cy, cx = ys // ctb, xs // ctb
    if vertical:
        on = (xs % ctb == 0) & (xs > 0)
        py, px = cy, np.maximum(cx - 1, 0)
    else:
        on = (ys % ctb == 0) & (ys > 0)
        py, px = np.maximum(cy - 1, 0), cx
    slice_cross = on & (addr[cy, cx] != addr[py, px])
    tile_cross = on & (tile[cy, cx] != tile[py, px])
    disabled = np.array([p.slices[k].deblocking_disabled for k in range(p.n_slices)], bool)[sl[cy, cx]]
    return slice_cross, tile_cross, disabled


def test_slices_that_allow_every_crossing_equal_one_slice():
    one, many = gen(31), gen(31, lf_across_slices=1, lf_across_tiles=1, **TILED)
    assert many.pic.contents.ctb_slice and not one.pic.contents.ctb_slice
    for a, b in zip(oracle.hevc_bs_map(one.pic), oracle.hevc_bs_map(many.pic)):
        assert (a == b).all()
    for a, b in zip(run(one, 31), run(many, 31)):
        assert (a == b).all()


def test_no_bs_across_refused_slice_and_tile_boundaries():
    one, many = gen(32), gen(32, lf_across_slices=0, lf_across_tiles=0, **TILED)
    crossed = 0
    for k, (a, b) in enumerate(zip(oracle.hevc_bs_map(one.pic), oracle.hevc_bs_map(many.pic))):
        sc, tc, _ = edge_ctbs(many, vertical=k == 0)
        cut = sc | tc
        crossed += int((a[cut] != 0).sum())
        assert (b[cut] == 0).all() and (b[~cut] == a[~cut]).all()
    assert crossed > 20  # the one-slice twin filters those edges


def test_tiles_only_cut_when_their_flag_is_off():
    one, many = gen(33), gen(33, lf_across_slices=1, lf_across_tiles=0, **TILED)
    for k, (a, b) in enumerate(zip(oracle.hevc_bs_map(one.pic), oracle.hevc_bs_map(many.pic))):
        _, tc, _ = edge_ctbs(many, vertical=k == 0)
        assert (b[tc] == 0).all() and (b[~tc] == a[~tc]).all()


def test_slices_with_deblocking_disabled_carry_no_bs():
    one, many = gen(34), gen(34, lf_across_slices=1, lf_across_tiles=1, **TILED)
    p = many.pic.contents
    p.slices[1].deblocking_disabled = p.slices[3].deblocking_disabled = 1
    for k, (a, b) in enumerate(zip(oracle.hevc_bs_map(one.pic), oracle.hevc_bs_map(many.pic))):
        _, _, dis = edge_ctbs(many, vertical=k == 0)
        assert (b[dis] == 0).all() and (b[~dis] == a[~dis]).all()
        assert (a[dis] != 0).any()


@pytest.mark.parametrize("beta,tc", [(6, -4), (-6, 2)])
def test_slice_offsets_equal_the_same_picture_offsets(beta, tc):
    """Every slice with offsets (beta, tc) filters as one slice with those
    picture offsets; and they differ from the default offsets' result."""
    one, many = gen(35), gen(35, lf_across_slices=1, lf_across_tiles=1, **TILED)
    p = many.pic.contents
    for k in range(p.n_slices):
        p.slices[k].beta_offset, p.slices[k].tc_offset = beta, tc
    default = run(one, 35)
    one.pic.contents.beta_offset, one.pic.contents.tc_offset = beta, tc
    want = run(one, 35)
    got = run(many, 35)
    for a, b in zip(want, got):
        assert (a == b).all()
    assert any((a != b).any() for a, b in zip(want, default))


def exempt_masks(g):
    """Luma masks of samples in CUs exempt from both loop filters and from
    deblocking only (palette)."""
    p = g.pic.contents
    lf = np.zeros((p.height, p.width), bool)
    dbk = np.zeros_like(lf)
    for i in range(p.n_cus):
        cu = p.cus[i]
        s = 1 << cu.log2_size
        if cu.exempt & _abi.RV_CU_LF_EXEMPT:
            lf[cu.y:cu.y + s, cu.x:cu.x + s] = True
        elif cu.exempt & _abi.RV_CU_DBK_EXEMPT:
            dbk[cu.y:cu.y + s, cu.x:cu.x + s] = True
    return lf, dbk


def test_exempt_samples_keep_their_values():
    """Samples of exempt CUs keep the reconstructed value through deblocking
    (and, bit 0, through SAO), where the twin without exemptions changes
    them."""
    g, plain = gen(36, exempt_percent=30), gen(36)
    lf, dbk = exempt_masks(g)
    assert lf.any() and dbk.any()
    pre, deb, fin = run(g, 36, 2), run(g, 36, 3), run(g, 36)
    pdeb, pfin = run(plain, 36, 3), run(plain, 36)
    for c in range(3):
        m_lf, m_dbk = (lf, dbk) if c == 0 else (lf[::2, ::2], dbk[::2, ::2])
        assert (fin[c][m_lf] == pre[c][m_lf]).all()
        assert (deb[c][m_dbk | m_lf] == pre[c][m_dbk | m_lf]).all()
    assert (pdeb[0][lf | dbk] != pre[0][lf | dbk]).any() and (pfin[0][lf] != pre[0][lf]).any()
    # Palette CUs are left alone by deblocking only: SAO still runs there.
    assert (fin[0][dbk] != deb[0][dbk]).any()


def test_sao_does_not_read_across_refused_boundaries():
    """A horizontal-class edge-offset CTB whose left neighbour may not be
    read leaves its first column as deblocked; with the neighbour readable
    (the one-slice twin) the same column changes somewhere."""
    seed = 37
    many = gen(seed, lf_across_slices=0, lf_across_tiles=0, **TILED)
    one = gen(seed)
    deb, fin = run(many, seed, 3), run(many, seed)
    odeb, ofin = run(one, seed, 3), run(one, seed)
    sl, addr, tile, ctb = tables(many)
    p = many.pic.contents
    hc, wc = sl.shape
    checked = changed = 0
    for ry in range(hc):
        for rx in range(1, wc):
            sp = p.sao[3 * (ry * wc + rx)]
            if sp.type_idx != 2 or (sp.band_or_class & 3) != 0:
                continue
            if addr[ry, rx] == addr[ry, rx - 1] and tile[ry, rx] == tile[ry, rx - 1]:
                continue
            x, y0, y1 = rx * ctb, ry * ctb, min((ry + 1) * ctb, p.height)
            assert (fin[0][y0:y1, x] == deb[0][y0:y1, x]).all()
            checked += 1
            changed += int((ofin[0][y0:y1, x] != odeb[0][y0:y1, x]).any())
    assert checked and changed
