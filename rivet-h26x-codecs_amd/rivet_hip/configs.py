"""The five benchmark workloads as generator settings.

Each picture is generated from (workload seed, stream, picture) through
`subseed`, so a stream's pictures are the same alone and in any mix.
Reference pictures live in DPB slots 0 (list 0) and 1 (list 1); a
single-picture job writes slot 2.
"""
from __future__ import annotations

from . import RV_CODEC_H264, RV_CODEC_HEVC, subseed
from ._abi import RV_H264_IMPLICIT

BASE_SEED = 0x5256_4854_3236_3478  # "RVHT26x4"

# 1: HEVC Main 8-bit 4:2:0, 416x240 P-picture.
HEVC_C1 = dict(width=416, height=240, bit_depth=8, b_picture=0, bi_percent=0, explicit_wp=0, intra_percent=0,
               nz_permille=100, lap_b=8, qp_min=22, qp_max=37, mv_range=256, deblock=1, sao=1,
               poc=1, ref_poc=[0, 0], ref_slot=[0, 0])
# Slices and tiles with loop-filter rules seeded per slice and exempt CUs:
# a 416x240 Main10 B-picture (7 x 4 CTBs of 64x64) in 3 x 2 tiles and 5 slices.
SLICED = dict(width=416, height=240, bit_depth=10, b_picture=1, bi_percent=50, explicit_wp=0, intra_percent=30,
              nz_permille=150, lap_b=16, qp_min=22, qp_max=40, mv_range=128, deblock=1, sao=1, poc=2,
              ref_poc=[0, 4], ref_slot=[0, 1], tile_cols=3, tile_rows=2, n_slices=5, lf_across_slices=-1,
              lf_across_tiles=0, slice_vary=1, exempt_percent=15)
# 2: HEVC Main10 1920x1080 B-picture, explicit WP (denominators 6), intra
# 8x8 CUs with DST 4x4 TUs, ~15 % nonzero coefficients at Laplacian b=16.
HEVC_C2 = dict(width=1920, height=1080, bit_depth=10, b_picture=1, bi_percent=60, explicit_wp=1, intra_percent=50,
               nz_permille=150, lap_b=16, qp_min=22, qp_max=37, mv_range=256, deblock=1, sao=1,
               poc=2, ref_poc=[0, 4], ref_slot=[0, 1])
# 3: one stream of the batched 8 x 3840x2160 HEVC Main workload (P or B).
HEVC_C3_P = dict(HEVC_C1, width=3840, height=2160)
HEVC_C3_B = dict(HEVC_C1, width=3840, height=2160, b_picture=1, bi_percent=60, poc=2, ref_poc=[0, 4],
                 ref_slot=[0, 1])
# 4: H.264 High 8-bit 1920x1080 (coded 1088) B-picture, implicit weights.
H264_C4 = dict(width=1920, height=1088, b_picture=1, bi_percent=100, weighting=RV_H264_IMPLICIT, intra_percent=10,
               nz_permille=120, lap_b=12, qp_min=20, qp_max=40, mv_range=512, deblock=1,
               poc=2, ref_poc=[0, 4], ref_slot=[0, 1])

VISIBLE = {(1920, 1088): (1920, 1080), (856, 480): (854, 480)}


def visible_luma(width: int, height: int) -> int:
    w, h = VISIBLE.get((width, height), (width, height))
    return w * h


def c3_stream(stream: int, seed: int = BASE_SEED):
    """Config 3, stream k: its own seed, P or B 50/50."""
    s = subseed(seed, 3, stream, 0)
    cfg = HEVC_C3_B if (s >> 7) & 1 else HEVC_C3_P
    return cfg, subseed(seed, 3, stream, 1)


def single(config: int, seed: int = BASE_SEED):
    """(codec, cfg, picture seed) of configs 1, 2, 4."""
    if config == 1:
        return RV_CODEC_HEVC, HEVC_C1, subseed(seed, 1, 0, 1)
    if config == 2:
        return RV_CODEC_HEVC, HEVC_C2, subseed(seed, 2, 0, 1)
    if config == 4:
        return RV_CODEC_H264, H264_C4, subseed(seed, 4, 0, 1)
    raise ValueError(config)


def ref_seed(pic_seed: int, slot: int) -> int:
    return subseed(pic_seed, 9, slot, 0)
