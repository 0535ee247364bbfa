"""Golden cases: seeded pictures of the five workloads.

TEST INFRASTRUCTURE. For each case, tests/golden/reconstruction.json holds
hashes of three things:
- the generated work lists;
- the two reference pictures;
- the oracle's reconstruction.

make_golden.py writes that file. tests/test_golden.py re-derives every
hash on the CPU, so a change to the generator or the oracle shows up.
tests/test_gpu_golden.py checks the device's reconstruction against the
committed output hashes directly.

The reference has no vectors for this path to take: its dsp tests are
differential, and its decoder hashes need its parser and vendored streams.
These fixtures are this build's own.
"""
from __future__ import annotations

import ctypes as C
import hashlib
import json
from pathlib import Path

import rivet_hip as rh
from rivet_hip import _abi
from rivet_hip import configs as cf

HERE = Path(__file__).resolve().parent
GOLDEN = HERE / "reconstruction.json"


def _h264_small(kind: str):
    cfg = dict(cf.H264_C4, width=352, height=288)
    if kind == "p":
        cfg.update(b_picture=0, weighting=0, poc=1, ref_poc=[0, 0])
    elif kind == "intra_device":
        cfg.update(bi_percent=50, intra_percent=40, intra_device=1, constrained_intra_pred=1, levels=1)
    else:
        cfg.update(bi_percent=50, intra_percent=30)
    return cfg


# name -> (codec, generator settings, picture seed)
CASES = {
    **{f"hevc_c1_seed{k}": (rh.RV_CODEC_HEVC, cf.HEVC_C1, rh.subseed(cf.BASE_SEED, 1, k, 1)) for k in range(3)},
    "hevc_c2": cf.single(2),
    **{f"hevc_c3_stream{k}": (rh.RV_CODEC_HEVC, *cf.c3_stream(k)) for k in range(2)},
    "hevc_b_wp_intra_8bit_416x240": (rh.RV_CODEC_HEVC, dict(cf.HEVC_C2, width=416, height=240, bit_depth=8), 12345),
    "hevc_slices_tiles_exempt_10bit_416x240": (rh.RV_CODEC_HEVC, cf.SLICED, 0x5117),
    "h264_c4": cf.single(4),
    "h264_p_352x288": (rh.RV_CODEC_H264, _h264_small("p"), 0x264),
    "h264_b_mixed_352x288": (rh.RV_CODEC_H264, _h264_small("b_mixed"), 0x265),
    "h264_b_intra_device_352x288": (rh.RV_CODEC_H264, _h264_small("intra_device"), 0x266),
}


def _bytes(ptr, n: int, st) -> bytes:
    return C.string_at(ptr, n * C.sizeof(st)) if n else b""


def worklist_digest(codec: int, pic) -> str:
    """SHA-256 of a picture description: its scalar fields and every list."""
    c = pic.contents
    h = hashlib.sha256()
    def plain(v):
        return [plain(x) for x in v] if isinstance(v, C.Array) else v

    scalars = {}
    for name, _ in type(c)._fields_:
        v = getattr(c, name)
        if isinstance(v, (int, C.Array)):
            scalars[name] = plain(v)
    h.update(json.dumps(scalars, sort_keys=True, default=str).encode())
    if codec == rh.RV_CODEC_HEVC:
        ctb = 1 << c.log2_ctb
        n_sao = ((c.width + ctb - 1) // ctb) * ((c.height + ctb - 1) // ctb) * 3
        n_ctb = n_sao // 3 if c.ctb_slice else 0
        parts = [(c.cus, c.n_cus, _abi.HevcCu), (c.pus, c.n_pus, _abi.HevcPu), (c.tus, c.n_tus, _abi.HevcTu),
                 (c.nz, c.n_nz, _abi.Nz), (c.intra_pred, c.n_intra_pred, C.c_uint16),
                 (c.sao, n_sao if c.sao else 0, _abi.Sao), (c.slices, c.n_slices if n_ctb else 0, _abi.HevcSlice),
                 (c.ctb_slice, n_ctb, C.c_uint16), (c.ctb_slice_addr, n_ctb, C.c_uint32),
                 (c.ctb_tile, n_ctb, C.c_uint16), (c.ctb_ts, n_ctb, C.c_uint32)]
    else:
        parts = [(c.mbs, c.n_mbs, _abi.H264Mb), (c.parts, c.n_parts, _abi.H264Part), (c.blks, c.n_blks, _abi.H264Blk),
                 (c.coefs, c.n_coefs, C.c_int32), (c.intra_pred, c.n_intra_pred, C.c_uint16)]
    for ptr, n, st in parts:
        h.update(_bytes(ptr, n, st))
    return h.hexdigest()


def planes_digest(planes) -> list:
    return [hashlib.sha256(p.tobytes()).hexdigest() for p in planes]


def references(codec: int, cfg: dict, seed: int):
    bd = cfg.get("bit_depth", 8)
    return [rh.gen_planes(cfg["width"], cfg["height"], bd, cf.ref_seed(seed, k)) for k in range(2)]


def compute(name: str, with_oracle: bool = True) -> dict:
    import oracle

    codec, cfg, seed = CASES[name]
    g = rh.GenPicture(codec, cfg, seed)
    refs = references(codec, cfg, seed)
    out = {
        "codec": "hevc" if codec == rh.RV_CODEC_HEVC else "h264",
        "seed": seed,
        "width": cfg["width"],
        "height": cfg["height"],
        "bit_depth": cfg.get("bit_depth", 8),
        "worklist_sha256": worklist_digest(codec, g.pic),
        "references_sha256": [planes_digest(r) for r in refs],
    }
    if with_oracle:
        out["output_sha256"] = planes_digest(oracle.reconstruct(codec, g.pic, refs + [None], 0))
    return out


def load() -> dict:
    return json.loads(GOLDEN.read_text())
