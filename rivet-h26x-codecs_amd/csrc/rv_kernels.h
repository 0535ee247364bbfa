// Launch entry points of the kernel translation units.
#pragma once

#include <hip/hip_runtime.h>

#include "rv_device.h"

namespace rv {

// HEVC (hevc.hip).
void hevc_stage_items(HevcPicDev& P);
// Stage kernels come in variants (variant 0 serves any picture; the
// others only pictures of one depth class).
int hevc_stage_variants(int stage);
int hevc_variant_depth(int stage, int variant);  // 0 any, 1 8-bit only, 2 deeper only
const char* hevc_variant_name(int stage, int variant);
hipError_t hevc_launch_stage(int stage, const BatchDev& b, int total, hipStream_t st, int variant = 0);
// One per-entry call: `blocks` workgroups of 256 threads.
hipError_t hevc_entry_launch(const HevcEntry& e, int blocks, hipStream_t st);

// H.264 (h264.hip).
void h264_stage_items(H264PicDev& P);
hipError_t h264_launch_stage(int stage, const BatchDev& b, int total, hipStream_t st);
// 64-bit words of the deblocking exchange of one picture.
size_t h264_deblock_xch_words(int mbw, int mbh);
hipError_t h264_entry_launch(const H264Entry& e, int blocks, hipStream_t st);

// Picture hash (hash.hip): the three planes of one picture, each `w` x `h`
// samples of `ss` bytes at a row pitch of `pitch` bytes.
struct HashPlanes {
    const uint8_t* p[3];
    int w[3], h[3], pitch[3];
    int ss;
};
// Device scratch for hash_launch (including the 48-byte result at its
// start: 3 x 16 bytes, 3 x uint16 or 3 x uint32 by type).
size_t hash_scratch_bytes(int type, const HashPlanes& H);
hipError_t hash_launch(int type, const HashPlanes& H, void* scratch, hipStream_t st);

// The MD5 of each picture packed planar (Y, Cb, Cr rows cropped to w x h,
// samples of ss bytes, low byte first): 16 bytes per picture into out.
struct FrameMd5 {
    const uint8_t* p[3];
    int w[3], h[3], pitch[3];  // pitch in bytes
    int ss;
};
hipError_t frame_md5_launch(const FrameMd5* frames, int n, uint8_t* out, hipStream_t st);

// Plain copy of `bytes` (a multiple of 4096) with 16-byte vectors.
hipError_t launch_copy_probe(void* dst, const void* src, int64_t bytes, hipStream_t st);

}  // namespace rv
