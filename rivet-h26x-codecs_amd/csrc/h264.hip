// H.264 picture reconstruction on gfx950 (stages filled in below).
#include <hip/hip_runtime.h>

#include "rv_device.h"
#include "rv_kernels.h"

namespace rv {

void h264_stage_items(H264PicDev& P) {
    for (int s = 0; s < kStages; ++s) P.items[s] = 0;
}

hipError_t h264_launch_stage(int, const BatchDev&, int total, hipStream_t) {
    return total <= 0 ? hipSuccess : hipErrorInvalidValue;
}

}  // namespace rv
