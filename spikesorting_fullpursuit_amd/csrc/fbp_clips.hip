// Multi-channel spike clips: the gather of get_clips (SF/processing/clip_utils.py:501-632,
// SF = spikesorting_fullpursuit/), which full binary pursuit runs over every found spike and
// every channel (SF/overlap/full_binary_pursuit.py:495-502).
//
//   out[k][q * W + j] = V[chan[q]][ev[k] + w0 + j]     (W = w1 - w0, q < nchan, j < W)
//
// A gather moves values without arithmetic, so it runs on raw 2-, 4- or 8-byte elements and
// the clips are bit-identical to the recording whatever its type (int16, float32, float64).
// One thread per output element, consecutive threads on consecutive samples of one channel
// row: reads and writes are coalesced.  Samples outside [0, N) and channels outside [0, C)
// read as 0 (the host passes only events whose clips lie inside the recording, as get_clips
// does, and validated channels).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "../../include/fbp.h"

namespace fbp {
namespace clips {

template <class E>
__global__ __launch_bounds__(256) void k_get_clips(const E* __restrict__ V, int64_t ldv, int64_t N, int C,
                                                   const int32_t* __restrict__ chan, int nchan,
                                                   const int64_t* __restrict__ ev, int64_t nev, int w0, int W,
                                                   E* __restrict__ out)
{
    const int64_t per = (int64_t)nchan * W;
    const int64_t total = nev * per;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t k = e / per;
        const int64_t r = e - k * per;
        const int q = (int)(r / W), j = (int)(r - (int64_t)q * W);
        const int64_t x = ev[k] + w0 + j;
        const int c = chan[q];
        out[e] = (x >= 0 && x < N && c >= 0 && c < C) ? V[(int64_t)c * ldv + x] : E(0);
    }
}

thread_local std::string t_err;

}  // namespace clips
}  // namespace fbp

using namespace fbp::clips;

extern "C" const char* fbp_clips_last_error(void) { return t_err.c_str(); }

extern "C" int fbp_get_clips(const void* V, int32_t elem_bytes, int64_t ldv, int64_t N, int32_t C,
                             const int32_t* chans, int32_t nchan, const int64_t* event_indices, int64_t n_events,
                             int32_t w0, int32_t w1, void* clips, void* stream)
{
    if (elem_bytes != 2 && elem_bytes != 4 && elem_bytes != 8) {
        t_err = "elem_bytes must be 2, 4 or 8";
        return FBP_E_INVALID;
    }
    if (N <= 0 || C <= 0 || nchan <= 0 || n_events < 0 || ldv < N) {
        t_err = "N, C and nchan must be positive, n_events >= 0 and ldv >= N";
        return FBP_E_INVALID;
    }
    if (w1 <= w0) {
        t_err = "the clip window must hold at least one sample (w1 > w0)";
        return FBP_E_INVALID;
    }
    if (n_events == 0) return FBP_OK;
    if (!V || !chans || !event_indices || !clips) {
        t_err = "null device pointer";
        return FBP_E_INVALID;
    }
    const int64_t total = n_events * (int64_t)nchan * (w1 - w0);
    const int64_t blocks = std::min<int64_t>((total + 255) / 256, (int64_t)1 << 20);
    const dim3 g((unsigned)blocks), b(256);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int W = (int)(w1 - w0);
    if (elem_bytes == 2)
        hipLaunchKernelGGL(k_get_clips<uint16_t>, g, b, 0, s, static_cast<const uint16_t*>(V), ldv, N, (int)C, chans,
                           (int)nchan, event_indices, n_events, (int)w0, W, static_cast<uint16_t*>(clips));
    else if (elem_bytes == 4)
        hipLaunchKernelGGL(k_get_clips<uint32_t>, g, b, 0, s, static_cast<const uint32_t*>(V), ldv, N, (int)C, chans,
                           (int)nchan, event_indices, n_events, (int)w0, W, static_cast<uint32_t*>(clips));
    else
        hipLaunchKernelGGL(k_get_clips<uint64_t>, g, b, 0, s, static_cast<const uint64_t*>(V), ldv, N, (int)C, chans,
                           (int)nchan, event_indices, n_events, (int)w0, W, static_cast<uint64_t*>(clips));
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        t_err = std::string("fbp_get_clips: ") + hipGetErrorString(e);
        return FBP_E_HIP;
    }
    return FBP_OK;
}
