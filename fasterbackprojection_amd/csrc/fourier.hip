// Frequency-domain paths: temporal band-pass filter_H, LoGFFT and LCT.
#include "fbp_internal.h"

namespace fbp {
void fft_cache_destroy(fbp_context) {}
}  // namespace fbp

using namespace fbp;

extern "C" {

uint32_t fbp_filter_h_padded_length(uint32_t nt, float delta_t, float wl_mean, float wl_sigma) {
    (void)nt; (void)delta_t; (void)wl_mean; (void)wl_sigma;
    return 0;
}
fbp_status fbp_filter_h(fbp_context ctx, float*, uint64_t, uint32_t, float, float, float, int32_t,
                        void*) {
    return fail(ctx, FBP_ERR_UNSUPPORTED, "filter_h: not built yet");
}
fbp_status fbp_filter_log_fft(fbp_context ctx, float*, fbp_uvec3, float, void*) {
    return fail(ctx, FBP_ERR_UNSUPPORTED, "log_fft: not built yet");
}
fbp_status fbp_lct(fbp_context ctx, const float*, const fbp_lct_params*, float*, void*) {
    return fail(ctx, FBP_ERR_UNSUPPORTED, "lct: not built yet");
}

}  // extern "C"
