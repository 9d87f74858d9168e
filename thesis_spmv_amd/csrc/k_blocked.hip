// CSB and BlockCOH device formats (built in a later step).
#include "api_internal.hpp"

namespace pwm_impl {

int build_csb(pwm_matrix* h, const int32_t*, const int32_t*, const double*) {
    return fail(h, PWM_ERR_UNSUPPORTED, "CSB format not built yet");
}
int build_blockcoh(pwm_matrix* h, const int32_t*, const int32_t*, const double*) {
    return fail(h, PWM_ERR_UNSUPPORTED, "BlockCOH format not built yet");
}
hipError_t csb_launch(pwm_matrix*, const void*, void*, const pwmk::PowerArgs*) { return hipErrorNotSupported; }
hipError_t coh_launch(pwm_matrix*, const void*, void*, const pwmk::PowerArgs*) { return hipErrorNotSupported; }
int64_t csb_partials(const pwm_matrix*) { return 0; }
int64_t coh_partials(const pwm_matrix*) { return 0; }

}  // namespace pwm_impl

extern "C" int pwm_csb_layout(pwm_matrix* h, int32_t*, int64_t*, uint32_t*, double*, int32_t*) {
    return pwm_impl::fail(h, PWM_ERR_UNSUPPORTED, "CSB format not built yet");
}
extern "C" int pwm_blockcoh_band(pwm_matrix* h, int, int64_t*, int32_t*, uint16_t*, uint16_t*, double*, int16_t*,
                                 int16_t*, int32_t*) {
    return pwm_impl::fail(h, PWM_ERR_UNSUPPORTED, "BlockCOH format not built yet");
}
