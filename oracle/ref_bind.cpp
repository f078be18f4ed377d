// Binding that exposes the reference's own fp16/bf16 converters
// (inc/ff_utils.h:90-127, compiled from the TensileLite tree by oracle/Makefile)
// to the oracle tests through a C ABI.  Test infrastructure only: the output
// lands in oracle/_ref/ and is loaded only by tests/.
#include <stdint.h>
#include "ff_utils.h"

extern "C" {
int16_t ref_cvt_f32_to_f16(float f) { return feifei::cvtFP32toFP16(f); }
float ref_cvt_f16_to_f32(uint16_t h) { return feifei::cvtFP16toFP32(h); }
int16_t ref_cvt_f32_to_bf16(float f) { return feifei::cvtFP32toBF16(f); }
float ref_cvt_bf16_to_f32(uint16_t b) { return feifei::cvtBF16toFP32(b); }
void ref_cvt_f32_to_f16_n(const float* s, uint16_t* d, int64_t n) {
  for (int64_t i = 0; i < n; i++) d[i] = (uint16_t)feifei::cvtFP32toFP16(s[i]);
}
}
