/* FP32 transcription of the reference kernels: the timed CPU baseline
 * (TEST INFRASTRUCTURE ONLY). */
#define REAL float
#define ORC_F64 0
#define SFX(name) name##_f32
#include "tss_oracle_phys.inc"
