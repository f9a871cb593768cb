// The attention backward kernels (dK/dV and dQ) in a translation unit of
// their own, so that they can be built with other code generation options
// than the forward kernel: with MFMA accumulators in VGPRs, the pipelined
// backward kernels copy far fewer values between AGPRs and VGPRs around
// their softmax, while the forward kernel is faster without.
#define EF_ATTN_BWD_TU 1
#include "attention.hip"
