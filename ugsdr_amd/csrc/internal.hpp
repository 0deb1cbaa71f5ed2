// Declarations shared between the library's translation units.
#pragma once

#include "ctx.hpp"

namespace ugsdr_hip {

int acquire_impl(ugsdr_hip_ctx* ctx, const ugsdr_acq_request* r, ugsdr_acq_result* out, float* surfaces,
                 float* peaks);
int acquire_reduced(ugsdr_hip_ctx* ctx, const ugsdr_acq_request* r, ugsdr_acq_result* out, float* surfaces,
                    float* peaks);
// DFT of d (n complex, device) in place on ctx's stream, tmp n complex of
// scratch; the inverse divides by n (DiscreteFourierTransform conventions)
int dft_device(ugsdr_hip_ctx* ctx, float2* d, float2* tmp, int64_t n, int inverse);
// conj(DFT(replica)) of length n_fft into out (device), from chips or samples
int replica_spectrum(ugsdr_hip_ctx* ctx, const int8_t* chips, int64_t n_chips, int64_t repeats, const float* samples,
                     int64_t n_fft, float2* out);
// whether the upsampled-code search handles a code of code_len chips held
// for U samples each (corr1023.hip)
bool upsampled_code_supported(int64_t code_len, int64_t U);
// the FP32 upsampled-code search's storage order of a code's conj spectrum
// (in place, on ctx's stream): 1023-chip codes in the prime-factor input
// order of the wave-autonomous kernels, 1023 M-chip codes with M = 4, 8, 16
// in that of the prime-factor phase transforms, others unchanged
// (corr1023.hip)
int code_spectrum_layout(ugsdr_hip_ctx* ctx, int64_t code_len, float2* spec);
// resolves the no-alloc workspace's device address on `device` (current),
// once per device (noalloc.hip)
int noalloc_scratch_init(int device);

}  // namespace ugsdr_hip

namespace ugsdr_hip {
// converts an FP32 spectrum in rep.spec to the reduced storage type in place
// (new allocation), recording the scale
int replica_reduce(ugsdr_hip_ctx* ctx, Replica& rep, int precision);
// reduced precision on the upsampled-code form: codes of N1 * 341 chips
// (N1 in 12, 24) held for U samples (U a Z-prep factor)
bool reduced_ucode_supported(int64_t code_len, int64_t U);
// the replica's code spectrum to 16-bit pairs scaled by 1/sqrt(code_len)
// (reduced-precision replicas)
int code_spec_reduce(ugsdr_hip_ctx* ctx, Replica& rep, int precision);
// in place, on st: nblocks blocks of code_len 16-bit words from natural order
// into the reduced upsampled-code search's prime-factor storage order
int upm_permute(uint32_t* blocks, int64_t nblocks, int64_t code_len, hipStream_t st);
// copies the spectrum back as FP32 complex (undoing the storage scale)
int replica_read(ugsdr_hip_ctx* ctx, const Replica& rep, float* out);
}  // namespace ugsdr_hip
