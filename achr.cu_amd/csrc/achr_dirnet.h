// Direction network state (see achr_dirnet.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace achr {

// Parameters are one FP32 array: W1 (H x L), b1 (H), W2 (H x H), b2 (H),
// W3 (O x H), b3 (O); weights row major (out x in).
struct DirNet {
  int L = 0, H = 0, O = 0;
  float* params = nullptr;
  float* grads = nullptr;
  float* adam_m = nullptr;
  float* adam_v = nullptr;
  int adam_t = 0;
  // activations for a batch of B latents (2B rows through the layers)
  int cap_B = 0;
  float *z = nullptr, *z2 = nullptr, *a1 = nullptr, *h1 = nullptr, *a2 = nullptr, *h2 = nullptr;
  float *f = nullptr, *v = nullptr, *dv = nullptr, *df = nullptr;
  float* scratch = nullptr;
  size_t scratch_cap = 0;
  void* blas = nullptr;     // rocblas_handle of the FP32 training GEMMs (created on first use)
  float* cov = nullptr;     // O x O: V^T V, then the error V^T V / B - C
  float* target = nullptr;  // O x O target covariance C
  float target_norm2 = 1.0f;

  size_t param_count() const;
  hipError_t alloc(int latent, int width, int outdim);
  void release();
  hipError_t ensure_batch(int B);
  hipError_t ensure_cov();
  void init_params(uint64_t seed, std::vector<float>& host) const;
  hipError_t gaussian_latents(int B, uint64_t seed, uint32_t iter, hipStream_t s);
  hipError_t forward(int B, int prec, hipStream_t s);
  hipError_t backward(int B, float cnorm2, hipStream_t s);
  hipError_t adam_step(float lr, int t, hipStream_t s);
};

}  // namespace achr
