// Learned preconditioner for the implicit Maxwell solve: a small 3-D convolution U-Net in
// bfloat16 on MFMA, trained on the GPU by its own forward and backward kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

namespace picgpu {

// Parameter layout of the flat float32 vector (shared with the host and the CPU check):
// for each 3x3x3 convolution W[27][cout][cin] then b[cout]; then the 1x1x1 head Wh[3][16], bh[3].
struct ConvSpec {
	int cin, cout, level;
	long long w_off, b_off;
};

class UNet {
public:
	// d: grid dims of the solver vectors (owned nodes), each divisible by 4.
	UNet(const int d[3]);
	~UNet();
	UNet(const UNet&) = delete;
	UNet& operator=(const UNet&) = delete;

	static long long param_count();
	void set_params(const float* host, hipStream_t s);
	void get_params(float* host, hipStream_t s) const;
	void get_grads(float* host, hipStream_t s) const;
	const int* dims() const { return d_; }

	// z = r + s y(r / s, mu / mu_scale), s = ||r|| / sqrt(3V); r, z: [3][V], mu: [9][V] (float32, device)
	void forward(const float* r, const float* mu, float mu_scale, float* z, hipStream_t s);
	// forward + backward of ||z - e||^2 / ||e||^2; gradients are added to the gradient buffer.
	// The loss is left on the device (loss_ptr) so training does not synchronise per sample.
	void accumulate_gradient(const float* r, const float* mu, float mu_scale, const float* e, hipStream_t s);
	const double* loss_ptr() const { return dscal_ + 4; }
	void zero_grad(hipStream_t s);
	// Adam step on the averaged gradient of `batch` samples; refreshes the bf16 weights.
	void adam_step(float lr, int batch, hipStream_t s);

private:
	int d_[3];
	long long V_;
	int adam_t_ = 0;
	std::vector<ConvSpec> convs_;
	float *params_ = nullptr, *grads_ = nullptr, *m_ = nullptr, *v_ = nullptr;
	uint16_t *wf_ = nullptr, *wb_ = nullptr;  // bf16 forward W[t][co][ci] and adjoint W[26-t]^T weights
	// activations (bf16, channels last) and gradients (float32 / bf16)
	uint16_t *x0_ = nullptr, *e1_ = nullptr, *p1_ = nullptr, *e2_ = nullptr, *p2_ = nullptr, *e3_ = nullptr, *d3_ = nullptr,
	         *u2_ = nullptr, *d2_ = nullptr, *u1_ = nullptr, *d1_ = nullptr;
	uint16_t* dz_ = nullptr;                            // masked gradient, bf16, largest level
	uint16_t *dzc_ = nullptr, *xc_ = nullptr;           // channel-major copies of dz_ and a layer's input
	float* rtmp_ = nullptr;                             // slice sums of the weight-gradient reductions
	float *g0_ = nullptr, *g1_ = nullptr, *g2_ = nullptr;  // float32 gradient scratch per level
	float *g0b_ = nullptr, *g1b_ = nullptr;                 // skip-gradient copies
	float* z_ = nullptr;                                  // forward output for the loss
	float* partial_ = nullptr;
	long long partial_len_ = 0;
	double* dscal_ = nullptr;  // [0] sum r^2, [1] s, [2] ||e||^2, [3] ||z - e||^2, [4] loss
	double* dpart_ = nullptr;
	void refresh_weights(hipStream_t s);
	void conv_forward(const ConvSpec& c, const uint16_t* x, uint16_t* y, int lvl, hipStream_t s);
	void conv_backward(const ConvSpec& c, const float* dout, const uint16_t* act, const uint16_t* input, float* din, int lvl,
	                   hipStream_t s);
	long long level_voxels(int lvl) const { return V_ >> (3 * lvl); }
};

}  // namespace picgpu
