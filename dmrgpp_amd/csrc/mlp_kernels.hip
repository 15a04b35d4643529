// The learned Davidson preconditioner: a per-element MLP (4 -> 64 -> 64 -> 1, tanh), FP32,
// with its inference, gradient and optimizer kernels.
//
// Parameters are one flat FP32 array: W1 (64 x 4, row major: W1[o][i]), b1 (64), W2 (64 x 64,
// W2[o][i]), b2 (64), W3 (64), b3 (1). A forward pass per element x (4 features):
//   h1 = tanh(W1 x + b1), h2 = tanh(W2 h1 + b2), y = W3 . h2 + b3.
// The gradient of the mean squared error over a batch is accumulated per workgroup in a fixed
// sample order and the workgroup partials are summed in index order, so a gradient is bitwise
// reproducible.
#include "dmk_internal.h"
#include "mlp.h"

#include <hip/hip_runtime.h>

namespace dmk {

namespace {

constexpr int kT      = 256; // threads per workgroup
constexpr int kPass   = 64;  // samples per pass of the gradient kernel
constexpr int kLd     = kMlpHidden + 1; // padded LDS row

__device__ inline void mlp_forward_one(const float* __restrict__ P, const float x[kMlpIn],
                                       float h1[kMlpHidden], float h2[kMlpHidden], float& y)
{
	const float* W1 = P + kMlpOffW1;
	const float* b1 = P + kMlpOffB1;
	const float* W2 = P + kMlpOffW2;
	const float* b2 = P + kMlpOffB2;
	const float* W3 = P + kMlpOffW3;
#pragma unroll
	for (int o = 0; o < kMlpHidden; ++o) {
		float a = b1[o];
#pragma unroll
		for (int i = 0; i < kMlpIn; ++i)
			a += W1[o * kMlpIn + i] * x[i];
		h1[o] = tanhf(a);
	}
	float out = P[kMlpOffB3];
	for (int o = 0; o < kMlpHidden; ++o) {
		float a = b2[o];
#pragma unroll 16
		for (int i = 0; i < kMlpHidden; ++i)
			a += W2[o * kMlpHidden + i] * h1[i];
		h2[o] = tanhf(a);
		out += W3[o] * h2[o];
	}
	y = out;
}

} // namespace

// y[s] = MLP(x[s]) for count samples (features row major, kMlpIn per sample).
__global__ __launch_bounds__(kT) void mlp_forward_kernel(const float* __restrict__ params,
                                                         const float* __restrict__ x, int64_t count,
                                                         float* __restrict__ y)
{
	__shared__ float P[kMlpParams];
	for (int i = threadIdx.x; i < kMlpParams; i += kT)
		P[i] = params[i];
	__syncthreads();
	for (int64_t s = int64_t(blockIdx.x) * kT + threadIdx.x; s < count; s += int64_t(gridDim.x) * kT) {
		float xi[kMlpIn], h1[kMlpHidden], h2[kMlpHidden], out;
#pragma unroll
		for (int i = 0; i < kMlpIn; ++i)
			xi[i] = x[s * kMlpIn + i];
		mlp_forward_one(P, xi, h1, h2, out);
		y[s] = out;
	}
}

// Per-workgroup partial gradient of L = (1/N) sum_s (y_s - t_s)^2 over samples
// [0, count) of (x, t), N = count: partial[wg * (kMlpParams + 1) + p], the last entry holding
// the partial sum of squared errors. Workgroup wg takes passes of kPass samples wg, wg + G, ...
// in order. Each pass runs layer by layer over the pass's samples with all threads (every
// activation and delta lives in LDS), then folds the pass into per-thread accumulators.
__global__ __launch_bounds__(kT) void mlp_grad_kernel(const float* __restrict__ params,
                                                      const float* __restrict__ x,
                                                      const float* __restrict__ t, int64_t count,
                                                      float* __restrict__ partial)
{
	__shared__ float P[kMlpParams];
	__shared__ float X[kPass][kMlpIn];
	__shared__ float H1[kPass][kLd];
	__shared__ float H2[kPass][kLd];
	__shared__ float DA2[kPass][kLd];
	__shared__ float DA1[kPass][kLd];
	__shared__ float DY[kPass];
	__shared__ float SE[kPass];
	for (int i = threadIdx.x; i < kMlpParams; i += kT)
		P[i] = params[i];
	const int   tid   = threadIdx.x;
	const float scale = 2.0f / float(count);
	constexpr int Q   = kPass * kMlpHidden / kT; // (sample, unit) entries per thread per layer
	float acc2[kMlpHidden * kMlpHidden / kT];
#pragma unroll
	for (int q = 0; q < kMlpHidden * kMlpHidden / kT; ++q)
		acc2[q] = 0.0f;
	float acc1 = 0.0f, accv = 0.0f, accse = 0.0f;
	const int64_t npass = (count + kPass - 1) / kPass;
	for (int64_t pass = blockIdx.x; pass < npass; pass += gridDim.x) {
		const int64_t s0    = pass * kPass;
		const int     valid = int(count - s0 < kPass ? count - s0 : kPass);
		__syncthreads();
		// Features of the pass (zero rows past the end).
		for (int e = tid; e < kPass * kMlpIn; e += kT) {
			const int sm = e / kMlpIn;
			X[sm][e % kMlpIn] = (sm < valid) ? x[(s0 + sm) * kMlpIn + e % kMlpIn] : 0.0f;
		}
		__syncthreads();
		// Layer 1.
#pragma unroll 1
		for (int q = 0; q < Q; ++q) {
			const int e = tid + kT * q, sm = e / kMlpHidden, o = e % kMlpHidden;
			float     a = P[kMlpOffB1 + o];
#pragma unroll
			for (int i = 0; i < kMlpIn; ++i)
				a += P[kMlpOffW1 + o * kMlpIn + i] * X[sm][i];
			H1[sm][o] = tanhf(a);
		}
		__syncthreads();
		// Layer 2.
#pragma unroll 1
		for (int q = 0; q < Q; ++q) {
			const int e = tid + kT * q, sm = e / kMlpHidden, o = e % kMlpHidden;
			float     a = P[kMlpOffB2 + o];
#pragma unroll 8
			for (int i = 0; i < kMlpHidden; ++i)
				a += P[kMlpOffW2 + o * kMlpHidden + i] * H1[sm][i];
			H2[sm][o] = tanhf(a);
		}
		__syncthreads();
		// Output, error and dL/dy.
		if (tid < kPass) {
			float y = P[kMlpOffB3];
			for (int o = 0; o < kMlpHidden; ++o)
				y += P[kMlpOffW3 + o] * H2[tid][o];
			const float err = (tid < valid) ? y - t[s0 + tid] : 0.0f;
			DY[tid]         = scale * err;
			SE[tid]         = err * err;
		}
		__syncthreads();
		// Delta of layer 2.
#pragma unroll 1
		for (int q = 0; q < Q; ++q) {
			const int e = tid + kT * q, sm = e / kMlpHidden, o = e % kMlpHidden;
			const float h = H2[sm][o];
			DA2[sm][o]    = DY[sm] * P[kMlpOffW3 + o] * (1.0f - h * h);
		}
		__syncthreads();
		// Delta of layer 1.
#pragma unroll 1
		for (int q = 0; q < Q; ++q) {
			const int e = tid + kT * q, sm = e / kMlpHidden, i = e % kMlpHidden;
			float     d = 0.0f;
#pragma unroll 8
			for (int o = 0; o < kMlpHidden; ++o)
				d += P[kMlpOffW2 + o * kMlpHidden + i] * DA2[sm][o];
			const float h = H1[sm][i];
			DA1[sm][i]    = d * (1.0f - h * h);
		}
		__syncthreads();
		// Fold the pass into the accumulators (samples in order).
#pragma unroll
		for (int q = 0; q < kMlpHidden * kMlpHidden / kT; ++q) {
			const int e = tid + kT * q;
			const int o = e / kMlpHidden, i = e % kMlpHidden;
			float     a = 0.0f;
#pragma unroll 4
			for (int sm = 0; sm < kPass; ++sm)
				a += DA2[sm][o] * H1[sm][i];
			acc2[q] += a;
		}
		{
			const int o = tid / kMlpIn, i = tid % kMlpIn; // tid < 256 = 64 * 4
			float     a = 0.0f;
			for (int sm = 0; sm < kPass; ++sm)
				a += DA1[sm][o] * X[sm][i];
			acc1 += a;
		}
		if (tid < 3 * kMlpHidden + 1) {
			float a = 0.0f;
			if (tid < kMlpHidden)
				for (int sm = 0; sm < kPass; ++sm)
					a += DA1[sm][tid];
			else if (tid < 2 * kMlpHidden)
				for (int sm = 0; sm < kPass; ++sm)
					a += DA2[sm][tid - kMlpHidden];
			else if (tid < 3 * kMlpHidden)
				for (int sm = 0; sm < kPass; ++sm)
					a += DY[sm] * H2[sm][tid - 2 * kMlpHidden];
			else
				for (int sm = 0; sm < kPass; ++sm)
					a += DY[sm];
			accv += a;
		}
		if (tid == 0)
			for (int sm = 0; sm < kPass; ++sm)
				accse += SE[sm];
	}
	float* out = partial + int64_t(blockIdx.x) * (kMlpParams + 1);
#pragma unroll
	for (int q = 0; q < kMlpHidden * kMlpHidden / kT; ++q)
		out[kMlpOffW2 + tid + kT * q] = acc2[q];
	out[kMlpOffW1 + tid] = acc1;
	if (tid < kMlpHidden)
		out[kMlpOffB1 + tid] = accv;
	else if (tid < 2 * kMlpHidden)
		out[kMlpOffB2 + tid - kMlpHidden] = accv;
	else if (tid < 3 * kMlpHidden)
		out[kMlpOffW3 + tid - 2 * kMlpHidden] = accv;
	else if (tid == 3 * kMlpHidden)
		out[kMlpOffB3] = accv;
	if (tid == 0)
		out[kMlpParams] = accse;
}

// grad[p] = sum over workgroups (in index order) of the partials, p <= kMlpParams.
__global__ __launch_bounds__(kT) void mlp_grad_reduce_kernel(const float* __restrict__ partial,
                                                             int nparts, float* __restrict__ grad)
{
	for (int p = blockIdx.x * kT + threadIdx.x; p <= kMlpParams; p += gridDim.x * kT) {
		float s = 0.0f;
		for (int w = 0; w < nparts; ++w)
			s += partial[int64_t(w) * (kMlpParams + 1) + p];
		grad[p] = s;
	}
}

// Adam step (bias-corrected, step number `step` >= 1).
__global__ __launch_bounds__(kT) void mlp_adam_kernel(float* __restrict__ params,
                                                      const float* __restrict__ grad,
                                                      float* __restrict__ m, float* __restrict__ v,
                                                      float lr, float b1, float b2, float eps,
                                                      int step)
{
	const float c1 = 1.0f - powf(b1, float(step));
	const float c2 = 1.0f - powf(b2, float(step));
	for (int p = blockIdx.x * kT + threadIdx.x; p < kMlpParams; p += gridDim.x * kT) {
		const float g = grad[p];
		const float mm = b1 * m[p] + (1.0f - b1) * g;
		const float vv = b2 * v[p] + (1.0f - b2) * g * g;
		m[p]           = mm;
		v[p]           = vv;
		params[p] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
	}
}

int mlp_grad_parts(int64_t count)
{
	int64_t passes = (count + kPass - 1) / kPass;
	return int(passes < 256 ? (passes < 1 ? 1 : passes) : 256);
}

void launch_mlp_forward(const float* params, const float* x, int64_t count, float* y, hipStream_t s)
{
	if (count <= 0)
		return;
	int64_t b = (count + kT - 1) / kT;
	if (b > 2048)
		b = 2048;
	hipLaunchKernelGGL(mlp_forward_kernel, dim3(unsigned(b)), dim3(kT), 0, s, params, x, count, y);
	DMK_HIP_CHECK(hipGetLastError());
}

void launch_mlp_grad(const float* params, const float* x, const float* t, int64_t count,
                     float* partial, float* grad, hipStream_t s)
{
	const int parts = mlp_grad_parts(count);
	hipLaunchKernelGGL(mlp_grad_kernel, dim3(unsigned(parts)), dim3(kT), 0, s, params, x, t, count,
	                   partial);
	DMK_HIP_CHECK(hipGetLastError());
	hipLaunchKernelGGL(mlp_grad_reduce_kernel, dim3(unsigned((kMlpParams + kT) / kT)), dim3(kT), 0,
	                   s, partial, parts, grad);
	DMK_HIP_CHECK(hipGetLastError());
}

void launch_mlp_adam(float* params, const float* grad, float* m, float* v, float lr, int step,
                     hipStream_t s)
{
	hipLaunchKernelGGL(mlp_adam_kernel, dim3(unsigned((kMlpParams + kT - 1) / kT)), dim3(kT), 0, s,
	                   params, grad, m, v, lr, 0.9f, 0.999f, 1e-8f, step);
	DMK_HIP_CHECK(hipGetLastError());
}

} // namespace dmk
