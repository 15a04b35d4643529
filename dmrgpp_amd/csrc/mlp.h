// Per-element MLP preconditioner of the Davidson solver (4 -> 64 -> 64 -> 1, tanh, FP32):
// parameters, training set, and the device training loop.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace dmk {

constexpr int kMlpIn     = 4;
constexpr int kMlpHidden = 64;
constexpr int kMlpOffW1  = 0;
constexpr int kMlpOffB1  = kMlpOffW1 + kMlpHidden * kMlpIn;
constexpr int kMlpOffW2  = kMlpOffB1 + kMlpHidden;
constexpr int kMlpOffB2  = kMlpOffW2 + kMlpHidden * kMlpHidden;
constexpr int kMlpOffW3  = kMlpOffB2 + kMlpHidden;
constexpr int kMlpOffB3  = kMlpOffW3 + kMlpHidden;
constexpr int kMlpParams = kMlpOffB3 + 1; // 4545

void launch_mlp_forward(const float* params, const float* x, int64_t count, float* y, hipStream_t s);
void launch_mlp_grad(const float* params, const float* x, const float* t, int64_t count,
                     float* partial, float* grad, hipStream_t s);
void launch_mlp_adam(float* params, const float* grad, float* m, float* v, float lr, int step,
                     hipStream_t s);
int  mlp_grad_parts(int64_t count);

class Mlp {
public:
	explicit Mlp(uint64_t seed, int64_t sample_capacity);
	~Mlp();
	Mlp(const Mlp&)            = delete;
	Mlp& operator=(const Mlp&) = delete;

	hipStream_t  stream() const { return stream_; }
	const float* params() const { return d_params_; }
	void         get_params(float* host) const;
	void         set_params(const float* host);

	// Gradient (host, kMlpParams) and mean squared error of the device batch (x, t).
	double gradient(const float* x, const float* t, int64_t count, float* grad_host);

	// Training set (device): count samples of kMlpIn features and one target.
	float*  samples_x() { return d_x_; }
	float*  samples_t() { return d_t_; }
	int64_t samples() const { return count_; }
	int64_t capacity() const { return cap_; }
	// Reserves room for `more` samples at the end of the set and returns the first index; the
	// caller fills them on the stream.
	int64_t append(int64_t more);
	void    clear_samples() { count_ = 0; }

	// Adam over minibatches of `batch` consecutive samples (cycling through the set);
	// returns the mean squared error of the last minibatch and of the whole set at the end.
	void train(int iters, int64_t batch, float lr, double* loss_last, double* loss_all);

private:
	hipStream_t stream_   = nullptr;
	float*      d_params_ = nullptr;
	float*      d_m_      = nullptr;
	float*      d_v_      = nullptr;
	float*      d_grad_   = nullptr; // kMlpParams + 1 (loss)
	float*      d_part_   = nullptr; // 256 x (kMlpParams + 1)
	float*      d_x_      = nullptr;
	float*      d_t_      = nullptr;
	int64_t     cap_      = 0;
	int64_t     count_    = 0;
	int         step_     = 0;
};

} // namespace dmk
