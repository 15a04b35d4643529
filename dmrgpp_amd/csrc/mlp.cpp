#include "mlp.h"

#include "dmk_internal.h"
#include "pool.h"

#include <cmath>
#include <random>
#include <stdexcept>

namespace dmk {

Mlp::Mlp(uint64_t seed, int64_t sample_capacity)
{
	if (sample_capacity < 1)
		throw std::invalid_argument("mlp: sample capacity must be positive");
	DMK_HIP_CHECK(hipStreamCreate(&stream_));
	auto dev = [](size_t n) { return static_cast<float*>(pool_device_alloc(n * sizeof(float))); };
	d_params_ = dev(kMlpParams);
	d_m_      = dev(kMlpParams);
	d_v_      = dev(kMlpParams);
	d_grad_   = dev(kMlpParams + 1);
	d_part_   = dev(size_t(256) * (kMlpParams + 1));
	cap_      = sample_capacity;
	d_x_      = dev(size_t(cap_) * kMlpIn);
	d_t_      = dev(size_t(cap_));
	// Glorot-style uniform initialization of the hidden layers, zero biases, from a fixed-seed
	// generator.
	std::vector<float>                    p(kMlpParams, 0.0f);
	std::mt19937_64                       rng(seed);
	auto fill = [&](int off, int fan_in, int fan_out) {
		const float                           r = std::sqrt(6.0f / float(fan_in + fan_out));
		std::uniform_real_distribution<float> u(-r, r);
		for (int i = 0; i < fan_in * fan_out; ++i)
			p[size_t(off + i)] = u(rng);
	};
	fill(kMlpOffW1, kMlpIn, kMlpHidden);
	fill(kMlpOffW2, kMlpHidden, kMlpHidden);
	// W3 and b3 start at zero: an untrained network predicts no correction.
	set_params(p.data());
	std::vector<float> zero(kMlpParams, 0.0f);
	DMK_HIP_CHECK(hipMemcpy(d_m_, zero.data(), kMlpParams * sizeof(float), hipMemcpyHostToDevice));
	DMK_HIP_CHECK(hipMemcpy(d_v_, zero.data(), kMlpParams * sizeof(float), hipMemcpyHostToDevice));
}

Mlp::~Mlp()
{
	if (stream_)
		(void)hipStreamSynchronize(stream_);
	for (float* p : { d_params_, d_m_, d_v_, d_grad_, d_part_, d_x_, d_t_ })
		pool_device_free(p);
	if (stream_)
		(void)hipStreamDestroy(stream_);
}

void Mlp::get_params(float* host) const
{
	DMK_HIP_CHECK(hipMemcpyAsync(host, d_params_, kMlpParams * sizeof(float), hipMemcpyDeviceToHost,
	                             stream_));
	DMK_HIP_CHECK(hipStreamSynchronize(stream_));
}

void Mlp::set_params(const float* host)
{
	DMK_HIP_CHECK(hipMemcpyAsync(d_params_, host, kMlpParams * sizeof(float), hipMemcpyHostToDevice,
	                             stream_));
	DMK_HIP_CHECK(hipStreamSynchronize(stream_));
	step_ = 0;
}

double Mlp::gradient(const float* x, const float* t, int64_t count, float* grad_host)
{
	if (count < 1 || !x || !t)
		throw std::invalid_argument("mlp_gradient: empty batch");
	launch_mlp_grad(d_params_, x, t, count, d_part_, d_grad_, stream_);
	std::vector<float> g(kMlpParams + 1);
	DMK_HIP_CHECK(hipMemcpyAsync(g.data(), d_grad_, (kMlpParams + 1) * sizeof(float),
	                             hipMemcpyDeviceToHost, stream_));
	DMK_HIP_CHECK(hipStreamSynchronize(stream_));
	if (grad_host)
		for (int i = 0; i < kMlpParams; ++i)
			grad_host[i] = g[size_t(i)];
	return double(g[kMlpParams]) / double(count);
}

int64_t Mlp::append(int64_t more)
{
	if (more < 0)
		throw std::invalid_argument("mlp: negative sample count");
	if (count_ + more > cap_)
		more = cap_ - count_;
	const int64_t first = count_;
	count_ += more;
	return first;
}

void Mlp::train(int iters, int64_t batch, float lr, double* loss_last, double* loss_all)
{
	if (count_ < 1)
		throw std::runtime_error("mlp_train: no training samples");
	if (batch < 1 || batch > count_)
		batch = count_;
	int64_t start = 0;
	for (int it = 0; it < iters; ++it) {
		if (start + batch > count_)
			start = 0;
		launch_mlp_grad(d_params_, d_x_ + start * kMlpIn, d_t_ + start, batch, d_part_, d_grad_,
		                stream_);
		launch_mlp_adam(d_params_, d_grad_, d_m_, d_v_, lr, ++step_, stream_);
		start += batch;
	}
	float l = 0;
	if (loss_last) {
		if (iters > 0) {
			DMK_HIP_CHECK(hipMemcpyAsync(&l, d_grad_ + kMlpParams, sizeof(float), hipMemcpyDeviceToHost,
			                             stream_));
			DMK_HIP_CHECK(hipStreamSynchronize(stream_));
		}
		*loss_last = double(l) / double(batch);
	}
	if (loss_all)
		*loss_all = gradient(d_x_, d_t_, count_, nullptr);
	DMK_HIP_CHECK(hipStreamSynchronize(stream_));
}

} // namespace dmk
