// Header-only C++ class with the signature of DMRG++'s plugin class
//   template <typename T> class BatchedGemm          (dmrg/GPUPlugin/BatchedGemm.h:11-56)
// implemented over the dmrgpp_amd C ABI (dmrgpp_amd.h). A DMRG++ build selecting the plugin
// (-DDMRG_BATCHED_GEMM_TYPE=Plugin) can use it in place of dmrg/GPUPlugin/BatchedGemm.{h,cpp}:
// the constructor takes the same vectors, apply_Htarget computes vout = H vin (beta = 0).
//
// Differences from the reference class, by design:
//  * apply_Htarget takes plain pointers (device pointers, or host pointers with
//    apply_Htarget_host) instead of Kokkos::View<..., SharedSpace>; a Kokkos caller passes
//    view.data();
//  * failures throw std::runtime_error carrying dmk_last_error() (the reference asserts);
//  * T = std::complex<double> runs through dmk_create_z / dmk_apply_z (the real embedding of the
//    complex blocks on the device).
#ifndef DMRGPP_AMD_BATCHED_GEMM_HPP
#define DMRGPP_AMD_BATCHED_GEMM_HPP

#include "dmrgpp_amd.h"

#include <complex>
#include <cstddef>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

namespace dmrgpp_amd {

template <typename T> class BatchedGemm {
public:
	using SizeType       = std::size_t;
	using VectorSizeType = std::vector<SizeType>;

	// Same arguments as BatchedGemm<T>::BatchedGemm (BatchedGemm.h:21-28). Amatrix/Bmatrix are
	// indexed outPatch + inPatch*npatches + k*npatches*npatches; nullptr means a zero block.
	BatchedGemm(SizeType noperator, SizeType npatches, VectorSizeType& left_patch_size,
	            VectorSizeType& right_patch_size, std::vector<T*>& Amatrix,
	            VectorSizeType& ld_Amatrix, std::vector<T*>& Bmatrix, VectorSizeType& ld_Bmatrix)
	{
		static_assert(std::is_same<T, double>::value || std::is_same<T, std::complex<double>>::value,
		              "BatchedGemm<T>: T must be double or std::complex<double>");
		const SizeType total = npatches * npatches * noperator;
		if (left_patch_size.size() < npatches || right_patch_size.size() < npatches
		    || Amatrix.size() < total || Bmatrix.size() < total || ld_Amatrix.size() < total
		    || ld_Bmatrix.size() < total)
			throw std::runtime_error("dmrgpp_amd::BatchedGemm: argument vectors too short");
		auto create = kComplex ? dmk_create_z : dmk_create_d;
		check(create(noperator, npatches, left_patch_size.data(), right_patch_size.data(),
		             reinterpret_cast<const double* const*>(Amatrix.data()), ld_Amatrix.data(),
		             reinterpret_cast<const double* const*>(Bmatrix.data()), ld_Bmatrix.data(),
		             &h_),
		      kComplex ? "dmk_create_z" : "dmk_create_d");
		dmk_info info;
		check(dmk_get_info(h_, &info), "dmk_get_info");
		n_ = size_t(info.n_complex);
	}

	~BatchedGemm() { dmk_destroy(h_); }

	BatchedGemm(const BatchedGemm&)            = delete;
	BatchedGemm& operator=(const BatchedGemm&) = delete;

	// vout = H vin on device vectors (BatchedGemm.h:30-31); waits for completion.
	void apply_Htarget(const T* vin, T* vout)
	{
		auto apply = kComplex ? dmk_apply_z : dmk_apply_d;
		check(apply(h_, reinterpret_cast<const double*>(vin), reinterpret_cast<double*>(vout)),
		      kComplex ? "dmk_apply_z" : "dmk_apply_d");
		check(dmk_sync(h_), "dmk_sync");
	}

	// Same on host vectors (copies in and out).
	void apply_Htarget_host(const T* vin, T* vout)
	{
		if (!kComplex) {
			check(dmk_apply_host_d(h_, reinterpret_cast<const double*>(vin),
			                       reinterpret_cast<double*>(vout)),
			      "dmk_apply_host_d");
			return;
		}
		const size_t bytes = n_ * sizeof(T);
		void*        dx    = nullptr;
		void*        dy    = nullptr;
		check(dmk_device_alloc(bytes, &dx), "dmk_device_alloc");
		int st = dmk_device_alloc(bytes, &dy);
		if (st == DMK_OK)
			st = dmk_memcpy_h2d(dx, vin, bytes);
		if (st == DMK_OK)
			st = dmk_apply_z(h_, static_cast<const double*>(dx), static_cast<double*>(dy));
		if (st == DMK_OK)
			st = dmk_sync(h_);
		if (st == DMK_OK)
			st = dmk_memcpy_d2h(vout, dy, bytes);
		const std::string msg = (st == DMK_OK) ? std::string() : dmk_last_error();
		dmk_device_free(dx);
		if (dy)
			dmk_device_free(dy);
		if (st != DMK_OK)
			throw std::runtime_error("apply_Htarget_host: " + msg);
	}

	void set_precision(int precision) { check(dmk_set_precision(h_, precision), "dmk_set_precision"); }

	dmk_handle* handle() const { return h_; }

private:
	static void check(int status, const char* what)
	{
		if (status != DMK_OK)
			throw std::runtime_error(std::string(what) + ": " + dmk_last_error());
	}

	static constexpr bool kComplex = std::is_same<T, std::complex<double>>::value;
	dmk_handle*           h_       = nullptr;
	size_t                n_       = 0; // vector length in elements of T
};

} // namespace dmrgpp_amd

#endif // DMRGPP_AMD_BATCHED_GEMM_HPP
