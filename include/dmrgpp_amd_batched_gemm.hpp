// Header-only C++ class with the signature of DMRG++'s plugin class
//   template <typename T> class BatchedGemm          (dmrg/GPUPlugin/BatchedGemm.h:11-56)
// implemented over the dmrgpp_amd C ABI (dmrgpp_amd.h). A DMRG++ build selecting the plugin
// (-DDMRG_BATCHED_GEMM_TYPE=Plugin) can use it in place of dmrg/GPUPlugin/BatchedGemm.{h,cpp}:
// the constructor takes the same vectors, apply_Htarget computes vout = H vin (beta = 0).
//
// apply_Htarget takes the reference's Kokkos::View<const KokkosScalar*, Kokkos::SharedSpace> /
// Kokkos::View<KokkosScalar*, Kokkos::SharedSpace> pair unchanged (any view type with data() and
// extent(0): managed, device or host memory, see dmk_apply_shared), so
// BatchedGemmPluginSc::matrixVector (BatchedGemmPluginSc.h:139-161) compiles as it is; the header
// BatchedGemm.h beside this one gives the class the reference's global name.
//
// Differences from the reference class, by design:
//  * two pointer overloads as well: apply_Htarget(const T*, T*) on device vectors and
//    apply_Htarget_host on host vectors;
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
#include <utility>
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

	// The reference signature (BatchedGemm.h:30-31): vout = H vin on the views' memory (a
	// Kokkos::View in SharedSpace, or any view type with data() and extent(0) whose element type
	// has the layout of T, such as Kokkos::complex<double> for std::complex<double>); returns
	// when vout is written. Views shorter than the superblock vector are rejected.
	template <class VI, class VO, class = decltype(std::declval<const VI&>().data()),
	          class = decltype(std::declval<const VO&>().data()),
	          class = decltype(std::declval<const VI&>().extent(0)),
	          class = decltype(std::declval<const VO&>().extent(0))>
	void apply_Htarget(const VI& vin, const VO& vout)
	{
		static_assert(sizeof(*vin.data()) == sizeof(T) && sizeof(*vout.data()) == sizeof(T),
		              "apply_Htarget: view element type must have the size of T");
		if (size_t(vin.extent(0)) < n_ || size_t(vout.extent(0)) < n_)
			throw std::runtime_error("dmrgpp_amd::BatchedGemm::apply_Htarget: view of "
			                         + std::to_string(size_t(vin.extent(0))) + " / "
			                         + std::to_string(size_t(vout.extent(0)))
			                         + " elements for a superblock vector of "
			                         + std::to_string(n_));
		check(dmk_apply_shared(h_, static_cast<const void*>(vin.data()),
		                       static_cast<void*>(vout.data())),
		      "dmk_apply_shared");
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
	std::size_t size() const { return n_; } // superblock vector length in elements of T

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
