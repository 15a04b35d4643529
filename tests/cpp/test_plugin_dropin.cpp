// The call pattern of the reference adapter BatchedGemmPluginSc (BatchedGemmPluginSc.h:50-161)
// compiled against the drop-in header "BatchedGemm.h": the global template BatchedGemm<T>, the
// constructor from the adapter's vectors (:119-120) and apply_Htarget on a pair of views (:159).
// SharedView stands in for Kokkos::View<T*, Kokkos::SharedSpace> (the members the call uses:
// data() and extent(0)); its memory is hipMallocManaged, which is what Kokkos::SharedSpace
// allocates on ROCm. The same call runs on device and on host memory, and a view shorter than the
// superblock vector is rejected. Fixtures: test_BatchedGemm.cpp:171-258 (real) and
// test_KronConnections.cpp:132-178 (complex). Needs a GPU. Exit code 0 on success.
#include "BatchedGemm.h"

#include <hip/hip_runtime_api.h>

#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

template <typename T> struct SharedView {
	T*          ptr = nullptr;
	std::size_t len = 0;
	T*          data() const { return ptr; }
	std::size_t extent(int) const { return len; }
};

namespace {

int failures = 0;

template <typename T> T* alloc(const char* kind, std::size_t n)
{
	void* p = nullptr;
	if (!std::strcmp(kind, "managed")) {
		if (hipMallocManaged(&p, n * sizeof(T)) != hipSuccess)
			return nullptr;
	} else if (!std::strcmp(kind, "device")) {
		if (hipMalloc(&p, n * sizeof(T)) != hipSuccess)
			return nullptr;
	} else {
		p = new T[n];
	}
	return static_cast<T*>(p);
}

template <typename T> void release(const char* kind, T* p)
{
	if (!std::strcmp(kind, "host"))
		delete[] p;
	else
		(void)hipFree(p);
}

// The adapter's matrixVector body with the view pair in memory of the given kind.
template <typename T>
void run(BatchedGemm<T>& bg, const char* kind, const std::vector<T>& vin, std::vector<T>& vout)
{
	const std::size_t n  = vin.size();
	T*                di = alloc<T>(kind, n);
	T*                dout = alloc<T>(kind, n);
	if (!di || !dout) {
		std::printf("%s: allocation failed\n", kind);
		++failures;
		return;
	}
	// hipMemset is for device-accessible memory only: on a plain host array it fails and leaves
	// a sticky error for the next hipGetLastError of the process.
	(void)hipMemcpy(di, vin.data(), n * sizeof(T), hipMemcpyDefault);
	if (!std::strcmp(kind, "host"))
		std::memset(static_cast<void*>(dout), 0, n * sizeof(T));
	else if (hipMemset(dout, 0, n * sizeof(T)) != hipSuccess) {
		std::printf("%s: hipMemset failed\n", kind);
		++failures;
	}
	SharedView<const T> vin_view { di, n };
	SharedView<T>       vout_view { dout, n };
	bg.apply_Htarget(vin_view, vout_view);
	(void)hipMemcpy(vout.data(), dout, n * sizeof(T), hipMemcpyDefault);
	release(kind, di);
	release(kind, dout);
}

} // namespace

int main()
{
	const char* kinds[] = { "managed", "device", "host" };
	try {
		std::mt19937_64                        rng(123456789);
		std::uniform_real_distribution<double> dist(-1.0, 1.0);
		std::vector<double>                    A(4), B(9), vin(6);
		for (int i = 0; i < 2; i++)
			for (int j = 0; j < 2; j++)
				A[i + 2 * j] = dist(rng);
		for (int i = 0; i < 3; i++)
			for (int j = 0; j < 3; j++)
				B[i + 3 * j] = dist(rng);
		for (int i = 0; i < 6; i++)
			vin[i] = dist(rng);
		std::vector<double> expected(6);
		for (int i = 0; i < 3; i++)
			for (int j = 0; j < 2; j++) {
				double s = 0;
				for (int k = 0; k < 2; k++) {
					double bx = 0;
					for (int l = 0; l < 3; l++)
						bx += B[i + 3 * l] * vin[l + 3 * k];
					s += bx * A[j + 2 * k];
				}
				expected[i + 3 * j] = s;
			}
		std::vector<std::size_t> pLeft { 2 }, pRight { 3 }, ldA { 2 }, ldB { 3 };
		std::vector<double*>     aptr { A.data() }, bptr { B.data() };
		BatchedGemm<double>*     bg = new BatchedGemm<double>(1, 1, pLeft, pRight, aptr, ldA, bptr, ldB);
		for (const char* kind : kinds) {
			std::vector<double> vout(6, 0.0);
			run(*bg, kind, vin, vout);
			for (int i = 0; i < 6; i++)
				if (!(std::fabs(vout[i] - expected[i]) <= 1e-12 * std::fabs(expected[i]))) {
					std::printf("%s: mismatch at %d: %.17g vs %.17g\n", kind, i, vout[i], expected[i]);
					++failures;
				}
		}
		// A view shorter than the vector is an error, reported before any device work.
		std::vector<double> shortv(5, 0.0), outv(6, 0.0);
		bool                thrown = false;
		try {
			bg->apply_Htarget(SharedView<const double> { shortv.data(), shortv.size() },
			                  SharedView<double> { outv.data(), outv.size() });
		} catch (const std::runtime_error&) {
			thrown = true;
		}
		if (!thrown) {
			std::printf("short view accepted\n");
			++failures;
		}
		delete bg;

		using cplx = std::complex<double>;
		std::vector<cplx>        a { cplx(1, 1), cplx(1, -2), cplx(2, 1), cplx(-1, 1) };
		std::vector<cplx>        b { cplx(2, -1), cplx(3, 1), cplx(1, 2), cplx(2, 2) };
		std::vector<cplx*>       ap { &a[0], &a[1], &a[2], &a[3] }, bp { &b[0], &b[1], &b[2], &b[3] };
		std::vector<std::size_t> one2 { 1, 1 }, one4 { 1, 1, 1, 1 };
		BatchedGemm<cplx>        bz(1, 2, one2, one2, ap, one4, bp, one4);
		const cplx               want[2] = { cplx(7, 14), cplx(2, 4) };
		for (const char* kind : kinds) {
			std::vector<cplx> zout(2);
			run(bz, kind, std::vector<cplx> { cplx(1, 1), cplx(2, -1) }, zout);
			for (int i = 0; i < 2; i++)
				if (std::abs(zout[i] - want[i]) > 1e-13) {
					std::printf("%s complex: mismatch at %d\n", kind, i);
					++failures;
				}
		}
	} catch (const std::exception& e) {
		std::printf("exception: %s\n", e.what());
		return 2;
	}
	std::printf("test_plugin_dropin: %s\n", failures ? "FAILED" : "passed");
	return failures ? 1 : 0;
}
