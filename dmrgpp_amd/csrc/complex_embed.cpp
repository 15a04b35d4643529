#include "complex_embed.h"

#include <stdexcept>

namespace dmk {

RealEmbedding embed_complex(int64_t nC, int64_t np, const size_t* L, const size_t* R,
                            const double* const* Ac, const size_t* ldA, const double* const* Bc,
                            const size_t* ldB)
{
	if (np <= 0 || nC <= 0)
		throw std::runtime_error("dmk_create_z: npatches and noperator must be positive");
	if (!L || !R || !Ac || !Bc || !ldA || !ldB)
		throw std::runtime_error("dmk_create_z: null argument array");
	RealEmbedding e;
	e.np = 2 * np;
	e.nC = 2 * nC;
	e.L.resize(size_t(e.np));
	e.R.resize(size_t(e.np));
	for (int64_t I = 0; I < np; ++I) {
		e.L[size_t(I)] = e.L[size_t(I + np)] = L[I];
		e.R[size_t(I)] = e.R[size_t(I + np)] = R[I];
	}
	const size_t total = size_t(e.np * e.np * e.nC);
	e.A.assign(total, nullptr);
	e.B.assign(total, nullptr);
	e.ldA.assign(total, 1);
	e.ldB.assign(total, 1);
	for (size_t i = 0; i < total; ++i) {
		const size_t I = i % size_t(e.np);
		e.ldA[i]       = e.L[I];
		e.ldB[i]       = e.R[I];
	}
	auto ridx = [&](int64_t I, int64_t J, int64_t k) {
		return size_t(I + J * e.np + k * e.np * e.np);
	};
	// Splits a complex block into (re, im) and (-re, -im) column-major real blocks.
	auto split = [&](const double* p, size_t rows, size_t cols, size_t ld, double sign) {
		std::vector<double> re(rows * cols), im(rows * cols);
		for (size_t c = 0; c < cols; ++c)
			for (size_t r = 0; r < rows; ++r) {
				re[r + c * rows] = sign * p[2 * (r + c * ld)];
				im[r + c * rows] = sign * p[2 * (r + c * ld) + 1];
			}
		e.storage.push_back(std::move(re));
		e.storage.push_back(std::move(im));
		return std::make_pair(e.storage[e.storage.size() - 2].data(),
		                      e.storage[e.storage.size() - 1].data());
	};
	e.storage.reserve(size_t(6) * size_t(np * np * nC));
	for (int64_t k = 0; k < nC; ++k)
		for (int64_t J = 0; J < np; ++J)
			for (int64_t I = 0; I < np; ++I) {
				const size_t id = size_t(I + J * np + k * np * np);
				if (!Ac[id] || !Bc[id])
					continue;
				if (ldA[id] < L[I] || ldB[id] < R[I])
					throw std::runtime_error("dmk_create_z: leading dimension smaller than rows");
				auto a  = split(Ac[id], L[I], L[J], ldA[id], 1.0);
				auto am = split(Ac[id], L[I], L[J], ldA[id], -1.0);
				auto b  = split(Bc[id], R[I], R[J], ldB[id], 1.0);
				const double* Ar = a.first;
				const double* Ai = a.second;
				const double* nAr = am.first;
				const double* nAi = am.second;
				const double* Br = b.first;
				const double* Bi = b.second;
				auto put = [&](int64_t I2, int64_t J2, int64_t k2, const double* Ap, const double* Bp) {
					e.A[ridx(I2, J2, k2)] = Ap;
					e.B[ridx(I2, J2, k2)] = Bp;
				};
				// Re <- Re: Ar (x) Br, (-Ai) (x) Bi
				put(I, J, 2 * k, Ar, Br);
				put(I, J, 2 * k + 1, nAi, Bi);
				// Re <- Im: (-Ar) (x) Bi, (-Ai) (x) Br
				put(I, J + np, 2 * k, nAr, Bi);
				put(I, J + np, 2 * k + 1, nAi, Br);
				// Im <- Re: Ar (x) Bi, Ai (x) Br
				put(I + np, J, 2 * k, Ar, Bi);
				put(I + np, J, 2 * k + 1, Ai, Br);
				// Im <- Im: Ar (x) Br, (-Ai) (x) Bi
				put(I + np, J + np, 2 * k, Ar, Br);
				put(I + np, J + np, 2 * k + 1, nAi, Bi);
			}
	return e;
}

} // namespace dmk
