// Real embedding of a complex Kron problem.
//
// For complex blocks A = Ar + i Ai, B = Br + i Bi and x = xr + i xi,
//   Re(B X A^T) = Br Xr Ar^T + Bi Xr (-Ai)^T + Bi Xi (-Ar)^T + Br Xi (-Ai)^T
//   Im(B X A^T) = Br Xr Ai^T + Bi Xr Ar^T + Br Xi Ar^T + Bi Xi (-Ai)^T
// so the complex problem on n elements equals a real problem on 2n elements whose patches are
// the original ones twice (real parts first, then imaginary parts) and whose terms are the
// eight real products above, two per (patch half, patch half) pair (connection indices 2k and
// 2k+1). Blocks that are identically zero
// (the imaginary parts of real factors) drop out in the plan's nonzero scan, so a real
// Hamiltonian stored as complex costs two real matvecs.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace dmk {

struct RealEmbedding {
	int64_t              nC = 0, np = 0; // of the real problem
	std::vector<size_t>  L, R;           // 2 * np_complex patches
	std::vector<const double*> A, B;     // np^2 * nC pointers (into storage)
	std::vector<size_t>  ldA, ldB;
	std::vector<std::vector<double>> storage;
};

// Blocks are interleaved complex (re, im) column-major arrays with leading dimensions in complex
// elements, as std::complex<double> matrices are stored.
RealEmbedding embed_complex(int64_t noperator, int64_t npatches, const size_t* L, const size_t* R,
                            const double* const* Acplx, const size_t* ldA,
                            const double* const* Bcplx, const size_t* ldB);

} // namespace dmk
