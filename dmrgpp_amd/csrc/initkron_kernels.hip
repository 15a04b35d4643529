// Packing of the Kron factors into the K-major panels on the device (the InitKron / packing step
// of a plan: setup_sparse_batch.cpp:281-343 copies the blocks column by column on the host).
//
// The planner uploads every distinct non-identity block once, compact and column major (rows x
// cols at raw_off), and a list of copy jobs. A job writes block rows as panel rows: element
// (i, j) of the block goes to pack[dst + i * lddst + j], and the `tail` elements after each row's
// last column are zeroed (row padding of the panels). One workgroup moves a 64 x 64 tile of one
// job through LDS: coalesced reads down the block's columns, coalesced writes along the panel
// rows.
#include "dmk_internal.h"

#include <hip/hip_runtime.h>

namespace dmk {

namespace {

__global__ __launch_bounds__(256) void pack_panels_kernel(const PackJob* __restrict__ jobs,
                                                          const PackTile* __restrict__ tiles,
                                                          const double* __restrict__ raw,
                                                          double* __restrict__ apack,
                                                          double* __restrict__ bpack)
{
	__shared__ double t[64][65];
	const PackTile pt = tiles[blockIdx.x];
	const PackJob  jb = jobs[pt.job];
	const double*  src = raw + jb.raw_off;
	double*        dst = (jb.to_b ? bpack : apack) + jb.dst;
	const int      tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
	// Read the tile column by column (i fastest).
	for (int j = ty; j < 64; j += 4) {
		const int i  = pt.i0 + tx;
		const int jj = pt.j0 + j;
		t[j][tx]     = (i < jb.rows && jj < jb.cols) ? src[i + int64_t(jj) * jb.rows] : 0.0;
	}
	__syncthreads();
	// Write panel rows (j fastest); the tile holding the last columns also zeroes the padding.
	const int jend = min(jb.cols, pt.j0 + 64);
	const bool last = pt.j0 + 64 >= jb.cols;
	for (int r = ty; r < 64; r += 4) {
		const int i = pt.i0 + r;
		if (i >= jb.rows)
			break;
		double*   row = dst + int64_t(i) * jb.lddst;
		const int jj  = pt.j0 + tx;
		if (jj < jend)
			row[jj] = t[tx][r];
		if (last && tx < jb.tail)
			row[jb.cols + tx] = 0.0;
	}
}

} // namespace

void launch_pack_panels(const PackJob* jobs, const PackTile* tiles, int ntiles, const double* raw,
                        double* apack, double* bpack, hipStream_t stream)
{
	if (ntiles <= 0)
		return;
	hipLaunchKernelGGL(pack_panels_kernel, dim3(unsigned(ntiles)), dim3(256), 0, stream, jobs,
	                   tiles, raw, apack, bpack);
}

} // namespace dmk
