// The InitKron step on the device: packing of the Kron factors into the K-major panels (the
// packing step of a plan: setup_sparse_batch.cpp:281-343 copies the blocks column by column on
// the host) and, for blocks that already live in device memory, the nonzero / identity scan
// (setup_nC.cpp:47-124) and the gather of the diagonals.
//
// A pack job writes block rows as panel rows: element (i, j) of the block (column major at src,
// leading dimension ldsrc) goes to pack[dst + i * lddst + j], and the `tail` elements after each
// row's last column are zeroed (row padding of the panels). One workgroup moves a 64 x 64 tile of
// one job through LDS: coalesced reads down the block's columns, coalesced writes along the panel
// rows.
#include "dmk_internal.h"

#include <hip/hip_runtime.h>

namespace dmk {

namespace {

__global__ __launch_bounds__(256) void pack_panels_kernel(const PackJob* __restrict__ jobs,
                                                          const PackTile* __restrict__ tiles,
                                                          double* __restrict__ apack,
                                                          double* __restrict__ bpack)
{
	__shared__ double t[64][65];
	const PackTile pt  = tiles[blockIdx.x];
	const PackJob  jb  = jobs[pt.job];
	double*        dst = (jb.to_b ? bpack : apack) + jb.dst;
	const int      tx  = threadIdx.x & 63, ty = threadIdx.x >> 6;
	// Read the tile column by column (i fastest), all 16 loads of a thread in flight before the
	// first LDS store.
	const int ir = pt.i0 + tx;
	double    v[16];
#pragma unroll
	for (int q = 0; q < 16; ++q) {
		const int jj = pt.j0 + ty + 4 * q;
		v[q]         = (ir < jb.rows && jj < jb.cols) ? jb.src[ir + int64_t(jj) * jb.ldsrc] : 0.0;
	}
#pragma unroll
	for (int q = 0; q < 16; ++q)
		t[ty + 4 * q][tx] = v[q];
	__syncthreads();
	// Write panel rows (j fastest); the tile holding the last columns also zeroes the padding.
	const int  jend = min(jb.cols, pt.j0 + 64);
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

// One 64 x 64 tile of a device-resident block: reads every element of the tile (coalesced down
// the columns) and ors the tile's findings into the block's two flags. The comparisons are those
// of the host scan: v != 0 for a nonzero (NaN counts), v != (i == j ? 1 : 0) for a departure
// from the identity.
__global__ __launch_bounds__(256) void scan_blocks_kernel(const ScanBlock* __restrict__ blocks,
                                                          const PackTile* __restrict__ tiles,
                                                          int* __restrict__ flags)
{
	const PackTile  pt = tiles[blockIdx.x];
	const ScanBlock b  = blocks[pt.job];
	const int       tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
	const int       i  = pt.i0 + tx;
	int             nz = 0, dep = 0;
	if (i < b.rows)
		for (int j = ty; j < 64; j += 4) {
			const int jj = pt.j0 + j;
			if (jj >= b.cols)
				break;
			const double v = b.p[i + int64_t(jj) * b.ld];
			nz |= v != 0.0;
			dep |= v != (i == jj ? 1.0 : 0.0);
		}
	nz  = __syncthreads_or(nz);
	dep = __syncthreads_or(dep);
	if (threadIdx.x == 0) {
		if (nz)
			atomicOr(flags + 2 * pt.job, 1);
		if (dep)
			atomicOr(flags + 2 * pt.job + 1, 1);
	}
}

__global__ __launch_bounds__(256) void gather_diag_kernel(const DiagJob* __restrict__ jobs,
                                                          double* __restrict__ out)
{
	const DiagJob jb = jobs[blockIdx.x];
	for (int64_t t = threadIdx.x; t < jb.count; t += blockDim.x)
		out[jb.off + t] = jb.p[t * (jb.ld + 1)];
}

} // namespace

void launch_pack_panels(const PackJob* jobs, const PackTile* tiles, int ntiles, double* apack,
                        double* bpack, hipStream_t stream)
{
	if (ntiles <= 0)
		return;
	hipLaunchKernelGGL(pack_panels_kernel, dim3(unsigned(ntiles)), dim3(256), 0, stream, jobs,
	                   tiles, apack, bpack);
}

void launch_scan_blocks(const ScanBlock* blocks, const PackTile* tiles, int ntiles, int* flags,
                        hipStream_t stream)
{
	if (ntiles <= 0)
		return;
	hipLaunchKernelGGL(scan_blocks_kernel, dim3(unsigned(ntiles)), dim3(256), 0, stream, blocks,
	                   tiles, flags);
}

void launch_gather_diag(const DiagJob* jobs, int njobs, double* out, hipStream_t stream)
{
	if (njobs <= 0)
		return;
	hipLaunchKernelGGL(gather_diag_kernel, dim3(unsigned(njobs)), dim3(256), 0, stream, jobs,
	                   out);
}

} // namespace dmk
