// Grouped multi-segment GEMM on operands in the caller's own layouts (column major, the right
// operand optionally transposed), for the variable-size batched GEMM (vbatch.cpp, the
// dmrg_Xgemm_vbatch replacement); the split-k reduction and the FP64 -> FP32 cast shared with
// the matvec (panel_kernels.hip holds the matvec's own GEMM kernels).
//
// One launch computes a list of independent output matrices ("tasks"), each the sum of one or
// more products ("segments") C += L * op(R).
//
// Tiling: 256 threads = 4 wave64s in a 2x2 grid; a workgroup owns a BM x BN output tile (64 or
// 128 each way) and each wave a (BM/2) x (BN/2) sub-tile held as MFMA 16x16 accumulators. K is
// walked in chunks of 16 staged through LDS, double buffered: the global loads of chunk c+1 are
// issued before the MFMAs of chunk c. Every output tile is owned by exactly one workgroup and
// its contraction order is fixed, so results are bitwise reproducible (no atomics).
#include "dmk_internal.h"

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <type_traits>

namespace dmk {

constexpr int kBK      = 16;
constexpr int kThreads = 256; // 4 wave64s in a 2x2 grid
constexpr int kPad     = 8;   // LDS row padding (elements)

template <typename T> struct MfmaOp;

// Operand layouts of the 16x16x4 MFMAs (both types): lane l supplies A[l%16][l/16] and
// B[l/16][l%16]. Result layouts differ (checked on gfx950): f64 holds D[l/16 + 4r][l%16] in
// register r, f32 holds D[4(l/16) + r][l%16].
// kM x kM output block and kK-deep contraction per instruction: lane l supplies operand
// A[l % kM][l / kM] and B[l / kM][l % kM]; register r of the result holds row row(l, r),
// column l % kM.
template <> struct MfmaOp<double> {
	static constexpr int kM    = 16;
	static constexpr int kK    = 4;
	static constexpr int kRegs = 4;
	using acc_t = double __attribute__((ext_vector_type(4)));
	__device__ static inline acc_t mma(double a, double b, acc_t c)
	{
		return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
	}
	__device__ static inline int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};

template <> struct MfmaOp<float> {
	static constexpr int kM    = 16;
	static constexpr int kK    = 4;
	static constexpr int kRegs = 4;
	using acc_t = float __attribute__((ext_vector_type(4)));
	__device__ static inline acc_t mma(float a, float b, acc_t c)
	{
		return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
	}
	__device__ static inline int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

// FP32 with 32x32x2 blocks: twice the flops per LDS operand read of the 16x16x4 form.
// Result register r of lane l holds row 8(r/4) + 4(l/32) + r%4 (checked on gfx950).
struct MfmaF32x32 {
	static constexpr int kM    = 32;
	static constexpr int kK    = 2;
	static constexpr int kRegs = 16;
	using acc_t = float __attribute__((ext_vector_type(16)));
	__device__ static inline acc_t mma(float a, float b, acc_t c)
	{
		return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
	}
	__device__ static inline int row(int lane, int r)
	{
		return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
	}
};

template <typename T> struct Bases {
	const T* p[BASE_COUNT];
};

// Per-thread register staging for one K chunk of a BM x BN tile.
template <typename T, int BM, int BN> struct Stage {
	static constexpr int NL = BM * kBK / kThreads;
	static constexpr int NR = BN * kBK / kThreads;
	T                    l[NL];
	T                    r[NR];
};

template <typename T, int BM, int BN>
__device__ inline void load_chunk(Stage<T, BM, BN>& st, const T* __restrict__ L,
                                  const T* __restrict__ R, const Seg& s, int m0, int n0, int M,
                                  int N, int k0, int tid)
{
	using S = Stage<T, BM, BN>;
	// L tile BM x 16: thread -> row tid % BM, columns tid / BM + (256 / BM) q (coalesced rows).
	{
		const int m  = m0 + (tid % BM);
		const int kb = k0 + (tid / BM);
#pragma unroll
		for (int q = 0; q < S::NL; ++q) {
			const int k = kb + (kThreads / BM) * q;
			st.l[q]     = (m < M && k < s.k) ? L[int64_t(m) + int64_t(k) * s.ldl] : T(0);
		}
	}
	if (s.rtrans) {
		// R stored N x K: element (k, n) at R[n + k*ldr]; coalesced along n.
		const int n  = n0 + (tid % BN);
		const int kb = k0 + (tid / BN);
#pragma unroll
		for (int q = 0; q < S::NR; ++q) {
			const int k = kb + (kThreads / BN) * q;
			st.r[q]     = (n < N && k < s.k) ? R[int64_t(n) + int64_t(k) * s.ldr] : T(0);
		}
	} else {
		// R stored K x N: element (k, n) at R[k + n*ldr]; coalesced along k.
		const int k  = k0 + (tid & 15);
		const int nb = n0 + (tid >> 4);
#pragma unroll
		for (int q = 0; q < S::NR; ++q) {
			const int n = nb + (kThreads / 16) * q;
			st.r[q]     = (n < N && k < s.k) ? R[int64_t(k) + int64_t(n) * s.ldr] : T(0);
		}
	}
}

template <typename T, int BM, int BN>
__device__ inline void store_chunk(const Stage<T, BM, BN>& st, T (*As)[BM + kPad],
                                   T (*Bs)[BN + kPad], int rtrans, int tid)
{
	using S = Stage<T, BM, BN>;
#pragma unroll
	for (int q = 0; q < S::NL; ++q)
		As[(tid / BM) + (kThreads / BM) * q][tid % BM] = st.l[q];
	if (rtrans) {
#pragma unroll
		for (int q = 0; q < S::NR; ++q)
			Bs[(tid / BN) + (kThreads / BN) * q][tid % BN] = st.r[q];
	} else {
#pragma unroll
		for (int q = 0; q < S::NR; ++q)
			Bs[tid & 15][(tid >> 4) + (kThreads / 16) * q] = st.r[q];
	}
}

// Advances (seg, k0) to the next K chunk of the task; returns false past the last one.
__device__ inline bool next_chunk(const Seg* __restrict__ segs, int seg_end, int& seg, int& k0)
{
	k0 += kBK;
	while (seg < seg_end && k0 >= segs[seg].k) {
		++seg;
		k0 = 0;
	}
	return seg < seg_end;
}

// The MFMAs of one staged K chunk for one wave. Partial: only the first fi_n x fj_n blocks
// (wave-uniform bounds), for tiles that overhang the task's edge.
template <bool Partial, typename T, typename Op, int BM, int BN, int FI, int FJ>
__device__ inline void mma_chunk(const T (*As)[BM + kPad], const T (*Bs)[BN + kPad], int lane,
                                 int wm, int wn, int fi_n, int fj_n,
                                 typename Op::acc_t (&acc)[FI][FJ])
{
#pragma unroll
	for (int kk = 0; kk < kBK; kk += Op::kK) {
		const int kr = kk + lane / Op::kM;
		T         a[FI], b[FJ];
#pragma unroll
		for (int i = 0; i < FI; ++i)
			if (!Partial || i < fi_n)
				a[i] = As[kr][wm + Op::kM * i + lane % Op::kM];
#pragma unroll
		for (int j = 0; j < FJ; ++j)
			if (!Partial || j < fj_n)
				b[j] = Bs[kr][wn + Op::kM * j + lane % Op::kM];
#pragma unroll
		for (int i = 0; i < FI; ++i)
#pragma unroll
			for (int j = 0; j < FJ; ++j)
				if (!Partial || (i < fi_n && j < fj_n))
					acc[i][j] = Op::mma(a[i], b[j], acc[i][j]);
	}
}

// T: operand/accumulate type; TO: output element type (the stage-2 output of the FP32 mode is
// written in FP64 for the FP64 Lanczos recurrence). BM x BN output tile per workgroup, each of
// the 4 waves owning a (BM/2) x (BN/2) sub-tile of (BM/32) x (BN/32) MFMA 16x16 blocks.
template <typename T, typename TO, int BM, int BN, typename Op = MfmaOp<T>>
__global__ __launch_bounds__(kThreads) void grouped_gemm_kernel(
    const Task* __restrict__ tasks, const Seg* __restrict__ segs, const Tile* __restrict__ tiles,
    Bases<T> bases, int64_t xstride, int64_t bxstride, TO* __restrict__ out_y, int64_t ystride,
    T* __restrict__ out_ws, int64_t wsstride, int beta)
{
	using acc_t = typename Op::acc_t;
	constexpr int WM = BM / 2, WN = BN / 2;
	constexpr int FI = WM / Op::kM, FJ = WN / Op::kM;
	static_assert(BM * kBK % kThreads == 0 && BN * kBK % kThreads == 0, "tile / thread mismatch");

	__shared__ T As[2][kBK][BM + kPad];
	__shared__ T Bs[2][kBK][BN + kPad];

	const Tile tl = tiles[blockIdx.x];
	if (tl.task < 0)
		return; // padding slot of the XCD-aware order
	const Task tk  = tasks[tl.task];
	const int  tid = threadIdx.x;
	const int  vec = blockIdx.y;

	// Per-vector operand bases (multi-vector launches put one right-hand side per grid row).
	// Selected through pick() rather than a dynamically indexed array so they stay in SGPRs.
	const T* b0   = bases.p[BASE_APACK];
	const T* b1   = bases.p[BASE_BPACK];
	const T* b2   = bases.p[BASE_X] + vec * xstride;
	const T* b3   = bases.p[BASE_BX] + vec * bxstride;
	auto     pick = [&](int id) -> const T* {
                return id == BASE_APACK ? b0 : id == BASE_BPACK ? b1 : id == BASE_X ? b2 : b3;
	};
	out_y += vec * ystride;
	out_ws += vec * wsstride;

	const int lane = tid & 63;
	const int wave = tid >> 6;
	const int wm   = (wave & 1) * WM;
	const int wn   = (wave >> 1) * WN;

	acc_t acc[FI][FJ];
#pragma unroll
	for (int i = 0; i < FI; ++i)
#pragma unroll
		for (int j = 0; j < FJ; ++j)
			acc[i][j] = acc_t {};
	// MFMA blocks of this wave that overlap the task (edge tiles): the others are skipped; their
	// accumulators stay zero and the epilogue drops them.
	const int fi_n = min(FI, max(0, (tk.m - tl.m0 - wm + Op::kM - 1) / Op::kM));
	const int fj_n = min(FJ, max(0, (tk.n - tl.n0 - wn + Op::kM - 1) / Op::kM));
	const bool full = fi_n == FI && fj_n == FJ;

	int seg = tk.seg_begin;
	int k0  = 0;
	while (seg < tk.seg_end && segs[seg].k <= 0)
		++seg;

	if (seg < tk.seg_end) {
		Stage<T, BM, BN> st;
		Seg              s = segs[seg];
		load_chunk<T, BM, BN>(st, pick(s.lbase) + s.loff, pick(s.rbase) + s.roff, s, tl.m0, tl.n0,
		                      tk.m, tk.n, k0, tid);
		store_chunk<T, BM, BN>(st, As[0], Bs[0], s.rtrans, tid);
		__syncthreads();
		int buf = 0;
		for (;;) {
			int  nseg    = seg;
			int  nk0     = k0;
			bool hasNext = next_chunk(segs, tk.seg_end, nseg, nk0);
			Seg  ns;
			if (hasNext) {
				ns = segs[nseg];
				load_chunk<T, BM, BN>(st, pick(ns.lbase) + ns.loff, pick(ns.rbase) + ns.roff, ns,
				                      tl.m0, tl.n0, tk.m, tk.n, nk0, tid);
			}
			if (full)
				mma_chunk<false, T, Op, BM, BN, FI, FJ>(As[buf], Bs[buf], lane, wm, wn, fi_n, fj_n, acc);
			else
				mma_chunk<true, T, Op, BM, BN, FI, FJ>(As[buf], Bs[buf], lane, wm, wn, fi_n, fj_n, acc);
			if (!hasNext)
				break;
			store_chunk<T, BM, BN>(st, As[buf ^ 1], Bs[buf ^ 1], ns.rtrans, tid);
			__syncthreads();
			buf ^= 1;
			seg = nseg;
			k0  = nk0;
		}
	}

	// Epilogue: register r of lane l holds row Op::row(l, r), column l % kM of each block.
	// Tasks writing the caller's vector convert to TO (and accumulate when beta); workspace
	// tasks (stage-1 panels, split-k partials) store the compute type and always overwrite.
	const bool to_y = tk.out_is_y != 0;
#pragma unroll
	for (int i = 0; i < FI; ++i)
#pragma unroll
		for (int j = 0; j < FJ; ++j)
#pragma unroll
			for (int r = 0; r < Op::kRegs; ++r) {
				const int row = tl.m0 + wm + Op::kM * i + Op::row(lane, r);
				const int col = tl.n0 + wn + Op::kM * j + lane % Op::kM;
				if (row < tk.m && col < tk.n) {
					const int     rr   = row % tk.rsplit;
					const int     blk  = row / tk.rsplit;
					const int64_t addr = tk.coff + int64_t(rr) + (int64_t(blk) * tk.n + col) * tk.ldc;
					if (to_y) {
						TO v = TO(acc[i][j][r]);
						if (beta)
							v += out_y[addr];
						out_y[addr] = v;
					} else {
						out_ws[addr] = acc[i][j][r];
					}
				}
			}
}

// Sums split-k partials in a fixed order (s = 0, 1, ...) into the output, accumulating in TO.
// Each thread owns kReduceLanes elements of the chunk (stride 256) and keeps their loads of one
// piece in flight together.
constexpr int kReduceLanes = 8;
static_assert(kReduceChunk == 256 * kReduceLanes, "reduce chunk / block mismatch");

template <typename T, typename TO>
__global__ __launch_bounds__(256) void splitk_reduce_kernel(const Reduce* __restrict__ red,
                                                            const ReduceChunk* __restrict__ chunks,
                                                            const T* __restrict__ ws,
                                                            int64_t wsstride, TO* __restrict__ out,
                                                            int64_t ystride, int beta)
{
	const ReduceChunk ch = chunks[blockIdx.x];
	const Reduce      rd = red[ch.r];
	ws += int64_t(blockIdx.y) * wsstride + rd.wsoff + ch.start;
	out += int64_t(blockIdx.y) * ystride + rd.yoff + ch.start;
	TO v[kReduceLanes];
#pragma unroll
	for (int j = 0; j < kReduceLanes; ++j)
		v[j] = 0;
	for (int s = 0; s < rd.nsplit; ++s) {
		const T* p = ws + int64_t(s) * rd.count;
#pragma unroll
		for (int j = 0; j < kReduceLanes; ++j) {
			const int e = int(threadIdx.x) + 256 * j;
			if (e < ch.len)
				v[j] += TO(p[e]);
		}
	}
#pragma unroll
	for (int j = 0; j < kReduceLanes; ++j) {
		const int e = int(threadIdx.x) + 256 * j;
		if (e < ch.len)
			out[e] = beta ? v[j] + out[e] : v[j];
	}
}

__global__ void cast_f64_to_f32_kernel(const double* __restrict__ x, float* __restrict__ y, int64_t n)
{
	int64_t i      = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	int64_t stride = int64_t(gridDim.x) * blockDim.x;
	for (; i < n; i += stride)
		y[i] = float(x[i]);
}

// Host-side launchers ------------------------------------------------------------------------

namespace {
template <typename T, typename TO, typename Op>
void launch_grouped_gemm_op(int bm, int bn, const Task* tasks, const Seg* segs, const Tile* tiles,
                            int ntiles, const Bases<T>& b, int64_t xstride, int64_t bxstride,
                            TO* out_y, int64_t ystride, T* out_ws, int64_t wsstride, int nvec,
                            int beta, hipStream_t stream)
{
	const dim3 grid { unsigned(ntiles), unsigned(nvec), 1u };
#define DMK_LAUNCH(BM, BN)                                                                          \
	hipLaunchKernelGGL((grouped_gemm_kernel<T, TO, BM, BN, Op>), grid, dim3(kThreads), 0,      \
	                   stream, tasks, segs, tiles, b, xstride, bxstride, out_y, ystride, out_ws,  \
	                   wsstride, beta)
	if (bm == 64 && bn == 64)
		DMK_LAUNCH(64, 64);
	else if (bm == 128 && bn == 64)
		DMK_LAUNCH(128, 64);
	else if (bm == 64 && bn == 128)
		DMK_LAUNCH(64, 128);
	else if (bm == 128 && bn == 128)
		DMK_LAUNCH(128, 128);
	else
		throw std::runtime_error("grouped gemm: unsupported tile shape");
#undef DMK_LAUNCH
}
} // namespace

// variant 1 selects the 32x32x2 MFMA form (FP32 only).
template <typename T, typename TO>
void launch_grouped_gemm(int bm, int bn, const Task* tasks, const Seg* segs, const Tile* tiles,
                         int ntiles, const T* const* bases, int64_t xstride, int64_t bxstride,
                         TO* out_y, int64_t ystride, T* out_ws, int64_t wsstride, int nvec,
                         int beta, hipStream_t stream, int variant)
{
	if (ntiles <= 0 || nvec <= 0)
		return;
	Bases<T> b;
	for (int i = 0; i < BASE_COUNT; ++i)
		b.p[i] = bases[i];
	if constexpr (std::is_same<T, float>::value) {
		if (variant == 1) {
			launch_grouped_gemm_op<T, TO, MfmaF32x32>(bm, bn, tasks, segs, tiles, ntiles, b,
			                                          xstride, bxstride, out_y, ystride, out_ws,
			                                          wsstride, nvec, beta, stream);
			return;
		}
	}
	launch_grouped_gemm_op<T, TO, MfmaOp<T>>(bm, bn, tasks, segs, tiles, ntiles, b, xstride,
	                                         bxstride, out_y, ystride, out_ws, wsstride, nvec,
	                                         beta, stream);
}

template <typename T, typename TO>
void launch_splitk_reduce(const Reduce* red, const ReduceChunk* chunks, int nchunks, const T* ws,
                          int64_t wsstride, TO* out, int64_t ystride, int nvec, int beta,
                          hipStream_t stream)
{
	if (nchunks <= 0 || nvec <= 0)
		return;
	const dim3 grid { unsigned(nchunks), unsigned(nvec), 1u };
	hipLaunchKernelGGL((splitk_reduce_kernel<T, TO>), grid, dim3(256), 0, stream, red, chunks, ws,
	                   wsstride, out, ystride, beta);
}

#define DMK_INST(T, TO)                                                                            \
	template void launch_grouped_gemm<T, TO>(int, int, const Task*, const Seg*, const Tile*, int, \
	                                         const T* const*, int64_t, int64_t, TO*, int64_t, T*,  \
	                                         int64_t, int, int, hipStream_t, int);                \
	template void launch_splitk_reduce<T, TO>(const Reduce*, const ReduceChunk*, int, const T*,   \
	                                          int64_t, TO*, int64_t, int, int, hipStream_t);
DMK_INST(double, double)
#undef DMK_INST
template void launch_splitk_reduce<float, double>(const Reduce*, const ReduceChunk*, int,
                                                  const float*, int64_t, double*, int64_t, int,
                                                  int, hipStream_t);

void launch_cast_f64_to_f32(const double* x, float* y, int64_t n, hipStream_t stream)
{
	if (n <= 0)
		return;
	int64_t blocks = (n + 255) / 256;
	if (blocks > 4096)
		blocks = 4096;
	hipLaunchKernelGGL(cast_f64_to_f32_kernel, dim3(unsigned(blocks)), dim3(256), 0, stream, x, y, n);
}

} // namespace dmk
