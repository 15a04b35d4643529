// BF16x3 split-precision variant of the grouped multi-segment GEMM (gfx950).
//
// Every operand value v is held as a pair of bf16 numbers, hi = bf16(v) and lo = bf16(v - hi),
// and a product is evaluated as lo_a*hi_b + hi_a*lo_b + hi_a*hi_b with three
// v_mfma_f32_16x16x32_bf16 instructions accumulating in FP32 (the lo*lo term, ~2^-16 relative,
// is dropped). The representation error of the pair is ~2^-17 relative, so a matvec carries an
// error of order 1e-5 relative; the Lanczos driver recovers FP64 accuracy with an FP64 restart.
//
// Same task / segment / tile tables as the FP64/FP32 kernel (kron_kernels.hip). Differences:
//  * K chunks of 32 (one 16x16x32 MFMA step), LDS tiles stored with k contiguous
//    ([row][k]) so that each lane's 8-element fragment is one 16-byte LDS read;
//  * the stage-1 epilogue writes the product as a (hi, lo) pair, ready for stage 2;
//  * split-k partials are FP32 and the stage-2 output is FP64.
// MFMA 16x16x32 bf16 lane layout (checked on gfx950): lane l supplies A[l%16][8(l/16)+i] and
// B[8(l/16)+i][l%16], i = 0..7, and holds D[4(l/16)+r][l%16] in register r.
#include "dmk_internal.h"

#include <hip/hip_runtime.h>

#include <stdexcept>

namespace dmk {

namespace {

using bf16_t = __bf16;
typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));
typedef float  f4_t __attribute__((ext_vector_type(4)));

constexpr int kK16      = 32;  // K chunk
constexpr int kPad16    = 8;   // row padding (elements); rows stay 16-byte aligned
constexpr int kThreads  = 256; // 4 wave64s in a 2x2 grid

struct Bases16 {
	const bf16_t* hi[BASE_COUNT];
	const bf16_t* lo[BASE_COUNT];
};

template <int BM, int BN> struct Stage16 {
	static constexpr int NL = BM * kK16 / kThreads;
	static constexpr int NR = BN * kK16 / kThreads;
	bf16_t               lh[NL], ll[NL], rh[NR], rl[NR];
};

template <int BM, int BN>
__device__ inline void load16(Stage16<BM, BN>& st, const bf16_t* __restrict__ Lh,
                              const bf16_t* __restrict__ Ll, const bf16_t* __restrict__ Rh,
                              const bf16_t* __restrict__ Rl, const Seg& s, int m0, int n0, int M,
                              int N, int k0, int tid)
{
	using S = Stage16<BM, BN>;
	const bf16_t zero = bf16_t(0.0f);
	{
		const int m  = m0 + (tid % BM);
		const int kb = k0 + (tid / BM);
#pragma unroll
		for (int q = 0; q < S::NL; ++q) {
			const int  k  = kb + (kThreads / BM) * q;
			const bool ok = (m < M && k < s.k);
			const int64_t o = int64_t(m) + int64_t(k) * s.ldl;
			st.lh[q] = ok ? Lh[o] : zero;
			st.ll[q] = ok ? Ll[o] : zero;
		}
	}
	if (s.rtrans) {
		const int n  = n0 + (tid % BN);
		const int kb = k0 + (tid / BN);
#pragma unroll
		for (int q = 0; q < S::NR; ++q) {
			const int     k  = kb + (kThreads / BN) * q;
			const bool    ok = (n < N && k < s.k);
			const int64_t o  = int64_t(n) + int64_t(k) * s.ldr;
			st.rh[q]         = ok ? Rh[o] : zero;
			st.rl[q]         = ok ? Rl[o] : zero;
		}
	} else {
		const int k  = k0 + (tid % kK16);
		const int nb = n0 + (tid / kK16);
#pragma unroll
		for (int q = 0; q < S::NR; ++q) {
			const int     n  = nb + (kThreads / kK16) * q;
			const bool    ok = (n < N && k < s.k);
			const int64_t o  = int64_t(k) + int64_t(n) * s.ldr;
			st.rh[q]         = ok ? Rh[o] : zero;
			st.rl[q]         = ok ? Rl[o] : zero;
		}
	}
}

template <int BM, int BN>
__device__ inline void store16(const Stage16<BM, BN>& st, bf16_t (*Ah)[kK16 + kPad16],
                               bf16_t (*Al)[kK16 + kPad16], bf16_t (*Bh)[kK16 + kPad16],
                               bf16_t (*Bl)[kK16 + kPad16], int rtrans, int tid)
{
	using S = Stage16<BM, BN>;
#pragma unroll
	for (int q = 0; q < S::NL; ++q) {
		const int m = tid % BM, k = (tid / BM) + (kThreads / BM) * q;
		Ah[m][k]    = st.lh[q];
		Al[m][k]    = st.ll[q];
	}
	if (rtrans) {
#pragma unroll
		for (int q = 0; q < S::NR; ++q) {
			const int n = tid % BN, k = (tid / BN) + (kThreads / BN) * q;
			Bh[n][k]    = st.rh[q];
			Bl[n][k]    = st.rl[q];
		}
	} else {
#pragma unroll
		for (int q = 0; q < S::NR; ++q) {
			const int k = tid % kK16, n = (tid / kK16) + (kThreads / kK16) * q;
			Bh[n][k]    = st.rh[q];
			Bl[n][k]    = st.rl[q];
		}
	}
}

__device__ inline bool next_chunk16(const Seg* __restrict__ segs, int seg_end, int& seg, int& k0)
{
	k0 += kK16;
	while (seg < seg_end && k0 >= segs[seg].k) {
		++seg;
		k0 = 0;
	}
	return seg < seg_end;
}

// WSPAIR: workspace tasks store a (hi, lo) bf16 pair (stage-1 panels) instead of FP32
// (split-k partials).
template <int BM, int BN, bool WSPAIR>
__global__ __launch_bounds__(kThreads) void grouped_gemm_bf16x3_kernel(
    const Task* __restrict__ tasks, const Seg* __restrict__ segs, const Tile* __restrict__ tiles,
    Bases16 bases, int64_t xstride, int64_t bxstride, double* __restrict__ out_y, int64_t ystride,
    float* __restrict__ out_f, bf16_t* __restrict__ out_h, bf16_t* __restrict__ out_l,
    int64_t wsstride, int beta)
{
	constexpr int WM = BM / 2, WN = BN / 2;
	constexpr int FI = WM / 16, FJ = WN / 16;

	__shared__ __attribute__((aligned(16))) bf16_t Ah[2][BM][kK16 + kPad16];
	__shared__ __attribute__((aligned(16))) bf16_t Al[2][BM][kK16 + kPad16];
	__shared__ __attribute__((aligned(16))) bf16_t Bh[2][BN][kK16 + kPad16];
	__shared__ __attribute__((aligned(16))) bf16_t Bl[2][BN][kK16 + kPad16];

	const Tile tl = tiles[blockIdx.x];
	if (tl.task < 0)
		return;
	const Task tk  = tasks[tl.task];
	const int  tid = threadIdx.x;
	const int  vec = blockIdx.y;

	const bf16_t* h0 = bases.hi[BASE_APACK];
	const bf16_t* h1 = bases.hi[BASE_BPACK];
	const bf16_t* h2 = bases.hi[BASE_X] + vec * xstride;
	const bf16_t* h3 = bases.hi[BASE_BX] + vec * bxstride;
	const bf16_t* l0 = bases.lo[BASE_APACK];
	const bf16_t* l1 = bases.lo[BASE_BPACK];
	const bf16_t* l2 = bases.lo[BASE_X] + vec * xstride;
	const bf16_t* l3 = bases.lo[BASE_BX] + vec * bxstride;
	auto pick_h = [&](int id) -> const bf16_t* {
		return id == BASE_APACK ? h0 : id == BASE_BPACK ? h1 : id == BASE_X ? h2 : h3;
	};
	auto pick_l = [&](int id) -> const bf16_t* {
		return id == BASE_APACK ? l0 : id == BASE_BPACK ? l1 : id == BASE_X ? l2 : l3;
	};
	out_y += vec * ystride;
	out_f += vec * wsstride;
	out_h += vec * wsstride;
	out_l += vec * wsstride;

	const int lane = tid & 63;
	const int wave = tid >> 6;
	const int wm   = (wave & 1) * WM;
	const int wn   = (wave >> 1) * WN;
	const int kq   = 8 * (lane >> 4);

	f4_t acc[FI][FJ];
#pragma unroll
	for (int i = 0; i < FI; ++i)
#pragma unroll
		for (int j = 0; j < FJ; ++j)
			acc[i][j] = f4_t { 0, 0, 0, 0 };

	int seg = tk.seg_begin;
	int k0  = 0;
	while (seg < tk.seg_end && segs[seg].k <= 0)
		++seg;

	if (seg < tk.seg_end) {
		Stage16<BM, BN> st;
		Seg             s = segs[seg];
		load16<BM, BN>(st, pick_h(s.lbase) + s.loff, pick_l(s.lbase) + s.loff,
		               pick_h(s.rbase) + s.roff, pick_l(s.rbase) + s.roff, s, tl.m0, tl.n0, tk.m,
		               tk.n, k0, tid);
		store16<BM, BN>(st, Ah[0], Al[0], Bh[0], Bl[0], s.rtrans, tid);
		__syncthreads();
		int buf = 0;
		for (;;) {
			int  nseg    = seg;
			int  nk0     = k0;
			bool hasNext = next_chunk16(segs, tk.seg_end, nseg, nk0);
			Seg  ns;
			if (hasNext) {
				ns = segs[nseg];
				load16<BM, BN>(st, pick_h(ns.lbase) + ns.loff, pick_l(ns.lbase) + ns.loff,
				               pick_h(ns.rbase) + ns.roff, pick_l(ns.rbase) + ns.roff, ns, tl.m0,
				               tl.n0, tk.m, tk.n, nk0, tid);
			}
			bf8_t ah[FI], al[FI], bh[FJ], bl[FJ];
#pragma unroll
			for (int i = 0; i < FI; ++i) {
				ah[i] = *reinterpret_cast<const bf8_t*>(&Ah[buf][wm + 16 * i + (lane & 15)][kq]);
				al[i] = *reinterpret_cast<const bf8_t*>(&Al[buf][wm + 16 * i + (lane & 15)][kq]);
			}
#pragma unroll
			for (int j = 0; j < FJ; ++j) {
				bh[j] = *reinterpret_cast<const bf8_t*>(&Bh[buf][wn + 16 * j + (lane & 15)][kq]);
				bl[j] = *reinterpret_cast<const bf8_t*>(&Bl[buf][wn + 16 * j + (lane & 15)][kq]);
			}
#pragma unroll
			for (int i = 0; i < FI; ++i)
#pragma unroll
				for (int j = 0; j < FJ; ++j) {
					acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
					acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
					acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
				}
			if (!hasNext)
				break;
			store16<BM, BN>(st, Ah[buf ^ 1], Al[buf ^ 1], Bh[buf ^ 1], Bl[buf ^ 1], ns.rtrans, tid);
			__syncthreads();
			buf ^= 1;
			seg = nseg;
			k0  = nk0;
		}
	}

	const bool to_y = tk.out_is_y != 0;
#pragma unroll
	for (int i = 0; i < FI; ++i)
#pragma unroll
		for (int j = 0; j < FJ; ++j)
#pragma unroll
			for (int r = 0; r < 4; ++r) {
				const int row = tl.m0 + wm + 16 * i + 4 * (lane >> 4) + r;
				const int col = tl.n0 + wn + 16 * j + (lane & 15);
				if (row < tk.m && col < tk.n) {
					const int     rr   = row % tk.rsplit;
					const int     blk  = row / tk.rsplit;
					const int64_t addr = tk.coff + int64_t(rr) + (int64_t(blk) * tk.n + col) * tk.ldc;
					const float   v    = acc[i][j][r];
					if (to_y) {
						double y = double(v);
						if (beta)
							y += out_y[addr];
						out_y[addr] = y;
					} else if (WSPAIR) {
						const bf16_t hi = bf16_t(v);
						out_h[addr]     = hi;
						out_l[addr]     = bf16_t(v - float(hi));
					} else {
						out_f[addr] = v;
					}
				}
			}
}

// hi = bf16(x), lo = bf16(x - hi), the residual taken in double.
__global__ void split_f64_bf16x2_kernel(const double* __restrict__ x, bf16_t* __restrict__ hi,
                                        bf16_t* __restrict__ lo, int64_t n)
{
	int64_t stride = int64_t(gridDim.x) * blockDim.x;
	for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
		const double v = x[i];
		const bf16_t h = bf16_t(float(v));
		hi[i]          = h;
		lo[i]          = bf16_t(float(v - double(float(h))));
	}
}

} // namespace

void launch_split_f64_bf16x2(const double* x, void* hi, void* lo, int64_t n, hipStream_t stream)
{
	if (n <= 0)
		return;
	int64_t blocks = (n + 255) / 256;
	if (blocks > 4096)
		blocks = 4096;
	hipLaunchKernelGGL(split_f64_bf16x2_kernel, dim3(unsigned(blocks)), dim3(256), 0, stream, x,
	                   static_cast<bf16_t*>(hi), static_cast<bf16_t*>(lo), n);
}

// bases_hi/lo: the four operand bases as bf16 arrays. Stage 1 passes ws_pair = true (workspace
// outputs are bf16 pairs at out_h/out_l), stage 2 false (FP32 partials at out_f).
void launch_grouped_gemm_bf16x3(int bm, int bn, const Task* tasks, const Seg* segs,
                                const Tile* tiles, int ntiles, const void* const* bases_hi,
                                const void* const* bases_lo, int64_t xstride, int64_t bxstride,
                                double* out_y, int64_t ystride, float* out_f, void* out_h,
                                void* out_l, int64_t wsstride, bool ws_pair, int nvec, int beta,
                                hipStream_t stream)
{
	if (ntiles <= 0 || nvec <= 0)
		return;
	Bases16 b;
	for (int i = 0; i < BASE_COUNT; ++i) {
		b.hi[i] = static_cast<const bf16_t*>(bases_hi[i]);
		b.lo[i] = static_cast<const bf16_t*>(bases_lo[i]);
	}
	bf16_t*    oh = static_cast<bf16_t*>(out_h);
	bf16_t*    ol = static_cast<bf16_t*>(out_l);
	const dim3 grid { unsigned(ntiles), unsigned(nvec), 1u };
#define DMK_LAUNCH16(BM, BN, PAIR)                                                                  \
	hipLaunchKernelGGL((grouped_gemm_bf16x3_kernel<BM, BN, PAIR>), grid, dim3(kThreads), 0,     \
	                   stream, tasks, segs, tiles, b, xstride, bxstride, out_y, ystride, out_f, oh, \
	                   ol, wsstride, beta)
#define DMK_SHAPES16(PAIR)                                                                          \
	if (bm == 64 && bn == 64)                                                                  \
		DMK_LAUNCH16(64, 64, PAIR);                                                        \
	else if (bm == 128 && bn == 64)                                                            \
		DMK_LAUNCH16(128, 64, PAIR);                                                       \
	else if (bm == 64 && bn == 128)                                                            \
		DMK_LAUNCH16(64, 128, PAIR);                                                       \
	else if (bm == 128 && bn == 128)                                                           \
		DMK_LAUNCH16(128, 128, PAIR);                                                      \
	else                                                                                       \
		throw std::runtime_error("grouped gemm bf16x3: unsupported tile shape");
	if (ws_pair) {
		DMK_SHAPES16(true)
	} else {
		DMK_SHAPES16(false)
	}
#undef DMK_SHAPES16
#undef DMK_LAUNCH16
}

} // namespace dmk
