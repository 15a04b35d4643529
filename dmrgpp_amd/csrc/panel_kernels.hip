// Grouped multi-segment GEMM on K-major operand panels for CDNA4 (gfx950): the kernels of the
// superblock matvec in all three precision modes, and the per-call preparation of the input.
//
// One launch computes a list of independent output matrices ("tasks"), each the sum of one or
// more products ("segments") C += L * R. The matvec uses two launches (kron_plan.cpp):
//   stage 1: for every (I,J) with general terms, [B_k1; B_k2; ...] * X_J (row-stacked B
//            blocks), written into the stage-2 panel BX_I;
//   stage 2: Y_I^T = A_I * BX_I^T + A * X_J^T (identity right factor) + X_J^T * B^T (identity
//            left factor), i.e. the transpose of the output patch, so that consecutive lanes
//            write consecutive elements of the column-major Y_I.
// This is the two-stage evaluation of GPUPlugin/apply_Htarget_sparse_not_ready.cpp:199-449 with
// the identity factors of H_L (x) I and I (x) H_R contracted directly.
//
// Every operand is K-major with rows zero padded to a multiple of kPanelPad (64) elements:
// element (m, k) of a left operand at loff + k + m * ldl, element (k, n) of a right operand at
// roff + k + n * ldr. The planner packs the factors that way, the input vector is rewritten that
// way on every call (prep_x_kernel: X_J column major with padded columns, and X_J transposed),
// and stage 1 writes BX_I that way; split-k pieces start on multiples of kPanelPad. So every K
// chunk of a segment (BK <= 64 elements from a multiple of BK) lies inside its rows, every
// global load is a 16-byte vector along k, every LDS store and fragment read a 16-byte vector,
// and the only edge handling is a row clamp (rows past the task edge are computed and dropped).
//
// Tiling: 256 threads = 4 wave64s in a 2x2 grid; a workgroup owns a BM x BN output tile and
// each wave a (BM/2) x (BN/2) sub-tile of MFMA blocks. K is walked in chunks of BK staged
// through LDS. In both kernels (panel_gemm_kernel for FP64 / FP32, panel_bf16x3_kernel for the
// hi and lo parts of BF16x3) the chunks move from global memory straight into a ring of two or
// three LDS stages (LDS-DMA), with no staging registers and no address arithmetic per chunk,
// and chunk c+1 (or c+1 and c+2) is in flight during the MFMAs of chunk c.
// LDS rows hold the BK elements of one operand row as 16-byte slots, XOR-swizzled by
// row (slot kv of row r at kv ^ swz(r)) so that the 16 rows one 16-lane pass reads fall on
// distinct banks, without padding.
// Within a chunk the lanes' k order is permuted (lane group g takes k = g * NF + s at MFMA step
// s) so that a lane's operands for all NF steps are NF consecutive elements of one LDS row; the
// sum over k is the same. Every output tile is owned by one workgroup and its contraction order
// is fixed: results are bitwise reproducible (no atomics; split-k partials are summed in a fixed
// order by splitk_reduce_kernel).
#include "dmk_internal.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>
#include <type_traits>

namespace dmk {

namespace {

constexpr int kPThreads = 256;

typedef double d2_t __attribute__((ext_vector_type(2)));
typedef float  f4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));

template <typename T> struct Vec16;
template <> struct Vec16<double> {
	using type = d2_t;
};
template <> struct Vec16<float> {
	using type = f4_t;
};
template <> struct Vec16<__bf16> {
	using type = bf8_t;
};

// MFMA forms. kM x kM output block, kK-deep contraction per instruction; lane l supplies
// A[l % kM][g] and B[g][l % kM] for the k index g = l / kM (the bf16 forms: 8 consecutive k
// from 8 * (l / kM)); register r of the result holds row row(l, r), column l % kM
// (layouts checked on gfx950).
struct OpF64 {
	using T                    = double;
	static constexpr int kM    = 16;
	static constexpr int kK    = 4;
	static constexpr int kRegs = 4;
	using acc_t                = double __attribute__((ext_vector_type(4)));
	__device__ static inline acc_t mma(double a, double b, acc_t c)
	{
		return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
	}
	__device__ static inline int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};

struct OpF32 {
	using T                    = float;
	static constexpr int kM    = 16;
	static constexpr int kK    = 4;
	static constexpr int kRegs = 4;
	using acc_t                = float __attribute__((ext_vector_type(4)));
	__device__ static inline acc_t mma(float a, float b, acc_t c)
	{
		return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
	}
	__device__ static inline int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

struct OpF32x32 {
	using T                    = float;
	static constexpr int kM    = 32;
	static constexpr int kK    = 2;
	static constexpr int kRegs = 16;
	using acc_t                = float __attribute__((ext_vector_type(16)));
	__device__ static inline acc_t mma(float a, float b, acc_t c)
	{
		return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
	}
	__device__ static inline int row(int lane, int r)
	{
		return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
	}
};

struct OpBf16 {
	static constexpr int kM    = 16;
	static constexpr int kK    = 32; // 8 consecutive k per lane, 4 lane groups
	static constexpr int kRegs = 4;
	using acc_t                = float __attribute__((ext_vector_type(4)));
	__device__ static inline acc_t mma(bf8_t a, bf8_t b, acc_t c)
	{
		return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
	}
	__device__ static inline int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

struct OpBf16x32 {
	static constexpr int kM    = 32;
	static constexpr int kK    = 16; // 8 consecutive k per lane, 2 lane groups
	static constexpr int kRegs = 16;
	using acc_t                = float __attribute__((ext_vector_type(16)));
	__device__ static inline acc_t mma(bf8_t a, bf8_t b, acc_t c)
	{
		return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
	}
	__device__ static inline int row(int lane, int r)
	{
		return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
	}
};

template <typename T> struct PBasesT {
	const T* p[PBASE_COUNT];
};

// Address of element (r, c) of a task's output.
__device__ inline int64_t out_addr(const PTask& tk, int r, int c)
{
	const int rr  = r % tk.rsplit;
	const int blk = r / tk.rsplit;
	return tk.coff + int64_t(rr) * tk.sm + int64_t(blk) * tk.n + c;
}

// Element i of a descriptor array, read through the constant address space: scalar loads, which
// an LDS-DMA load in flight (a vector memory operation that writes memory) does not make wait.
template <typename S> __device__ inline S load_const(const S* p, int i)
{
#if defined(__HIP_DEVICE_COMPILE__)
	typedef const __attribute__((address_space(4))) S* cptr;
	return ((cptr)p)[i];
#else
	return p[i]; // the host pass only checks the kernels
#endif
}

// Advances (seg, k0) to the next K chunk of the task, with K the current segment's length (held
// in registers, so the common step reads no descriptor); false past the last chunk. newseg is
// set when the chunk starts a new segment.
template <int BK>
__device__ inline bool next_chunk(const PSeg* __restrict__ segs, int seg_end, int K, int& seg,
                                  int& k0, bool& newseg)
{
	k0 += BK;
	newseg = false;
	if (k0 < K)
		return true;
	k0     = 0;
	newseg = true;
	do
		++seg;
	while (seg < seg_end && load_const(segs, seg).k <= 0);
	return seg < seg_end;
}

// Number of BK chunks of a task (segments with k <= 0 have none).
template <int BK>
__device__ inline int chunk_count(const PSeg* __restrict__ segs, int seg_begin, int seg_end)
{
	int n = 0;
	for (int q = seg_begin; q < seg_end; ++q)
		n += max(0, (load_const(segs, q).k + BK - 1) / BK);
	return n;
}

// LDS layout of one operand tile (ROWS x BK elements, VW-element 16-byte vectors, VPR vectors
// per row).
template <typename E, int ROWS, int BK> struct TileLoad {
	using vec_t              = typename Vec16<E>::type;
	static constexpr int VW  = 16 / int(sizeof(E));
	static constexpr int VPR = BK / VW;
	static_assert(ROWS * VPR % kPThreads == 0, "tile / thread mismatch");
	// Slot of vector kv of LDS row r. A 16-byte LDS access is served 8 lanes (32 banks) or 16
	// lanes (64 banks) per pass; the fragment reads put consecutive rows on consecutive lanes,
	// so the XOR mask f(r) keeps any 8 consecutive rows, and any 16 consecutive rows that start
	// on the same bank, on distinct slots: f is a permutation of 0..VPR-1 on rows 8a..8a+7, on
	// the even and on the odd rows of 16a..16a+15 (VPR = 8), and on each residue class mod 2
	// within 8 rows and mod 4 within 16 rows (VPR = 4).
	__device__ static inline int slot(int r, int kv)
	{
		int f;
		if constexpr (VPR == 8)
			f = (r & 7) ^ ((r >> 3) & 1);
		else if constexpr (VPR == 4)
			f = ((r >> 1) & 3) ^ ((r >> 3) & 1);
		else
			f = r & (VPR - 1);
		return kv ^ f;
	}
};

// LDS address-space view of a __shared__ pointer (the destination operand of an LDS-DMA load).
[[maybe_unused]] __device__ inline __attribute__((address_space(3))) void* lds_ptr(void* p)
{
	return (__attribute__((address_space(3))) void*)p;
}

// s_waitcnt vmcnt(N) lgkmcnt(0): this wave's vector memory operations but the last N issued have
// completed, and its LDS reads too (gfx9 encoding: vmcnt bits [3:0] and [15:14], expcnt [6:4],
// lgkmcnt [11:8]).
template <int N> __device__ inline void wait_vm_lgkm0()
{
	static_assert(N >= 0 && N < 64, "vmcnt range");
	__builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4));
}

// A workgroup barrier without the memory fence of __syncthreads (which would wait for every
// LDS-DMA load in flight, the prefetched chunks included); the memory clobber keeps the
// compiler from moving LDS accesses across it. Callers wait for what they need first.
__device__ inline void barrier_no_fence() { asm volatile("s_barrier" ::: "memory"); }

// Direct global -> LDS loads (LDS-DMA, global_load_lds_dwordx4) of one operand tile of ROWS x
// BK elements. The 16-byte LDS slots of the tile are filled 64 at a time: instruction q of wave
// w writes slots (4 q + w) * 64 ... + 63, lane l slot (4 q + w) * 64 + l, and the slot's row r
// and vector kv follow from the XOR swizzle of TileLoad (slot kv ^ f(r) of row r holds vector
// kv; the map is an involution per row). Every lane's source is the uniform segment base plus
// a per-lane 32-bit offset fixed for the segment, so a chunk costs no vector ALU work. Rows past
// the task edge are clamped to its last row (computed and dropped). Every chunk lies inside its
// rows: rows are zero padded to a multiple of kPanelPad >= BK elements and split-k cuts fall on
// multiples of kPanelPad.
template <typename E, int ROWS, int BK> struct TileDma {
	using TL                 = TileLoad<E, ROWS, BK>;
	using vec_t              = typename TL::vec_t;
	static constexpr int VW  = TL::VW;
	static constexpr int VPR = TL::VPR;
	static constexpr int NI  = ROWS * VPR / kPThreads; // instructions per wave per chunk
	static_assert(NI >= 1 && ROWS * VPR % kPThreads == 0, "tile / thread mismatch");
	static_assert(kPanelPad % BK == 0, "chunks must fit the row padding");
	const E* base;    // the segment's operand at the tile's first row (wave-uniform)
	uint32_t off[NI]; // byte offsets of the lane's vectors from base

	__device__ inline void setup(const E* seg, int ld, int row0, int nrows, int lane, int wave)
	{
		base = seg + int64_t(row0) * ld;
#pragma unroll
		for (int q = 0; q < NI; ++q) {
			const int s  = (q * 4 + wave) * 64 + lane;
			const int r  = s / VPR;
			const int kv = TL::slot(r, s % VPR);
			const int dr = min(row0 + r, nrows - 1) - row0;
			off[q]       = uint32_t((dr * ld + kv * VW) * int(sizeof(E)));
		}
	}
	__device__ inline void dma(vec_t* stage, int k0, int wave) const
	{
		const char* b = reinterpret_cast<const char*>(base + k0);
#pragma unroll
		for (int q = 0; q < NI; ++q) {
			// The builtin exists for the device only; the host pass sees the kernel's signature.
#if defined(__HIP_DEVICE_COMPILE__)
			__builtin_amdgcn_global_load_lds(b + off[q], lds_ptr(stage + (q * 4 + wave) * 64), 16,
			                                 0, 0);
#else
			(void)b;
			(void)stage;
#endif
		}
	}
};

// FP64 / FP32: T operands, TO output for tasks writing the caller's vector (the FP32 mode writes
// its stage-2 output in FP64 for the FP64 Lanczos recurrence). Workspace tasks (stage-1 panels,
// split-k partials) store T and always overwrite.
//
// The K chunks go through a ring of NS LDS stages filled by LDS-DMA: with NS = 2 chunk c+1 lands
// during the MFMAs of chunk c, with NS = 3 chunks c+1 and c+2. Each wave waits for its own loads
// of the next chunk (vmcnt) and then meets the others at a barrier, after which the chunk is in
// LDS for all of them; a stage is refilled only after the barrier that follows its last reads.
// The stages are distinct __shared__ arrays and the ring is unrolled NS times, so that the
// compiler can tell the stage being read from the stages being filled.
//
// With bins != nullptr the launch is persistent and balanced: workgroup b runs the items
// [bins[b], bins[b + 1]) one after another (SkItem: a chunk range of one tile, its result to the
// task's output or to a partial slot of out_ws for the fix-up); else workgroup b runs tile b
// whole.
template <typename TO, int BM, int BN, int BK, int NS, typename Op>
__global__ __launch_bounds__(kPThreads) void panel_gemm_kernel(
    const PTask* __restrict__ tasks, const PSeg* __restrict__ segs, const Tile* __restrict__ tiles,
    const SkItem* __restrict__ items, const int32_t* __restrict__ bins,
    PBasesT<typename Op::T> bases, PStrides vs, TO* __restrict__ out_y, int64_t ystride,
    typename Op::T* __restrict__ out_ws, int64_t wsstride, int beta)
{
	using T     = typename Op::T;
	using acc_t = typename Op::acc_t;
	using LA    = TileDma<T, BM, BK>;
	using LB    = TileDma<T, BN, BK>;
	using SA    = TileLoad<T, BM, BK>;
	using SB    = TileLoad<T, BN, BK>;
	using vec_t = typename LA::vec_t;
	constexpr int VW  = LA::VW;
	constexpr int VPR = LA::VPR; // 16-byte slots per LDS row
	constexpr int WM = BM / 2, WN = BN / 2;
	constexpr int FI = WM / Op::kM, FJ = WN / Op::kM;
	constexpr int NF = BK / Op::kK;          // consecutive k per lane per chunk
	constexpr int NW = LA::NI + LB::NI;      // LDS-DMA instructions per wave per chunk
	static_assert(NF % VW == 0, "fragment / vector mismatch");
	static_assert(NS == 2 || NS == 3, "two or three stages");

	__shared__ vec_t A0[BM * VPR], B0[BN * VPR], A1[BM * VPR], B1[BN * VPR];
	__shared__ vec_t A2[NS == 3 ? BM * VPR : 1], B2[NS == 3 ? BN * VPR : 1];

	const int tid = threadIdx.x;
	const int vec = blockIdx.y;

	const T* b0   = bases.p[PBASE_APACK];
	const T* b1   = bases.p[PBASE_BPACK];
	const T* b2   = bases.p[PBASE_XP] + vec * vs.xp;
	const T* b3   = bases.p[PBASE_XT] + vec * vs.xt;
	const T* b4   = bases.p[PBASE_BX] + vec * vs.bx;
	auto     pick = [&](int id) -> const T* {
                return id == PBASE_APACK ? b0 : id == PBASE_BPACK ? b1 : id == PBASE_XP ? b2 : id == PBASE_XT ? b3 : b4;
	};
	out_y += vec * ystride;
	out_ws += vec * wsstride;

	const int lane = tid & 63;
	const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
	const int wm   = (wave & 1) * WM;
	const int wn   = (wave >> 1) * WN;
	const int g    = lane / Op::kM; // lane group: k slice g * NF .. g * NF + NF - 1
	const int lr   = lane % Op::kM;

	const int it_end = bins ? bins[blockIdx.x + 1] : 1;
	for (int it = bins ? bins[blockIdx.x] : 0; it < it_end; ++it) {
		Tile tl;
		int  skip = 0, nch_item = -1, slot = -1; // chunks to skip, chunks to run (-1: all), slot
		if (bins) {
			const SkItem w = items[it];
			tl             = Tile { w.task, w.m0, w.n0, 0 };
			skip           = w.c0;
			nch_item       = w.nch;
			slot           = w.slot;
		} else {
			tl = tiles[blockIdx.x];
		}
		if (tl.task < 0)
			continue; // padding slot of the XCD-aware order
		const PTask tk = load_const(tasks, tl.task);

		acc_t acc[FI][FJ];
#pragma unroll
		for (int i = 0; i < FI; ++i)
#pragma unroll
			for (int j = 0; j < FJ; ++j)
				acc[i][j] = acc_t {};

		// The item's first chunk: `skip` whole chunks into the task's segments (with k > 0).
		int seg = tk.seg_begin;
		while (seg < tk.seg_end && load_const(segs, seg).k <= 0)
			++seg;
		int kstart = 0;
		while (seg < tk.seg_end) {
			const int nc = (load_const(segs, seg).k + BK - 1) / BK;
			if (skip < nc) {
				kstart = skip * BK;
				break;
			}
			skip -= nc;
			do
				++seg;
			while (seg < tk.seg_end && load_const(segs, seg).k <= 0);
		}

		if (seg < tk.seg_end && nch_item != 0) {
			// Load cursor: the next chunk to fetch (segment seg, offset ck0). Past the tile's last
			// chunk the last one is fetched again, into a stage that is not read any more, so
			// that every step has the same loads in flight (past an item's last chunk, the
			// task's next chunk is fetched and never read).
			LA   la;
			LB   lb;
			PSeg s      = load_const(segs, seg);
			int  ck0    = kstart;
			int  lastk  = kstart;
			bool cvalid = true;
			la.setup(pick(s.lbase) + s.loff, s.ldl, tl.m0, tk.m, lane, wave);
			lb.setup(pick(s.rbase) + s.roff, s.ldr, tl.n0, tk.n, lane, wave);
			auto issue = [&](vec_t* sa, vec_t* sb) {
				const int kk = cvalid ? ck0 : lastk;
				la.dma(sa, kk, wave);
				lb.dma(sb, kk, wave);
				lastk = kk;
				if (cvalid) {
					bool newseg;
					cvalid = next_chunk<BK>(segs, tk.seg_end, s.k, seg, ck0, newseg);
					if (cvalid && newseg) {
						s = load_const(segs, seg);
						la.setup(pick(s.lbase) + s.loff, s.ldl, tl.m0, tk.m, lane, wave);
						lb.setup(pick(s.rbase) + s.roff, s.ldr, tl.n0, tk.n, lane, wave);
					}
				}
			};
			auto compute = [&](const vec_t* sa, const vec_t* sb) {
				// A lane's NF consecutive k of one LDS row per operand block, one 16-byte vector
				// (VW of the NF MFMA steps) at a time, so that only VW steps of fragments are live.
#pragma unroll
				for (int v = 0; v < NF / VW; ++v) {
					vec_t xa[FI], xb[FJ];
#pragma unroll
					for (int i = 0; i < FI; ++i) {
						const int r = wm + Op::kM * i + lr;
						xa[i]       = sa[r * VPR + SA::slot(r, g * (NF / VW) + v)];
					}
#pragma unroll
					for (int j = 0; j < FJ; ++j) {
						const int r = wn + Op::kM * j + lr;
						xb[j]       = sb[r * VPR + SB::slot(r, g * (NF / VW) + v)];
					}
#pragma unroll
					for (int e = 0; e < VW; ++e)
#pragma unroll
						for (int i = 0; i < FI; ++i)
#pragma unroll
							for (int j = 0; j < FJ; ++j)
								acc[i][j] = Op::mma(T(xa[i][e]), T(xb[j][e]), acc[i][j]);
				}
			};
			const int nch =
			    nch_item > 0 ? nch_item : chunk_count<BK>(segs, seg, tk.seg_end) - kstart / BK;
			if constexpr (NS == 2) {
				// At the top of each step stage 0 holds chunk c; chunk c+1 goes to stage 1.
				issue(A0, B0);
				wait_vm_lgkm0<0>();
				barrier_no_fence();
				int c = 0;
				for (; c + 2 <= nch; c += 2) {
					issue(A1, B1);
					compute(A0, B0);
					wait_vm_lgkm0<0>();
					barrier_no_fence();
					if (c + 2 == nch) {
						compute(A1, B1);
						break;
					}
					issue(A0, B0);
					compute(A1, B1);
					wait_vm_lgkm0<0>();
					barrier_no_fence();
				}
				if (c + 1 == nch)
					compute(A0, B0);
			} else {
				// At the top of each step stage 0 holds chunk c and chunk c+1 is in flight to
				// stage 1; chunk c+2 goes to stage 2, which was last read one step earlier.
				issue(A0, B0);
				issue(A1, B1);
				wait_vm_lgkm0<NW>();
				barrier_no_fence();
				int c = 0;
				for (; c + 3 <= nch; c += 3) {
					issue(A2, B2);
					compute(A0, B0);
					wait_vm_lgkm0<NW>();
					barrier_no_fence();
					issue(A0, B0);
					compute(A1, B1);
					wait_vm_lgkm0<NW>();
					barrier_no_fence();
					issue(A1, B1);
					compute(A2, B2);
					wait_vm_lgkm0<NW>();
					barrier_no_fence();
				}
				// Zero, one or two chunks remain, in stages 0 and 1.
				if (c < nch) {
					compute(A0, B0);
					if (c + 1 < nch) {
						wait_vm_lgkm0<0>();
						barrier_no_fence();
						compute(A1, B1);
					}
				}
			}
			// Every load has landed before the stages are reused or the workgroup ends.
			wait_vm_lgkm0<0>();
		}

		if (slot >= 0) {
			// A partial of a tile cut between workgroups: the whole BM x BN tile, tile-local.
			T* part = out_ws + int64_t(slot) * (BM * BN);
#pragma unroll
			for (int i = 0; i < FI; ++i)
#pragma unroll
				for (int j = 0; j < FJ; ++j)
#pragma unroll
					for (int r = 0; r < Op::kRegs; ++r)
						part[(wm + Op::kM * i + Op::row(lane, r)) * BN + wn + Op::kM * j + lr] =
						    acc[i][j][r];
		} else {
			const bool to_y = tk.out_is_y != 0;
#pragma unroll
			for (int i = 0; i < FI; ++i)
#pragma unroll
				for (int j = 0; j < FJ; ++j)
#pragma unroll
					for (int r = 0; r < Op::kRegs; ++r) {
						const int row = tl.m0 + wm + Op::kM * i + Op::row(lane, r);
						const int col = tl.n0 + wn + Op::kM * j + lr;
						if (row < tk.m && col < tk.n) {
							const int64_t addr = out_addr(tk, row, col);
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
		// The next item's first stores go to LDS buffer 0, which the last chunk may have used.
		__syncthreads();
	}
}

// BF16x3: every operand value v is held as hi = bf16(v), lo = bf16(v - hi) in two arrays of the
// same layout, and a product is lo_a*hi_b + hi_a*lo_b + hi_a*hi_b on bf16 MFMAs accumulating in
// FP32 (lo*lo, ~2^-16 relative, is dropped). PAIR: workspace tasks store a (hi, lo) pair
// (stage-1 panels) instead of FP32 (split-k partials). Stage-2 output to the caller is FP64.
// The K chunks (hi and lo of both operands) go through a ring of NS LDS stages filled by
// LDS-DMA, as in panel_gemm_kernel.
template <int BM, int BN, int BK, int NS, bool PAIR, typename Op>
__global__ __launch_bounds__(kPThreads) void panel_bf16x3_kernel(
    const PTask* __restrict__ tasks, const PSeg* __restrict__ segs, const Tile* __restrict__ tiles,
    const SkItem* __restrict__ items, const int32_t* __restrict__ bins,
    PBasesT<__bf16> hi, PBasesT<__bf16> lo, PStrides vs, double* __restrict__ out_y,
    int64_t ystride, float* __restrict__ out_f, __bf16* __restrict__ out_h,
    __bf16* __restrict__ out_l, int64_t wsstride, int beta)
{
	using acc_t = typename Op::acc_t;
	using LA    = TileDma<__bf16, BM, BK>;
	using LB    = TileDma<__bf16, BN, BK>;
	using SA    = TileLoad<__bf16, BM, BK>;
	using SB    = TileLoad<__bf16, BN, BK>;
	constexpr int VPR = LA::VPR;
	constexpr int WM = BM / 2, WN = BN / 2;
	constexpr int FI = WM / Op::kM, FJ = WN / Op::kM;
	constexpr int NST = BK / Op::kK;         // MFMA steps per chunk
	constexpr int NG  = 64 / Op::kM;         // lane groups, 8 consecutive k each per step
	constexpr int NW  = 2 * (LA::NI + LB::NI); // LDS-DMA instructions per wave per chunk
	static_assert(NS == 2 || NS == 3, "two or three stages");

	// One stage: the hi and lo parts of both operand tiles of a chunk.
	struct Stage {
		bf8_t ah[BM * VPR], al[BM * VPR], bh[BN * VPR], bl[BN * VPR];
	};
	__shared__ Stage S0, S1;
	__shared__ bf8_t S2raw[NS == 3 ? sizeof(Stage) / sizeof(bf8_t) : 1];
	Stage& S2 = *reinterpret_cast<Stage*>(S2raw); // used only when NS == 3

	const int tid = threadIdx.x;
	const int vec = blockIdx.y;

	const int64_t vo[PBASE_COUNT] = { 0, 0, vec * vs.xp, vec * vs.xt, vec * vs.bx };
	auto pick = [&](const PBasesT<__bf16>& b, int id) -> const __bf16* {
		const __bf16* p = id == PBASE_APACK ? b.p[PBASE_APACK]
		                  : id == PBASE_BPACK ? b.p[PBASE_BPACK]
		                  : id == PBASE_XP    ? b.p[PBASE_XP] + vo[PBASE_XP]
		                  : id == PBASE_XT    ? b.p[PBASE_XT] + vo[PBASE_XT]
		                                      : b.p[PBASE_BX] + vo[PBASE_BX];
		return p;
	};
	out_y += vec * ystride;
	out_f += vec * wsstride;
	out_h += vec * wsstride;
	out_l += vec * wsstride;

	const int lane = tid & 63;
	const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
	const int wm   = (wave & 1) * WM;
	const int wn   = (wave >> 1) * WN;
	const int g    = lane / Op::kM;
	const int lr   = lane % Op::kM;

	const int it_end = bins ? bins[blockIdx.x + 1] : 1;
	for (int it = bins ? bins[blockIdx.x] : 0; it < it_end; ++it) {
		// The work item, as in panel_gemm_kernel.
		Tile tl;
		int  skip = 0, nch_item = -1, slot = -1;
		if (bins) {
			const SkItem w = items[it];
			tl             = Tile { w.task, w.m0, w.n0, 0 };
			skip           = w.c0;
			nch_item       = w.nch;
			slot           = w.slot;
		} else {
			tl = tiles[blockIdx.x];
		}
		if (tl.task < 0)
			continue;
		const PTask tk = load_const(tasks, tl.task);

		acc_t acc[FI][FJ];
#pragma unroll
		for (int i = 0; i < FI; ++i)
#pragma unroll
			for (int j = 0; j < FJ; ++j)
				acc[i][j] = acc_t {};

		int seg = tk.seg_begin;
		while (seg < tk.seg_end && load_const(segs, seg).k <= 0)
			++seg;
		int kstart = 0;
		while (seg < tk.seg_end) {
			const int nc = (load_const(segs, seg).k + BK - 1) / BK;
			if (skip < nc) {
				kstart = skip * BK;
				break;
			}
			skip -= nc;
			do
				++seg;
			while (seg < tk.seg_end && load_const(segs, seg).k <= 0);
		}

		if (seg < tk.seg_end && nch_item != 0) {
			LA   lah, lal;
			LB   lbh, lbl;
			PSeg s      = load_const(segs, seg);
			int  ck0    = kstart;
			int  lastk  = kstart;
			bool cvalid = true;
			auto setup  = [&](const PSeg& sg) {
				lah.setup(pick(hi, sg.lbase) + sg.loff, sg.ldl, tl.m0, tk.m, lane, wave);
				lal.setup(pick(lo, sg.lbase) + sg.loff, sg.ldl, tl.m0, tk.m, lane, wave);
				lbh.setup(pick(hi, sg.rbase) + sg.roff, sg.ldr, tl.n0, tk.n, lane, wave);
				lbl.setup(pick(lo, sg.rbase) + sg.roff, sg.ldr, tl.n0, tk.n, lane, wave);
			};
			// LDS-DMA of the cursor's chunk into stage st (past the last chunk the last one
			// again, into a stage that is not read any more), then the cursor moves on.
			auto issue = [&](Stage& st) {
				const int kk = cvalid ? ck0 : lastk;
				lah.dma(st.ah, kk, wave);
				lal.dma(st.al, kk, wave);
				lbh.dma(st.bh, kk, wave);
				lbl.dma(st.bl, kk, wave);
				lastk = kk;
				if (cvalid) {
					bool newseg;
					cvalid = next_chunk<BK>(segs, tk.seg_end, s.k, seg, ck0, newseg);
					if (cvalid && newseg) {
						s = load_const(segs, seg);
						setup(s);
					}
				}
			};
			auto compute = [&](const Stage& st) {
#pragma unroll
				for (int stp = 0; stp < NST; ++stp) {
					// Step stp covers k in [stp * kK, stp * kK + kK): lane group g the 8 at 8 g.
					const int kv = stp * NG + g;
					bf8_t     ah[FI], al[FI], bh[FJ], bl[FJ];
#pragma unroll
					for (int i = 0; i < FI; ++i) {
						const int r = wm + Op::kM * i + lr;
						ah[i]       = st.ah[r * VPR + SA::slot(r, kv)];
						al[i]       = st.al[r * VPR + SA::slot(r, kv)];
					}
#pragma unroll
					for (int j = 0; j < FJ; ++j) {
						const int r = wn + Op::kM * j + lr;
						bh[j]       = st.bh[r * VPR + SB::slot(r, kv)];
						bl[j]       = st.bl[r * VPR + SB::slot(r, kv)];
					}
#pragma unroll
					for (int i = 0; i < FI; ++i)
#pragma unroll
						for (int j = 0; j < FJ; ++j) {
							acc[i][j] = Op::mma(al[i], bh[j], acc[i][j]);
							acc[i][j] = Op::mma(ah[i], bl[j], acc[i][j]);
							acc[i][j] = Op::mma(ah[i], bh[j], acc[i][j]);
						}
				}
			};
			setup(s);
			const int nch =
			    nch_item > 0 ? nch_item : chunk_count<BK>(segs, seg, tk.seg_end) - kstart / BK;
			if constexpr (NS == 2) {
				issue(S0);
				wait_vm_lgkm0<0>();
				barrier_no_fence();
				int c = 0;
				for (; c + 2 <= nch; c += 2) {
					issue(S1);
					compute(S0);
					wait_vm_lgkm0<0>();
					barrier_no_fence();
					if (c + 2 == nch) {
						compute(S1);
						break;
					}
					issue(S0);
					compute(S1);
					wait_vm_lgkm0<0>();
					barrier_no_fence();
				}
				if (c + 1 == nch)
					compute(S0);
			} else {
				issue(S0);
				issue(S1);
				wait_vm_lgkm0<NW>();
				barrier_no_fence();
				int c = 0;
				for (; c + 3 <= nch; c += 3) {
					issue(S2);
					compute(S0);
					wait_vm_lgkm0<NW>();
					barrier_no_fence();
					issue(S0);
					compute(S1);
					wait_vm_lgkm0<NW>();
					barrier_no_fence();
					issue(S1);
					compute(S2);
					wait_vm_lgkm0<NW>();
					barrier_no_fence();
				}
				if (c < nch) {
					compute(S0);
					if (c + 1 < nch) {
						wait_vm_lgkm0<0>();
						barrier_no_fence();
						compute(S1);
					}
				}
			}
			wait_vm_lgkm0<0>();
		}

		if (slot >= 0) {
			// A partial of a tile cut between workgroups (FP32, tile-local).
			float* part = out_f + int64_t(slot) * (BM * BN);
#pragma unroll
			for (int i = 0; i < FI; ++i)
#pragma unroll
				for (int j = 0; j < FJ; ++j)
#pragma unroll
					for (int r = 0; r < Op::kRegs; ++r)
						part[(wm + Op::kM * i + Op::row(lane, r)) * BN + wn + Op::kM * j + lr] =
						    acc[i][j][r];
		} else {
			const bool to_y = tk.out_is_y != 0;
#pragma unroll
			for (int i = 0; i < FI; ++i)
#pragma unroll
				for (int j = 0; j < FJ; ++j)
#pragma unroll
					for (int r = 0; r < Op::kRegs; ++r) {
						const int row = tl.m0 + wm + Op::kM * i + Op::row(lane, r);
						const int col = tl.n0 + wn + Op::kM * j + lr;
						if (row < tk.m && col < tk.n) {
							const int64_t addr = out_addr(tk, row, col);
							const float   v    = acc[i][j][r];
							if (to_y) {
								double y = double(v);
								if (beta)
									y += out_y[addr];
								out_y[addr] = y;
							} else if (PAIR) {
								const __bf16 h = __bf16(v);
								out_h[addr]    = h;
								out_l[addr]    = __bf16(v - float(h));
							} else {
								out_f[addr] = v;
							}
						}
					}
		}
		__syncthreads(); // LDS buffer 0 is reused by the next item
	}
}

// Fix-up of a balanced launch: the partials of each tile cut between workgroups, summed in slot
// order (the workgroups' order, so the result does not depend on timing) in FP64 into the
// task's output (the caller's vector; beta adds to it). One workgroup per cut tile and vector.
template <typename T>
__global__ __launch_bounds__(256) void sk_fixup_kernel(const SkFix* __restrict__ fixes,
                                                       const PTask* __restrict__ tasks,
                                                       const T* __restrict__ part,
                                                       int64_t wsstride, double* __restrict__ out_y,
                                                       int64_t ystride, int bm, int bn, int beta)
{
	const SkFix   f   = fixes[blockIdx.x];
	const PTask   tk  = tasks[f.task];
	const int64_t tsz = int64_t(bm) * bn;
	part += int64_t(blockIdx.y) * wsstride + int64_t(f.first) * tsz;
	out_y += int64_t(blockIdx.y) * ystride;
	for (int e = threadIdx.x; e < bm * bn; e += blockDim.x) {
		const int row = f.m0 + e / bn;
		const int col = f.n0 + e % bn;
		if (row >= tk.m || col >= tk.n)
			continue;
		double v = 0.0;
		for (int q = 0; q < f.count; ++q)
			v += double(part[q * tsz + e]);
		const int64_t addr = out_addr(tk, row, col);
		out_y[addr]        = beta ? v + out_y[addr] : v;
	}
}

// Input preparation, one 64 x 64 tile of one patch per workgroup: X_J (R x L, column major at
// xoff) is written column major with padded columns (XP: element (m, c) at pofs + m + c * ldp,
// zero for R <= m < ldp) and transposed (XT: element (m, c) at tofs + c + m * ldt, zero for
// L <= c < ldt), converted to the compute type.
template <typename E> struct PrepOut {
	E* xp;
	E* xt;
};

__device__ inline void put(double* p, int64_t i, double v) { p[i] = v; }
__device__ inline void put(float* p, int64_t i, double v) { p[i] = float(v); }

template <typename E>
__global__ __launch_bounds__(256) void prep_x_kernel(const PrepTile* __restrict__ tiles,
                                                     const PrepPatch* __restrict__ patches,
                                                     const double* __restrict__ x, int64_t ldx,
                                                     E* __restrict__ xp, E* __restrict__ xt,
                                                     PStrides vs)
{
	__shared__ double t[64][65];
	const PrepTile  pt = tiles[blockIdx.x];
	const PrepPatch pp = patches[pt.patch];
	const int       v  = blockIdx.y;
	x += v * ldx + pp.xoff;
	xp += v * vs.xp + pp.pofs;
	xt += v * vs.xt + pp.tofs;
	const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
	// Column-major read of the tile (m = m0 + tx fastest), all 16 loads of a thread in flight
	// before the first store; XP written the same way.
	const int m = pt.m0 + tx;
	double    val[16];
#pragma unroll
	for (int q = 0; q < 16; ++q) {
		const int col = pt.c0 + ty + 4 * q;
		val[q]        = (m < pp.R && col < pp.L) ? x[m + int64_t(col) * pp.R] : 0.0;
	}
#pragma unroll
	for (int q = 0; q < 16; ++q) {
		const int c   = ty + 4 * q;
		const int col = pt.c0 + c;
		t[c][tx]      = val[q];
		if (m < pp.ldp && col < pp.L)
			put(xp, m + int64_t(col) * pp.ldp, val[q]);
	}
	__syncthreads();
	// XT: c fastest.
	for (int r = ty; r < 64; r += 4) {
		const int m   = pt.m0 + r;
		const int col = pt.c0 + tx;
		if (m < pp.R && col < pp.ldt)
			put(xt, col + int64_t(m) * pp.ldt, t[tx][r]);
	}
}

// The same into (hi, lo) bf16 pairs: hi = bf16(v), lo = bf16(v - hi), the residual in double.
__global__ __launch_bounds__(256) void prep_x_bf16x2_kernel(
    const PrepTile* __restrict__ tiles, const PrepPatch* __restrict__ patches,
    const double* __restrict__ x, int64_t ldx, __bf16* __restrict__ xph, __bf16* __restrict__ xpl,
    __bf16* __restrict__ xth, __bf16* __restrict__ xtl, PStrides vs)
{
	__shared__ double t[64][65];
	const PrepTile  pt = tiles[blockIdx.x];
	const PrepPatch pp = patches[pt.patch];
	const int       v  = blockIdx.y;
	x += v * ldx + pp.xoff;
	const int64_t po = v * vs.xp + pp.pofs, to = v * vs.xt + pp.tofs;
	const int     tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
	const int     m  = pt.m0 + tx;
	double        val[16]; // all 16 loads of a thread in flight before the first store
#pragma unroll
	for (int q = 0; q < 16; ++q) {
		const int col = pt.c0 + ty + 4 * q;
		val[q]        = (m < pp.R && col < pp.L) ? x[m + int64_t(col) * pp.R] : 0.0;
	}
#pragma unroll
	for (int q = 0; q < 16; ++q) {
		const int c   = ty + 4 * q;
		const int col = pt.c0 + c;
		t[c][tx]      = val[q];
		if (m < pp.ldp && col < pp.L) {
			const __bf16  h = __bf16(float(val[q]));
			const int64_t i = po + m + int64_t(col) * pp.ldp;
			xph[i]          = h;
			xpl[i]          = __bf16(float(val[q] - double(float(h))));
		}
	}
	__syncthreads();
	for (int r = ty; r < 64; r += 4) {
		const int m   = pt.m0 + r;
		const int col = pt.c0 + tx;
		if (m < pp.R && col < pp.ldt) {
			const double  val = t[tx][r];
			const __bf16  h   = __bf16(float(val));
			const int64_t i   = to + col + int64_t(m) * pp.ldt;
			xth[i]            = h;
			xtl[i]            = __bf16(float(val - double(float(h))));
		}
	}
}

// hi = bf16(x), lo = bf16(x - hi), the residual taken in double (elementwise; the packed
// factors keep their layout).
__global__ void split_f64_bf16x2_kernel(const double* __restrict__ x, __bf16* __restrict__ hi,
                                        __bf16* __restrict__ lo, int64_t n)
{
	int64_t stride = int64_t(gridDim.x) * blockDim.x;
	for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
		const double v = x[i];
		const __bf16 h = __bf16(float(v));
		hi[i]          = h;
		lo[i]          = __bf16(float(v - double(float(h))));
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
	                   static_cast<__bf16*>(hi), static_cast<__bf16*>(lo), n);
}

// Host-side launchers -------------------------------------------------------------------------

// depth 1: two LDS stages (one chunk in flight), depth 2: three (two chunks in flight). kchunk 2
// doubles the K chunk (FP64 32, FP32 64 elements per stage: twice the MFMAs per barrier).
template <typename T, typename TO>
void launch_panel_gemm(int bm, int bn, int variant, int depth, int kchunk, const PTask* tasks,
                       const PSeg* segs, const Tile* tiles, int ntiles, const T* const* bases,
                       const PStrides& vs, TO* out_y, int64_t ystride, T* out_ws,
                       int64_t wsstride, int nvec, int beta, hipStream_t stream,
                       const PanelLaunchExtra& ex)
{
	if (!ex.occupancy && (ntiles <= 0 || nvec <= 0))
		return;
	PBasesT<T> b;
	for (int i = 0; i < PBASE_COUNT; ++i)
		b.p[i] = bases ? bases[i] : nullptr;
	const dim3 grid { unsigned(std::max(ntiles, 1)), unsigned(std::max(nvec, 1)), 1u };
	auto go = [&](auto kern) {
		if (ex.occupancy) {
			DMK_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(ex.occupancy, kern,
			                                                           kPThreads, 0));
			return;
		}
		hipLaunchKernelGGL(kern, grid, dim3(kPThreads), 0, stream, tasks, segs, tiles, ex.items,
		                   ex.bins, b, vs, out_y, ystride, out_ws, wsstride, beta);
	};
#define DMK_PLAUNCH(BM, BN, BK, OP)                                                                \
	do {                                                                                       \
		if (depth == 2)                                                                    \
			go(panel_gemm_kernel<TO, BM, BN, BK, 3, OP>);                             \
		else                                                                               \
			go(panel_gemm_kernel<TO, BM, BN, BK, 2, OP>);                             \
	} while (0)
#define DMK_PSHAPES(BK, OP)                                                                        \
	if (bm == 64 && bn == 64)                                                                  \
		DMK_PLAUNCH(64, 64, BK, OP);                                                       \
	else if (bm == 128 && bn == 64)                                                            \
		DMK_PLAUNCH(128, 64, BK, OP);                                                      \
	else if (bm == 64 && bn == 128)                                                            \
		DMK_PLAUNCH(64, 128, BK, OP);                                                      \
	else if (bm == 128 && bn == 128)                                                           \
		DMK_PLAUNCH(128, 128, BK, OP);                                                     \
	else                                                                                       \
		throw std::runtime_error("panel gemm: unsupported tile shape");
	// 96-wide tiles (48-wide wave tiles, so 16x16 MFMA blocks in either MFMA setting): they fit
	// patch sizes just above a multiple of 64, such as 1040, with less edge waste.
#define DMK_PSHAPES96(BK, OP)                                                                      \
	if (bm == 96 && bn == 96)                                                                  \
		DMK_PLAUNCH(96, 96, BK, OP);                                                       \
	else if (bm == 96 && bn == 64)                                                             \
		DMK_PLAUNCH(96, 64, BK, OP);                                                       \
	else if (bm == 64 && bn == 96)                                                             \
		DMK_PLAUNCH(64, 96, BK, OP);                                                       \
	else                                                                                       \
		throw std::runtime_error("panel gemm: unsupported tile shape");
	// 32-wide tiles (16-wide wave tiles, 16x16 MFMA blocks): small patches, such as the comb
	// shape's patches of 1-30 states, and stage-1 tasks whose row count is a few R_I.
#define DMK_PSHAPES32(BK, OP)                                                                      \
	if (bm == 64 && bn == 32)                                                                  \
		DMK_PLAUNCH(64, 32, BK, OP);                                                       \
	else if (bm == 32 && bn == 64)                                                             \
		DMK_PLAUNCH(32, 64, BK, OP);                                                       \
	else if (bm == 32 && bn == 32)                                                             \
		DMK_PLAUNCH(32, 32, BK, OP);                                                       \
	else if (bm == 128 && bn == 32)                                                            \
		DMK_PLAUNCH(128, 32, BK, OP);                                                      \
	else if (bm == 32 && bn == 128)                                                            \
		DMK_PLAUNCH(32, 128, BK, OP);                                                      \
	else                                                                                       \
		throw std::runtime_error("panel gemm: unsupported tile shape");
	// Double-depth K chunks: tiles of at most 64 + 64 rows.
#define DMK_PSHAPES_LONG(BK, OP)                                                                   \
	if (bm == 64 && bn == 64)                                                                  \
		DMK_PLAUNCH(64, 64, BK, OP);                                                       \
	else if (bm == 64 && bn == 32)                                                             \
		DMK_PLAUNCH(64, 32, BK, OP);                                                       \
	else if (bm == 32 && bn == 64)                                                             \
		DMK_PLAUNCH(32, 64, BK, OP);                                                       \
	else if (bm == 32 && bn == 32)                                                             \
		DMK_PLAUNCH(32, 32, BK, OP);                                                       \
	else                                                                                       \
		throw std::runtime_error("panel gemm: unsupported tile shape for double K chunks");
	const bool s96 = bm == 96 || bn == 96;
	const bool s32 = bm == 32 || bn == 32;
	if (kchunk == 2) {
		if constexpr (std::is_same<T, double>::value) {
			DMK_PSHAPES_LONG(32, OpF64)
		} else {
			if (variant == 1 && bm == 64 && bn == 64) {
				DMK_PLAUNCH(64, 64, 64, OpF32x32);
			} else {
				DMK_PSHAPES_LONG(64, OpF32)
			}
		}
		return;
	}
	if (kchunk != 1)
		throw std::runtime_error("panel gemm: kchunk must be 1 or 2");
	if constexpr (std::is_same<T, double>::value) {
		if (s96) {
			DMK_PSHAPES96(16, OpF64)
		} else if (s32) {
			DMK_PSHAPES32(16, OpF64)
		} else {
			DMK_PSHAPES(16, OpF64)
		}
	} else {
		if (s96) {
			DMK_PSHAPES96(32, OpF32)
		} else if (s32) {
			DMK_PSHAPES32(32, OpF32)
		} else if (variant == 1) {
			DMK_PSHAPES(32, OpF32x32)
		} else {
			DMK_PSHAPES(32, OpF32)
		}
	}
#undef DMK_PSHAPES_LONG
#undef DMK_PSHAPES32
#undef DMK_PSHAPES96
#undef DMK_PSHAPES
#undef DMK_PLAUNCH
}

template void launch_panel_gemm<double, double>(int, int, int, int, int, const PTask*, const PSeg*,
                                                const Tile*, int, const double* const*,
                                                const PStrides&, double*, int64_t, double*,
                                                int64_t, int, int, hipStream_t,
                                                const PanelLaunchExtra&);
template void launch_panel_gemm<float, float>(int, int, int, int, int, const PTask*, const PSeg*,
                                              const Tile*, int, const float* const*,
                                              const PStrides&, float*, int64_t, float*, int64_t,
                                              int, int, hipStream_t,
                                                const PanelLaunchExtra&);
template void launch_panel_gemm<float, double>(int, int, int, int, int, const PTask*, const PSeg*,
                                               const Tile*, int, const float* const*,
                                               const PStrides&, double*, int64_t, float*, int64_t,
                                               int, int, hipStream_t,
                                                const PanelLaunchExtra&);

// variant 1 selects the 32x32x16 bf16 MFMA form; depth 1: two LDS stages (one chunk in
// flight), depth 2: three (two chunks in flight).
void launch_panel_bf16x3(int bm, int bn, int variant, int depth, const PTask* tasks,
                         const PSeg* segs,
                         const Tile* tiles, int ntiles, const void* const* bases_hi,
                         const void* const* bases_lo, const PStrides& vs, double* out_y,
                         int64_t ystride, float* out_f, void* out_h, void* out_l, int64_t wsstride,
                         bool pair, int nvec, int beta, hipStream_t stream,
                         const PanelLaunchExtra& ex)
{
	if (!ex.occupancy && (ntiles <= 0 || nvec <= 0))
		return;
	PBasesT<__bf16> h, l;
	for (int i = 0; i < PBASE_COUNT; ++i) {
		h.p[i] = bases_hi ? static_cast<const __bf16*>(bases_hi[i]) : nullptr;
		l.p[i] = bases_lo ? static_cast<const __bf16*>(bases_lo[i]) : nullptr;
	}
	__bf16*    oh = static_cast<__bf16*>(out_h);
	__bf16*    ol = static_cast<__bf16*>(out_l);
	const dim3 grid { unsigned(std::max(ntiles, 1)), unsigned(std::max(nvec, 1)), 1u };
	auto go = [&](auto kern) {
		if (ex.occupancy) {
			DMK_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(ex.occupancy, kern,
			                                                           kPThreads, 0));
			return;
		}
		hipLaunchKernelGGL(kern, grid, dim3(kPThreads), 0, stream, tasks, segs, tiles, ex.items,
		                   ex.bins, h, l, vs, out_y, ystride, out_f, oh, ol, wsstride, beta);
	};
#define DMK_BLAUNCH_PF(BM, BN, NS, PAIR, OP) go(panel_bf16x3_kernel<BM, BN, 32, NS, PAIR, OP>)
#define DMK_BLAUNCH(BM, BN, PAIR, OP)                                                              \
	do {                                                                                       \
		if (depth == 2)                                                                    \
			DMK_BLAUNCH_PF(BM, BN, 3, PAIR, OP);                                       \
		else                                                                               \
			DMK_BLAUNCH_PF(BM, BN, 2, PAIR, OP);                                       \
	} while (0)
#define DMK_BSHAPES(PAIR, OP)                                                                      \
	if (bm == 64 && bn == 64)                                                                  \
		DMK_BLAUNCH(64, 64, PAIR, OP);                                                     \
	else if (bm == 128 && bn == 64)                                                            \
		DMK_BLAUNCH(128, 64, PAIR, OP);                                                    \
	else if (bm == 64 && bn == 128)                                                            \
		DMK_BLAUNCH(64, 128, PAIR, OP);                                                    \
	else if (bm == 128 && bn == 128)                                                           \
		DMK_BLAUNCH(128, 128, PAIR, OP);                                                   \
	else                                                                                       \
		throw std::runtime_error("panel gemm bf16x3: unsupported tile shape");
	if (pair) {
		if (variant == 1) {
			DMK_BSHAPES(true, OpBf16x32)
		} else {
			DMK_BSHAPES(true, OpBf16)
		}
	} else {
		if (variant == 1) {
			DMK_BSHAPES(false, OpBf16x32)
		} else {
			DMK_BSHAPES(false, OpBf16)
		}
	}
#undef DMK_BSHAPES
#undef DMK_BLAUNCH
#undef DMK_BLAUNCH_PF
}

template <typename T>
void launch_sk_fixup(const SkFix* fixes, int nfix, const PTask* tasks, const T* part,
                     int64_t wsstride, double* out_y, int64_t ystride, int bm, int bn, int nvec,
                     int beta, hipStream_t stream)
{
	if (nfix <= 0 || nvec <= 0)
		return;
	hipLaunchKernelGGL((sk_fixup_kernel<T>), dim3(unsigned(nfix), unsigned(nvec)), dim3(256), 0,
	                   stream, fixes, tasks, part, wsstride, out_y, ystride, bm, bn, beta);
}
template void launch_sk_fixup<double>(const SkFix*, int, const PTask*, const double*, int64_t,
                                      double*, int64_t, int, int, int, int, hipStream_t);
template void launch_sk_fixup<float>(const SkFix*, int, const PTask*, const float*, int64_t,
                                     double*, int64_t, int, int, int, int, hipStream_t);

template <typename E>
void launch_prep_x(const PrepTile* tiles, int ntiles, const PrepPatch* patches, const double* x,
                   int64_t ldx, E* xp, E* xt, const PStrides& vs, int nvec, hipStream_t stream)
{
	if (ntiles <= 0 || nvec <= 0)
		return;
	hipLaunchKernelGGL((prep_x_kernel<E>), dim3(unsigned(ntiles), unsigned(nvec)), dim3(256), 0,
	                   stream, tiles, patches, x, ldx, xp, xt, vs);
}
template void launch_prep_x<double>(const PrepTile*, int, const PrepPatch*, const double*, int64_t,
                                    double*, double*, const PStrides&, int, hipStream_t);
template void launch_prep_x<float>(const PrepTile*, int, const PrepPatch*, const double*, int64_t,
                                   float*, float*, const PStrides&, int, hipStream_t);

void launch_prep_x_bf16x2(const PrepTile* tiles, int ntiles, const PrepPatch* patches,
                          const double* x, int64_t ldx, void* xph, void* xpl, void* xth, void* xtl,
                          const PStrides& vs, int nvec, hipStream_t stream)
{
	if (ntiles <= 0 || nvec <= 0)
		return;
	hipLaunchKernelGGL(prep_x_bf16x2_kernel, dim3(unsigned(ntiles), unsigned(nvec)), dim3(256), 0,
	                   stream, tiles, patches, x, ldx, static_cast<__bf16*>(xph),
	                   static_cast<__bf16*>(xpl), static_cast<__bf16*>(xth),
	                   static_cast<__bf16*>(xtl), vs);
}

} // namespace dmk
