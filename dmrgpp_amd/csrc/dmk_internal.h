// Internal descriptors shared by the host planner (kron_plan.cpp) and the HIP kernels
// (panel_kernels.hip, kron_kernels.hip, initkron_kernels.hip). Not part of the public C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

namespace dmk {

// One fixed-order reduction of split-k partials into the output: for e < count,
//   out[yoff + e] (+)= sum_{s < nsplit} ws[wsoff + s * count + e].
struct Reduce {
	int64_t yoff;
	int64_t wsoff;
	int64_t count;
	int32_t nsplit;
	int32_t pad;
};

// One workgroup of the reduction launch: elements [start, start + len) of reduce entry r,
// len <= kReduceChunk.
constexpr int kReduceChunk = 2048;
struct ReduceChunk {
	int32_t r;
	int32_t len;
	int64_t start;
};

// One output tile of a task; the tile list is ordered by decreasing task cost.
struct Tile {
	int32_t task;
	int32_t m0;
	int32_t n0;
	int32_t pad;
};

// Balanced persistent launch (a grid of as many workgroups as the chip holds at once, each
// given an equal share of the launch's K chunks): workgroup w runs items [bins[w], bins[w+1]).
// An item is chunks [c0, c0 + nch) of tile (task, m0, n0); its result goes to the task's output
// (slot < 0: the item is the whole tile) or to partial slot `slot` (a BM x BN tile of the compute
// type, element (r, c) at r * BN + c), which the fix-up kernel sums for the tile.
struct SkItem {
	int32_t task;
	int32_t m0;
	int32_t n0;
	int32_t c0;
	int32_t nch;
	int32_t slot;
};
// A tile cut between workgroups: its partials are slots [first, first + count), summed in that
// order (the workgroup order), so the result does not depend on timing.
struct SkFix {
	int32_t task;
	int32_t m0;
	int32_t n0;
	int32_t first;
	int32_t count;
	int32_t pad;
};

// Extra arguments of a panel-kernel launch: the item list and the workgroups' item ranges of a
// balanced persistent launch (nullptr: one tile per workgroup); occupancy, when non-null,
// receives the kernel's resident workgroups per CU instead of a launch.
struct PanelLaunchExtra {
	const SkItem*  items     = nullptr;
	const int32_t* bins      = nullptr;
	int*           occupancy = nullptr;
};

// Descriptors of the matvec kernels on K-major panels (panel_kernels.hip). Operand bases:
enum PBaseId : int32_t
{
	PBASE_APACK = 0, // packed left-block factors, K-major
	PBASE_BPACK = 1, // packed right-block factors, K-major
	PBASE_XP    = 2, // input patches X_J column major with padded columns (K-major as a right
	                 // operand over rows of X_J, as a left operand over columns)
	PBASE_XT    = 3, // input patches transposed, padded rows
	PBASE_BX    = 4, // stage-1 products, K-major stage-2 panels
	PBASE_COUNT = 5
};

// One contribution C(MxN) += L(MxK) * R(KxN): element (m, k) of L at loff + k + m * ldl,
// element (k, n) of R at roff + k + n * ldr. Row starts and loff/roff are multiples of 8
// elements and rows are zero padded to a multiple of 8 past k.
struct PSeg {
	int64_t loff;
	int64_t roff;
	int32_t lbase;
	int32_t rbase;
	int32_t ldl;
	int32_t ldr;
	int32_t k;
	int32_t pad;
};

// One output matrix: element (r, c) at coff + (r % rsplit) * sm + (r / rsplit) * n + c.
struct PTask {
	int64_t coff;
	int32_t m;
	int32_t n;
	int32_t rsplit;
	int32_t sm;
	int32_t seg_begin;
	int32_t seg_end;
	int32_t out_is_y; // 1: the caller's output vector, 0: the workspace (compute type)
	int32_t pad;
};

// Per-vector element strides of the per-call operands in multi-vector launches.
struct PStrides {
	int64_t xp;
	int64_t xt;
	int64_t bx;
};

// Input preparation: a 64 x 64 tile (m0, c0) of patch `patch`.
struct PrepTile {
	int32_t patch;
	int32_t m0;
	int32_t c0;
	int32_t pad;
};
struct PrepPatch {
	int64_t xoff; // X_J in the input vector (R x L column major)
	int64_t pofs; // XP_J: element (m, c) at pofs + m + c * ldp
	int64_t tofs; // XT_J: element (m, c) at tofs + c + m * ldt
	int32_t R;
	int32_t L;
	int32_t ldp;
	int32_t ldt;
};

// Device-side packing of the factors (initkron_kernels.hip): block (rows x cols, column major at
// src with leading dimension ldsrc: in the uploaded copy of host blocks, or the caller's
// device-resident block itself) written as panel rows at dst with row stride lddst, then `tail`
// (< 64) zeros after each row.
struct PackJob {
	const double* src;
	int64_t       ldsrc;
	int64_t       dst;
	int32_t       rows;
	int32_t       cols;
	int32_t       lddst;
	int32_t       tail;
	int32_t       to_b; // 0: A pack, 1: B pack
	int32_t       pad;
};
struct PackTile {
	int32_t job;
	int32_t i0;
	int32_t j0;
	int32_t pad;
};

// Scan of a device-resident block (64 x 64 tiles listed as PackTile with job = block index):
// flags[2 d] becomes nonzero when block d holds a nonzero element, flags[2 d + 1] when it departs
// from the identity pattern (a diagonal element other than 1, an off-diagonal one other than 0).
struct ScanBlock {
	const double* p;
	int64_t       ld;
	int32_t       rows;
	int32_t       cols;
};

// Gather of `count` diagonal elements p[t * (ld + 1)] to out + off.
struct DiagJob {
	const double* p;
	int64_t       ld;
	int64_t       off;
	int64_t       count;
};

// Rows of K-major panels are zero padded to a multiple of this many elements, and split-k pieces
// start on multiples of it: every K chunk of a segment (at most 64 elements, from a multiple of
// its size) then lies inside the segment's rows, which the LDS-DMA loads of the FP64 / FP32
// panel kernel rely on (they have no masks).
constexpr int kPanelPad = 64;
inline int64_t panel_ld(int64_t k) { return (k + kPanelPad - 1) / kPanelPad * kPanelPad; }
// The panel kernels address the rows of one tile (at most 128) by 32-bit byte offsets from the
// tile's first row: leading dimensions up to 2^21 elements keep them below 2^31.
inline bool panel_ld_fits(int64_t ld) { return ld >= 0 && ld <= (int64_t(1) << 21); }

// Error reporting for the C ABI.
void        set_last_error(const std::string& msg);
const char* last_error();

} // namespace dmk

#define DMK_HIP_CHECK(expr)                                                                        \
	do {                                                                                       \
		hipError_t _e = (expr);                                                            \
		if (_e != hipSuccess) {                                                            \
			throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(_e)); \
		}                                                                                  \
	} while (0)
