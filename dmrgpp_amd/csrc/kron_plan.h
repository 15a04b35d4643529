// Host-side plan of one superblock Kron matvec: the packed factors resident in HBM, the task /
// segment / tile tables of the two grouped-GEMM stages, and the per-handle stream.
#pragma once

#include "dmk_internal.h"

#include <cstdint>
#include <string>
#include <vector>

namespace dmk {

enum Precision : int
{
	PREC_FP64   = 0,
	PREC_FP32   = 1,
	PREC_BF16X3 = 2,
};

struct PlanStats {
	int64_t n                 = 0; // superblock dimension (vector length)
	int64_t npatches          = 0;
	int64_t noperator         = 0;
	int64_t nterms            = 0; // nonzero (I,J,k) pairs (setup_nC total)
	int64_t nterms_identity   = 0; // of which one factor is an identity
	int64_t stage1_tasks      = 0;
	int64_t stage2_tasks      = 0;
	int64_t stage1_tiles      = 0;
	int64_t stage2_tiles      = 0;
	int64_t split_tasks       = 0; // stage-2 tasks cut into split-k pieces
	double  flops_algorithmic = 0; // sum over terms of min(method 1, method 2), identities as one GEMM
	double  flops_plugin      = 0; // the reference plugin's gflops1+gflops2 count (all terms dense)
	double  flops_executed    = 0; // what the two stages execute (before tile padding)
	double  flops_stage1      = 0; // executed by the stage-1 launch
	double  flops_stage2      = 0; // executed by the stage-2 launch
	double  bytes_compulsory  = 0; // non-identity factors once + x read + y write, FP64 bytes
	double  bytes_bx          = 0; // stage-1 scratch written and read back, FP64 bytes
	int64_t apack_elems       = 0;
	int64_t bpack_elems       = 0;
	int64_t bx_elems          = 0;
	int64_t ws_elems          = 0; // split-k partials
	int64_t max_patch_left    = 0;
	int64_t max_patch_right   = 0;
	// Plan construction (host wall time): block scan, staging of the uploaded blocks, upload
	// plus device packing; bytes uploaded.
	double ms_scan   = 0;
	double ms_stage  = 0;
	double ms_upload = 0;
	double raw_bytes = 0;
};

// Launch shape of a plan: output tile per stage, split-k granularity of stage 2 (target
// contraction length per piece; 0 disables), and the tile order.
struct Tuning {
	int bm1    = 64;
	int bn1    = 64;
	int bm2    = 64;
	int bn2    = 64;
	int splitk = 1024;
	int xcd    = 0; // 1: XCD-aware tile order
	int mfma32 = 0; // 1: 32x32 MFMA blocks (FP32 32x32x2, BF16x3 32x32x16); FP64 ignores it
	int depth  = 1; // global loads issued this many K chunks ahead (1 or 2)
	int kchunk = 1; // K chunk depth in units of the mode's base chunk (1 or 2)
	int persist = 0; // k > 0: stage 2 as a balanced persistent launch, k workgroups per CU at most

	bool operator==(const Tuning& o) const
	{
		return bm1 == o.bm1 && bn1 == o.bn1 && bm2 == o.bm2 && bn2 == o.bn2 && splitk == o.splitk
		       && xcd == o.xcd && mfma32 == o.mfma32 && depth == o.depth && kchunk == o.kchunk
		       && persist == o.persist;
	}
};

struct StageTables {
	int                bm = 64; // output tile shape of this stage's launch
	int                bn = 64;
	std::vector<PTask> tasks;
	std::vector<PSeg>  segs;
	std::vector<Tile>  tiles;
	PTask*             d_tasks = nullptr;
	PSeg*              d_segs  = nullptr;
	Tile*              d_tiles = nullptr;
	void               upload();
	void               release();
};

// The starting tuning of a precision mode (before any autotuning).
Tuning default_tuning(int precision);

// Whether the present blocks (both factors non-null) live in device or managed memory: true when
// all do, false when all are host memory; throws std::invalid_argument when they are mixed.
bool blocks_on_device(int64_t noperator, int64_t npatches, const double* const* Amatrix,
                      const double* const* Bmatrix);

// Process-wide autotuner cache (histogram signature -> tuning).
bool cached_tuning(const std::string& key, Tuning& t);
void store_tuning(const std::string& key, const Tuning& t);

class Plan {
public:
	// Same inputs as BatchedGemm<double>::BatchedGemm (GPUPlugin/BatchedGemm.h:21-28). The
	// blocks are all host memory (scanned on the host, staged and uploaded) or all device /
	// managed memory (scanned and packed on the device, read in place; no upload).
	Plan(int64_t noperator, int64_t npatches, const size_t* left_patch_size,
	     const size_t* right_patch_size, const double* const* Amatrix, const size_t* ld_Amatrix,
	     const double* const* Bmatrix, const size_t* ld_Bmatrix);
	~Plan();

	Plan(const Plan&)            = delete;
	Plan& operator=(const Plan&) = delete;

	// vout = H vin (beta = 0) or vout += H vin (beta = 1) for nvec vectors; device pointers.
	void apply(const double* vin, int64_t ldin, double* vout, int64_t ldout, int nvec, int beta);

	// Scratch of one single-vector matvec in the current precision, as three 256-byte aligned
	// regions of one buffer: stage-1 products at 0, split-k partials at *off_ws, input panels
	// at *off_xin. Returns the total size in bytes.
	size_t scratch_layout(size_t* off_ws, size_t* off_xin) const;

	// vout = H vin (one vector, beta = 0) with every scratch buffer carved from the caller's
	// device memory at `scratch` (`bytes` bytes, 256-byte aligned, at least scratch_layout()):
	// no allocation, no synchronisation, no hipMemsetAsync (the stage-1 region's zero padding is
	// restored by a fill kernel on every call). Enqueued on the plan's stream.
	void apply_on_scratch(const double* vin, double* vout, void* scratch, size_t bytes);

	void set_precision(int p);
	int  precision() const { return precision_; }

	// Rebuilds the launch tables for a new tuning (tile shapes, split-k, order). The tuning
	// belongs to the current precision mode: set_precision switches to the tuning last set
	// (or autotuned, or cached, or default) for the mode it selects.
	void          retune(const Tuning& t);
	const Tuning& tuning() const { return tuning_; }

	// Times candidate tunings on this problem (iters matvecs each, current precision) and keeps
	// the fastest; the choice is cached per block-size histogram signature for later plans.
	Tuning autotune(const double* vin, double* vout, int iters);

	// Signature of the block-size histogram (log2 bins of the stage-2 task shapes).
	std::string histogram_key() const;
	// The autotuner cache key: the histogram signature and the precision mode.
	std::string tuning_key() const;

	const PlanStats& stats() const { return stats_; }
	hipStream_t      stream() const { return stream_; }
	int64_t          n() const { return stats_.n; }
	bool             device_blocks() const { return device_blocks_; }

	// diag(H) into out (device, n): sum over the diagonal-patch terms (I, I, k) of
	// diag(A_k(I,I)) (x) diag(B_k(I,I)). Enqueued on the plan's stream.
	void diagonal(double* out);

	// Average device time (ms) of each stage's kernel and of the whole matvec over iters calls,
	// from HIP events recorded on the plan's stream around every launch.
	void profile(const double* vin, double* vout, int iters, double* ms_stage1, double* ms_stage2,
	             double* ms_total);

private:
	void ensure_scratch(int nvec);
	void ensure_fp32();
	void ensure_bf16();
	// The matvec on the given scratch buffers (the plan's own, or a caller's).
	void launch(const double* vin, int64_t ldin, double* vout, int64_t ldout, int nvec, int beta,
	            hipEvent_t mid, void* bx, void* ws, void* xin);

	int         precision_     = PREC_FP64;
	bool        device_blocks_ = false;
	Tuning      tuning_;
	Tuning      tuning_of_[3]; // per precision mode
	PlanStats   stats_;
	StageTables st1_, st2_;
	hipStream_t stream_ = nullptr;

	// Unsplit tasks of both stages as planned from the blocks (retune builds the tables).
	std::vector<PTask>  t1_tasks_, t2_tasks_;
	std::vector<PSeg>   t1_segs_, t2_segs_;
	std::vector<double> cost1_, cost2_;

	// Per-call input preparation (XP / XT panels of the input vector).
	std::vector<PrepTile>  prep_tiles_;
	std::vector<PrepPatch> prep_patches_;
	PrepTile*              d_prep_tiles_   = nullptr;
	PrepPatch*             d_prep_patches_ = nullptr;
	int64_t                xp_elems_       = 0;
	int64_t                xt_elems_       = 0;

	std::vector<Reduce>      reduce_;
	std::vector<ReduceChunk> chunks_;
	Reduce*                  d_reduce_ = nullptr;
	ReduceChunk*             d_chunks_ = nullptr;
	// Balanced persistent stage 2 (Tuning::persist): items, workgroup item ranges, cut tiles.
	std::vector<SkItem>  sk_items_;
	std::vector<int32_t> sk_bins_;
	std::vector<SkFix>   sk_fixes_;
	SkItem*              d_sk_items_ = nullptr;
	int32_t*             d_sk_bins_  = nullptr;
	SkFix*               d_sk_fixes_ = nullptr;
	// Resident stage-2 workgroups per CU of a tuning in the current precision (occupancy query
	// of the kernel the tuning selects).
	int stage2_occupancy(const Tuning& t) const;
	// Builds the balanced item lists of stage 2 from st2_'s XCD-ordered tiles.
	void build_balanced(const Tuning& t);

	std::vector<int64_t> off_;
	double*              d_apack_   = nullptr;
	double*              d_bpack_   = nullptr;
	float*               d_apack32_ = nullptr;
	float*               d_bpack32_ = nullptr;
	uint16_t*            d_apack16_ = nullptr; // bf16 (hi, lo) pairs: hi parts, then lo parts
	uint16_t*            d_bpack16_ = nullptr;
	void*                d_bx_      = nullptr; // stage-1 products, bx_bytes_ bytes
	void*                d_ws_      = nullptr; // split-k partials, ws_bytes_ bytes
	void*                d_xin_     = nullptr; // XP and XT panels of the input, xin_bytes_
	size_t               bx_bytes_  = 0;
	size_t               ws_bytes_  = 0;
	size_t               xin_bytes_ = 0;
	int                  bx_prec_   = -1; // precision whose zero padding d_bx_ holds

	hipEvent_t ev_[3] = { nullptr, nullptr, nullptr };

	// Diagonals of the (I, I, k) terms: per term diag(A) (L_I values) then diag(B) (R_I values)
	// at dg_off_[t]; terms of patch I are [dg_begin_[I], dg_begin_[I + 1]).
	std::vector<double>  dg_vals_;
	std::vector<int64_t> dg_off_;
	std::vector<int32_t> dg_begin_;
	std::vector<int32_t> pL_, pR_;
	void*                d_dg_ = nullptr; // uploaded tables (one allocation)
};

} // namespace dmk
