// extern "C" layer of dmrgpp_amd (declarations and reference citations in include/dmrgpp_amd.h).
#include "complex_embed.h"
#include "davidson.h"
#include "dmrgpp_amd.h"
#include "kron_plan.h"
#include "krylov.h"
#include "lanczos.h"
#include "noalloc_scratch.h"
#include "pool.h"
#include "tridiag.h"
#include "vbatch.h"

#include <hip/hip_runtime.h>

#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

struct dmk_mlp {
	std::unique_ptr<dmk::Mlp> net;
};

struct dmk_handle {
	std::unique_ptr<dmk::Plan>             plan;
	std::unique_ptr<dmk::LanczosWorkspace> lanczos;
	std::unique_ptr<dmk::KrylovWorkspace>  krylov;
	bool                                   is_complex = false; // plan is a real embedding
	int64_t                                nc         = 0;     // complex vector length
	double*                                zbuf       = nullptr; // split in / out of dmk_apply_z
	int64_t*                               smap       = nullptr; // sector map (device)
	double*                                sbuf       = nullptr; // gathered input, product
	~dmk_handle()
	{
		if ((zbuf || smap || sbuf) && plan)
			(void)hipStreamSynchronize(plan->stream());
		dmk::pool_device_free(zbuf);
		dmk::pool_device_free(smap);
		dmk::pool_device_free(sbuf);
	}
};

namespace dmk {

void density_diag(int nblocks, const int64_t* n, double* const* blocks, double* eigs);
void density_diag_device(int nblocks, const int64_t* n, const double* const* dblocks,
                         double* const* vecs, double* eigs);
void density_svd(int nblocks, const int64_t* rows, const int64_t* cols,
                 const double* const* blocks, double* const* u, double* s);
void launch_copy(double* y, const double* x, int64_t n, hipStream_t s);
void launch_sector_gather(const double* v, const int64_t* map, int64_t n, double* y, hipStream_t s);
void launch_sector_scatter_add(const double* x, const int64_t* map, int64_t n, double* v,
                               hipStream_t s);

namespace {
thread_local std::string g_last_error;
}

void set_last_error(const std::string& msg) { g_last_error = msg; }

const char* last_error() { return g_last_error.c_str(); }

} // namespace dmk

namespace {

template <typename F> int guarded(F&& f)
{
	try {
		return f();
	} catch (const std::invalid_argument& e) {
		dmk::set_last_error(e.what());
		return DMK_ERR_ARGUMENT;
	} catch (const std::runtime_error& e) {
		dmk::set_last_error(e.what());
		const std::string w = e.what();
		if (w.find("hip") != std::string::npos)
			return DMK_ERR_DEVICE;
		if (w.find("unsupported") != std::string::npos)
			return DMK_ERR_UNSUPPORTED;
		return DMK_ERR_ARGUMENT;
	} catch (const std::exception& e) {
		dmk::set_last_error(e.what());
		return DMK_ERR_DEVICE;
	}
}

dmk::Mlp& net_of(const dmk_mlp* m)
{
	if (!m || !m->net)
		throw std::invalid_argument("null dmk_mlp");
	return *m->net;
}

dmk::Plan& plan_of(dmk_handle* h)
{
	if (!h || !h->plan)
		throw std::invalid_argument("null dmk handle");
	return *h->plan;
}

} // namespace

extern "C" {

int dmk_create_d(size_t noperator, size_t npatches, const size_t* left_patch_size,
                 const size_t* right_patch_size, const double* const* Amatrix,
                 const size_t* ld_Amatrix, const double* const* Bmatrix, const size_t* ld_Bmatrix,
                 dmk_handle** handle)
{
	return guarded([&] {
		if (!handle)
			throw std::invalid_argument("dmk_create_d: null handle pointer");
		*handle = nullptr;
		auto h  = std::make_unique<dmk_handle>();
		h->plan = std::make_unique<dmk::Plan>(int64_t(noperator), int64_t(npatches),
		                                      left_patch_size, right_patch_size, Amatrix,
		                                      ld_Amatrix, Bmatrix, ld_Bmatrix);
		h->nc   = h->plan->n();
		*handle = h.release();
		return DMK_OK;
	});
}

int dmk_create_z(size_t noperator, size_t npatches, const size_t* left_patch_size,
                 const size_t* right_patch_size, const double* const* Amatrix,
                 const size_t* ld_Amatrix, const double* const* Bmatrix, const size_t* ld_Bmatrix,
                 dmk_handle** handle)
{
	return guarded([&] {
		if (!handle)
			throw std::invalid_argument("dmk_create_z: null handle pointer");
		*handle = nullptr;
		if (!Amatrix || !Bmatrix)
			throw std::invalid_argument("dmk_create_z: null argument array");
		if (dmk::blocks_on_device(int64_t(noperator), int64_t(npatches), Amatrix, Bmatrix))
			throw std::invalid_argument("dmk_create_z: complex blocks must be host memory "
			                            "(device-resident blocks: dmk_create_d)");
		dmk::RealEmbedding e = dmk::embed_complex(int64_t(noperator), int64_t(npatches),
		                                          left_patch_size, right_patch_size, Amatrix,
		                                          ld_Amatrix, Bmatrix, ld_Bmatrix);
		auto h  = std::make_unique<dmk_handle>();
		h->plan = std::make_unique<dmk::Plan>(e.nC, e.np, e.L.data(), e.R.data(), e.A.data(),
		                                      e.ldA.data(), e.B.data(), e.ldB.data());
		h->is_complex = true;
		h->nc         = h->plan->n() / 2;
		*handle       = h.release();
		return DMK_OK;
	});
}

void dmk_destroy(dmk_handle* handle) { delete handle; }

int dmk_apply_d(dmk_handle* handle, const double* vin, double* vout)
{
	return guarded([&] {
		dmk::Plan& p = plan_of(handle);
		if (handle->is_complex)
			throw std::invalid_argument("dmk_apply_d: complex handle (use dmk_apply_z)");
		p.apply(vin, p.n(), vout, p.n(), 1, 0);
		return DMK_OK;
	});
}

int dmk_apply_host_d(dmk_handle* handle, const double* vin, double* vout)
{
	return guarded([&] {
		dmk::Plan&   p     = plan_of(handle);
		const size_t bytes = size_t(p.n()) * sizeof(double);
		double*      dx    = nullptr;
		double*      dy    = nullptr;
		DMK_HIP_CHECK(hipMalloc(&dx, bytes));
		try {
			DMK_HIP_CHECK(hipMalloc(&dy, bytes));
			DMK_HIP_CHECK(hipMemcpyAsync(dx, vin, bytes, hipMemcpyHostToDevice, p.stream()));
			p.apply(dx, p.n(), dy, p.n(), 1, 0);
			DMK_HIP_CHECK(hipMemcpyAsync(vout, dy, bytes, hipMemcpyDeviceToHost, p.stream()));
			DMK_HIP_CHECK(hipStreamSynchronize(p.stream()));
		} catch (...) {
			(void)hipStreamSynchronize(p.stream());
			(void)hipFree(dx);
			(void)hipFree(dy);
			throw;
		}
		(void)hipFree(dx);
		(void)hipFree(dy);
		return DMK_OK;
	});
}

int dmk_apply_z(dmk_handle* handle, const double* vin, double* vout)
{
	return guarded([&] {
		dmk::Plan& p = plan_of(handle);
		if (!vin || !vout)
			throw std::invalid_argument("dmk_apply_z: null vector");
		const int64_t nc = handle->nc;
		if (!handle->zbuf)
			handle->zbuf = static_cast<double*>(dmk::pool_device_alloc(size_t(4 * nc) * sizeof(double)));
		double* xin  = handle->zbuf;
		double* xout = handle->zbuf + 2 * nc;
		dmk::launch_kry_deinterleave(vin, nc, xin, 2 * nc, nc, 1, p.stream());
		dmk::KrylovOperator op;
		op.plan     = &p;
		op.embedded = handle->is_complex;
		op.nc       = nc;
		op.apply(xin, xout, 1);
		dmk::launch_kry_interleave(xout, 2 * nc, vout, nc, nc, 1, p.stream());
		return DMK_OK;
	});
}

int dmk_apply_shared(dmk_handle* handle, const void* vin, void* vout)
{
	// Whether a kernel may read / write the vector in place: device memory, and managed memory
	// (Kokkos::SharedSpace allocations). Pageable and pinned host vectors are copied: the
	// matvec reads its input several times, which over PCIe would cost more than the copies.
	auto in_place = [](const void* ptr) {
		hipPointerAttribute_t a {};
		const hipError_t      e = hipPointerGetAttributes(&a, ptr);
		// Unregistered host memory makes the query fail; its status must not reach the launch
		// checks of this call.
		(void)hipGetLastError();
		if (e != hipSuccess)
			return false;
		return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged || a.isManaged;
	};
	return guarded([&] {
		dmk::Plan& p = plan_of(handle);
		if (!vin || !vout)
			throw std::invalid_argument("dmk_apply_shared: null vector");
		const bool    cplx  = handle->is_complex;
		const size_t  bytes = size_t(cplx ? 2 * handle->nc : p.n()) * sizeof(double);
		const bool    di    = in_place(vin);
		const bool    dout  = in_place(vout);
		double*       tmp   = nullptr; // device copies of the host vectors
		if (!di || !dout)
			tmp = static_cast<double*>(dmk::pool_device_alloc(2 * bytes));
		const double* x = di ? static_cast<const double*>(vin) : tmp;
		double*       y = dout ? static_cast<double*>(vout) : tmp + bytes / sizeof(double);
		try {
			if (!di)
				DMK_HIP_CHECK(hipMemcpyAsync(tmp, vin, bytes, hipMemcpyHostToDevice, p.stream()));
			const int st = cplx ? dmk_apply_z(handle, x, y) : dmk_apply_d(handle, x, y);
			if (st != DMK_OK)
				throw std::runtime_error(dmk::last_error());
			if (!dout)
				DMK_HIP_CHECK(hipMemcpyAsync(vout, y, bytes, hipMemcpyDeviceToHost, p.stream()));
			DMK_HIP_CHECK(hipStreamSynchronize(p.stream()));
		} catch (...) {
			(void)hipStreamSynchronize(p.stream());
			dmk::pool_device_free(tmp);
			throw;
		}
		dmk::pool_device_free(tmp);
		return DMK_OK;
	});
}

int dmk_apply_noalloc_d(dmk_handle* handle, const double* vin, double* vout)
{
	static_assert(DMK_NOALLOC_SCRATCH_BYTES == dmk::kNoallocScratchBytes,
	              "header and library disagree on the static scratch size");
	return guarded([&] {
		dmk::Plan& p = plan_of(handle);
		if (handle->is_complex)
			throw std::invalid_argument("dmk_apply_noalloc_d: complex handle (use dmk_apply_z)");
		p.apply_on_scratch(vin, vout, dmk::noalloc_scratch_device_ptr(),
		                   dmk::kNoallocScratchBytes);
		return DMK_OK;
	});
}

int dmk_noalloc_scratch_bytes(const dmk_handle* handle, size_t* bytes)
{
	return guarded([&] {
		if (!handle || !handle->plan || !bytes)
			throw std::invalid_argument("dmk_noalloc_scratch_bytes: null argument");
		*bytes = handle->plan->scratch_layout(nullptr, nullptr);
		return DMK_OK;
	});
}

int dmk_set_sector_map(dmk_handle* handle, const int64_t* map)
{
	return guarded([&] {
		dmk::Plan& p = plan_of(handle);
		if (!map)
			throw std::invalid_argument("dmk_set_sector_map: null map");
		if (handle->is_complex)
			throw std::invalid_argument("dmk_set_sector_map: complex handle");
		const int64_t     n = p.n();
		std::vector<char> seen(size_t(n), 0);
		for (int64_t i = 0; i < n; ++i) {
			if (map[i] < 0 || map[i] >= n || seen[size_t(map[i])])
				throw std::invalid_argument("dmk_set_sector_map: not a permutation of 0..n-1 (entry "
				                            + std::to_string(i) + ")");
			seen[size_t(map[i])] = 1;
		}
		if (!handle->smap)
			handle->smap = static_cast<int64_t*>(dmk::pool_device_alloc(size_t(n) * sizeof(int64_t)));
		DMK_HIP_CHECK(hipMemcpyAsync(handle->smap, map, size_t(n) * sizeof(int64_t),
		                             hipMemcpyHostToDevice, p.stream()));
		DMK_HIP_CHECK(hipStreamSynchronize(p.stream()));
		return DMK_OK;
	});
}

int dmk_matvec_sector_d(dmk_handle* handle, const double* vin, double* vout)
{
	return guarded([&] {
		dmk::Plan& p = plan_of(handle);
		if (!handle->smap)
			throw std::invalid_argument("dmk_matvec_sector_d: no sector map (dmk_set_sector_map)");
		if (!vin || !vout)
			throw std::invalid_argument("dmk_matvec_sector_d: null vector");
		const int64_t n = p.n();
		if (!handle->sbuf)
			handle->sbuf = static_cast<double*>(dmk::pool_device_alloc(size_t(2 * n) * sizeof(double)));
		dmk::launch_sector_gather(vin, handle->smap, n, handle->sbuf, p.stream());
		p.apply(handle->sbuf, n, handle->sbuf + n, n, 1, 0);
		dmk::launch_sector_scatter_add(handle->sbuf + n, handle->smap, n, vout, p.stream());
		return DMK_OK;
	});
}

int dmk_apply_multi_d(dmk_handle* handle, int nvec, const double* vin, size_t ldin, double* vout,
                      size_t ldout, int beta)
{
	return guarded([&] {
		dmk::Plan& p = plan_of(handle);
		if (beta != 0 && beta != 1)
			throw std::invalid_argument("dmk_apply_multi_d: beta must be 0 or 1");
		if (handle->is_complex)
			throw std::invalid_argument("dmk_apply_multi_d: complex handle (use dmk_apply_z)");
		p.apply(vin, int64_t(ldin), vout, int64_t(ldout), nvec, beta);
		return DMK_OK;
	});
}

int dmk_set_precision(dmk_handle* handle, int precision)
{
	return guarded([&] {
		plan_of(handle).set_precision(precision);
		return DMK_OK;
	});
}

int dmk_get_info(const dmk_handle* handle, dmk_info* info)
{
	return guarded([&] {
		if (!handle || !handle->plan || !info)
			throw std::invalid_argument("dmk_get_info: null argument");
		const dmk::PlanStats& s = handle->plan->stats();
		std::memset(info, 0, sizeof(*info));
		info->n                 = s.n;
		info->npatches          = s.npatches;
		info->noperator         = s.noperator;
		info->nterms            = s.nterms;
		info->nterms_identity   = s.nterms_identity;
		info->stage1_tasks      = s.stage1_tasks;
		info->stage2_tasks      = s.stage2_tasks;
		info->stage1_tiles      = s.stage1_tiles;
		info->stage2_tiles      = s.stage2_tiles;
		info->max_patch_left    = s.max_patch_left;
		info->max_patch_right   = s.max_patch_right;
		info->apack_elems       = s.apack_elems;
		info->bpack_elems       = s.bpack_elems;
		info->bx_elems          = s.bx_elems;
		info->flops_algorithmic = s.flops_algorithmic;
		info->flops_plugin      = s.flops_plugin;
		info->flops_executed    = s.flops_executed;
		info->flops_stage1      = s.flops_stage1;
		info->flops_stage2      = s.flops_stage2;
		info->bytes_compulsory  = s.bytes_compulsory;
		info->bytes_bx          = s.bytes_bx;
		info->precision         = handle->plan->precision();
		info->split_tasks       = s.split_tasks;
		info->ws_elems          = s.ws_elems;
		info->is_complex        = handle->is_complex ? 1 : 0;
		info->device_blocks     = handle->plan->device_blocks() ? 1 : 0;
		info->n_complex         = handle->nc;
		info->ms_scan           = s.ms_scan;
		info->ms_stage          = s.ms_stage;
		info->ms_upload         = s.ms_upload;
		info->raw_bytes         = s.raw_bytes;
		return DMK_OK;
	});
}

namespace {
void to_c(const dmk::Tuning& t, dmk_tuning* c)
{
	c->stage1_tile_m = t.bm1;
	c->stage1_tile_n = t.bn1;
	c->stage2_tile_m = t.bm2;
	c->stage2_tile_n = t.bn2;
	c->splitk        = t.splitk;
	c->xcd_order     = t.xcd;
	c->mfma32        = t.mfma32;
	c->prefetch      = t.depth;
	c->kchunk        = t.kchunk;
	c->persist       = t.persist;
}
} // namespace

int dmk_get_tuning(const dmk_handle* handle, dmk_tuning* tuning)
{
	return guarded([&] {
		if (!handle || !handle->plan || !tuning)
			throw std::invalid_argument("dmk_get_tuning: null argument");
		to_c(handle->plan->tuning(), tuning);
		return DMK_OK;
	});
}

int dmk_set_tuning(dmk_handle* handle, const dmk_tuning* tuning)
{
	return guarded([&] {
		if (!tuning)
			throw std::invalid_argument("dmk_set_tuning: null tuning");
		dmk::Tuning t;
		t.bm1    = tuning->stage1_tile_m;
		t.bn1    = tuning->stage1_tile_n;
		t.bm2    = tuning->stage2_tile_m;
		t.bn2    = tuning->stage2_tile_n;
		t.splitk = tuning->splitk;
		t.xcd    = tuning->xcd_order;
		t.mfma32 = tuning->mfma32;
		t.depth  = tuning->prefetch == 0 ? 1 : tuning->prefetch;
		t.kchunk = tuning->kchunk == 0 ? 1 : tuning->kchunk;
		t.persist = tuning->persist;
		plan_of(handle).retune(t);
		return DMK_OK;
	});
}

int dmk_autotune_d(dmk_handle* handle, const double* vin, double* vout, int iters,
                   dmk_tuning* chosen)
{
	return guarded([&] {
		if (!vin || !vout)
			throw std::invalid_argument("dmk_autotune_d: null vector");
		dmk::Tuning t = plan_of(handle).autotune(vin, vout, iters);
		if (chosen)
			to_c(t, chosen);
		return DMK_OK;
	});
}

int dmk_sync(dmk_handle* handle)
{
	return guarded([&] {
		DMK_HIP_CHECK(hipStreamSynchronize(plan_of(handle).stream()));
		return DMK_OK;
	});
}

int dmk_profile_apply_d(dmk_handle* handle, const double* vin, double* vout, int iters,
                        double* ms_stage1, double* ms_stage2, double* ms_total)
{
	return guarded([&] {
		plan_of(handle).profile(vin, vout, iters, ms_stage1, ms_stage2, ms_total);
		return DMK_OK;
	});
}

int dmk_lanczos_default_params(dmk_lanczos_params* params)
{
	return guarded([&] {
		if (!params)
			throw std::invalid_argument("dmk_lanczos_default_params: null pointer");
		dmk::LanczosParams d;
		std::memset(params, 0, sizeof(*params));
		params->max_steps        = d.max_steps;
		params->min_steps        = d.min_steps;
		params->tolerance        = d.tolerance;
		params->reortho          = d.reortho;
		params->precision        = d.precision;
		params->refine           = d.refine;
		params->refine_min_steps = d.refine_min_steps;
		params->refine_tolerance = d.refine_tolerance;
		params->graph            = d.graph;
		params->stop_residual    = d.stop_residual;
		params->switch_noise     = d.switch_noise;
		return DMK_OK;
	});
}

int dmk_lanczos_ground_d(dmk_handle* handle, const double* init, double* z,
                         const dmk_lanczos_params* params, dmk_lanczos_result* result)
{
	return guarded([&] {
		dmk::Plan& p = plan_of(handle);
		if (!params || !result)
			throw std::invalid_argument("dmk_lanczos_ground_d: null argument");
		dmk::LanczosParams lp;
		lp.max_steps        = params->max_steps;
		lp.min_steps        = params->min_steps;
		lp.tolerance        = params->tolerance;
		lp.reortho          = params->reortho;
		lp.precision        = params->precision;
		lp.refine           = params->refine;
		lp.refine_min_steps = params->refine_min_steps;
		lp.refine_tolerance = params->refine_tolerance;
		lp.graph            = params->graph;
		lp.stop_residual    = params->stop_residual;
		lp.switch_noise     = params->switch_noise;
		if (!(lp.switch_noise >= 0))
			throw std::invalid_argument("dmk_lanczos_ground_d: switch_noise must be >= 0");
		if (!handle->lanczos)
			handle->lanczos = std::make_unique<dmk::LanczosWorkspace>();
		dmk::LanczosReport r = dmk::lanczos_ground(p, *handle->lanczos, init, z, lp);
		std::memset(result, 0, sizeof(*result));
		result->energy          = r.energy;
		result->residual        = r.residual;
		result->steps           = r.steps;
		result->matvecs         = r.matvecs;
		result->energy_first    = r.energy_first;
		result->steps_first     = r.steps_first;
		result->matvecs_first   = r.matvecs_first;
		result->precision_first = r.precision_first;
		result->ms_total        = r.ms_total;
		result->ms_matvec       = r.ms_matvec;
		return DMK_OK;
	});
}

int dmk_tridiag_lowest(int m, const double* a, const double* b, double* eval, double* evec)
{
	return guarded([&] {
		if (m < 1 || !a || !b || !eval)
			throw std::invalid_argument("dmk_tridiag_lowest: bad arguments");
		*eval = dmk::tridiag_eigenvalue(a, b, m, 0);
		if (evec) {
			std::vector<double> v = dmk::tridiag_eigenvector(a, b, m, *eval);
			std::memcpy(evec, v.data(), size_t(m) * sizeof(double));
		}
		return DMK_OK;
	});
}

int dmk_tridiag_eigen(int m, const double* a, const double* b, double* evals, double* evecs)
{
	return guarded([&] {
		if (m < 1 || !a || !b || !evals || !evecs)
			throw std::invalid_argument("dmk_tridiag_eigen: bad arguments");
		dmk::tridiag_eigen_all(a, b, m, evals, evecs);
		return DMK_OK;
	});
}

int dmk_krylov_default_params(dmk_krylov_params* params)
{
	return guarded([&] {
		if (!params)
			throw std::invalid_argument("dmk_krylov_default_params: null pointer");
		dmk::KrylovParams d;
		std::memset(params, 0, sizeof(*params));
		params->max_steps      = d.max_steps;
		params->min_steps      = d.min_steps;
		params->tolerance      = d.tolerance;
		params->stop           = d.stop;
		params->precision      = d.precision;
		params->time_direction = d.time_direction;
		params->energy_shift   = d.energy_shift;
		return DMK_OK;
	});
}

int dmk_krylov_evolve_z(dmk_handle* handle, int nvec, const double* phi, size_t ldphi, int ntimes,
                        const double* times, double* out, size_t ldout,
                        const dmk_krylov_params* params, dmk_krylov_result* results,
                        dmk_krylov_stats* stats)
{
	return guarded([&] {
		dmk::Plan& p = plan_of(handle);
		if (!params)
			throw std::invalid_argument("dmk_krylov_evolve_z: null params");
		dmk::KrylovParams kp;
		kp.max_steps      = params->max_steps;
		kp.min_steps      = params->min_steps;
		kp.tolerance      = params->tolerance;
		kp.stop           = params->stop;
		kp.precision      = params->precision;
		kp.time_direction = params->time_direction;
		kp.energy_shift   = params->energy_shift;
		dmk::KrylovOperator op;
		op.plan     = &p;
		op.embedded = handle->is_complex;
		op.nc       = handle->nc;
		if (!handle->krylov)
			handle->krylov = std::make_unique<dmk::KrylovWorkspace>();
		double ms_total = 0, ms_mv = 0;
		int    launches = 0;
		std::vector<dmk::KrylovReport> reps = dmk::krylov_evolve(
		    op, *handle->krylov, nvec, phi, int64_t(ldphi), ntimes, times, out, int64_t(ldout), kp,
		    &ms_total, &ms_mv, &launches);
		if (results)
			for (int t = 0; t < nvec; ++t) {
				results[t].steps          = reps[size_t(t)].steps;
				results[t].breakdown      = reps[size_t(t)].breakdown;
				results[t].norm           = reps[size_t(t)].norm;
				results[t].error_estimate = reps[size_t(t)].error_estimate;
			}
		if (stats) {
			std::memset(stats, 0, sizeof(*stats));
			stats->ms_total        = ms_total;
			stats->ms_matvec       = ms_mv;
			stats->matvec_launches = launches;
		}
		return DMK_OK;
	});
}

int dmk_gemm_vbatch_d(char transa, char transb, const int* m_array, const int* n_array,
                      const int* k_array, double alpha, const double* const* a_array,
                      const int* lda_array, const double* const* b_array, const int* ldb_array,
                      double beta, double* const* c_array, const int* ldc_array,
                      size_t group_count, const int* group_size)
{
	return guarded([&] {
		dmk::gemm_vbatch(false, transa, transb, m_array, n_array, k_array, &alpha, a_array,
		                 lda_array, b_array, ldb_array, &beta, c_array, ldc_array, group_count,
		                 group_size);
		return DMK_OK;
	});
}

int dmk_gemm_vbatch_z(char transa, char transb, const int* m_array, const int* n_array,
                      const int* k_array, const double* alpha, const double* const* a_array,
                      const int* lda_array, const double* const* b_array, const int* ldb_array,
                      const double* beta, double* const* c_array, const int* ldc_array,
                      size_t group_count, const int* group_size)
{
	return guarded([&] {
		if (!alpha || !beta)
			throw std::invalid_argument("dmk_gemm_vbatch_z: null alpha or beta");
		dmk::gemm_vbatch(true, transa, transb, m_array, n_array, k_array, alpha, a_array,
		                 lda_array, b_array, ldb_array, beta, c_array, ldc_array, group_count,
		                 group_size);
		return DMK_OK;
	});
}

int dmk_density_diag_d(int nblocks, const int64_t* n, double* const* blocks, double* eigs)
{
	return guarded([&] {
		dmk::density_diag(nblocks, n, blocks, eigs);
		return DMK_OK;
	});
}

int dmk_density_diag_dev_d(int nblocks, const int64_t* n, const double* const* blocks,
                           double* const* vecs, double* eigs)
{
	return guarded([&] {
		dmk::density_diag_device(nblocks, n, blocks, vecs, eigs);
		return DMK_OK;
	});
}

int dmk_density_svd_d(int nblocks, const int64_t* rows, const int64_t* cols,
                      const double* const* blocks, double* const* u, double* s)
{
	return guarded([&] {
		dmk::density_svd(nblocks, rows, cols, blocks, u, s);
		return DMK_OK;
	});
}

int dmk_diagonal_d(dmk_handle* handle, double* diag)
{
	return guarded([&] {
		dmk::Plan& p = plan_of(handle);
		p.diagonal(diag);
		DMK_HIP_CHECK(hipStreamSynchronize(p.stream()));
		return DMK_OK;
	});
}

int dmk_mlp_create(uint64_t seed, int64_t sample_capacity, dmk_mlp** mlp)
{
	return guarded([&] {
		if (!mlp)
			throw std::invalid_argument("dmk_mlp_create: null pointer");
		*mlp   = nullptr;
		auto m = std::make_unique<dmk_mlp>();
		m->net = std::make_unique<dmk::Mlp>(seed, sample_capacity);
		*mlp   = m.release();
		return DMK_OK;
	});
}

void dmk_mlp_destroy(dmk_mlp* mlp) { delete mlp; }

int dmk_mlp_get_params(const dmk_mlp* mlp, float* params)
{
	return guarded([&] {
		if (!params)
			throw std::invalid_argument("dmk_mlp_get_params: null pointer");
		net_of(mlp).get_params(params);
		return DMK_OK;
	});
}

int dmk_mlp_set_params(dmk_mlp* mlp, const float* params)
{
	return guarded([&] {
		if (!params)
			throw std::invalid_argument("dmk_mlp_set_params: null pointer");
		net_of(mlp).set_params(params);
		return DMK_OK;
	});
}

int dmk_mlp_forward(dmk_mlp* mlp, const float* x, int64_t count, float* y)
{
	return guarded([&] {
		dmk::Mlp& m = net_of(mlp);
		if (count < 0 || (count > 0 && (!x || !y)))
			throw std::invalid_argument("dmk_mlp_forward: bad arguments");
		dmk::launch_mlp_forward(m.params(), x, count, y, m.stream());
		DMK_HIP_CHECK(hipStreamSynchronize(m.stream()));
		return DMK_OK;
	});
}

int dmk_mlp_gradient(dmk_mlp* mlp, const float* x, const float* t, int64_t count, float* grad,
                     double* loss)
{
	return guarded([&] {
		double l = net_of(mlp).gradient(x, t, count, grad);
		if (loss)
			*loss = l;
		return DMK_OK;
	});
}

int dmk_mlp_add_samples(dmk_mlp* mlp, const float* x, const float* t, int64_t count)
{
	return guarded([&] {
		dmk::Mlp& m = net_of(mlp);
		if (count < 0 || (count > 0 && (!x || !t)))
			throw std::invalid_argument("dmk_mlp_add_samples: bad arguments");
		const int64_t first = m.append(count);
		const int64_t got   = m.samples() - first;
		if (got > 0) {
			DMK_HIP_CHECK(hipMemcpyAsync(m.samples_x() + first * dmk::kMlpIn, x,
			                             size_t(got) * dmk::kMlpIn * sizeof(float),
			                             hipMemcpyDeviceToDevice, m.stream()));
			DMK_HIP_CHECK(hipMemcpyAsync(m.samples_t() + first, t, size_t(got) * sizeof(float),
			                             hipMemcpyDeviceToDevice, m.stream()));
			DMK_HIP_CHECK(hipStreamSynchronize(m.stream()));
		}
		return DMK_OK;
	});
}

int dmk_mlp_samples(const dmk_mlp* mlp, int64_t* count)
{
	return guarded([&] {
		if (!count)
			throw std::invalid_argument("dmk_mlp_samples: null pointer");
		*count = net_of(mlp).samples();
		return DMK_OK;
	});
}

int dmk_mlp_clear_samples(dmk_mlp* mlp)
{
	return guarded([&] {
		net_of(mlp).clear_samples();
		return DMK_OK;
	});
}

int dmk_mlp_train(dmk_mlp* mlp, int iters, int64_t batch, float lr, double* loss_last,
                  double* loss_all)
{
	return guarded([&] {
		if (iters < 0 || !(lr > 0))
			throw std::invalid_argument("dmk_mlp_train: bad arguments");
		net_of(mlp).train(iters, batch, lr, loss_last, loss_all);
		return DMK_OK;
	});
}

int dmk_davidson_default_params(dmk_davidson_params* params)
{
	return guarded([&] {
		if (!params)
			throw std::invalid_argument("dmk_davidson_default_params: null pointer");
		dmk::DavidsonParams d;
		std::memset(params, 0, sizeof(*params));
		params->max_matvecs     = d.max_matvecs;
		params->max_basis       = d.max_basis;
		params->tolerance       = d.tolerance;
		params->precond         = d.precond;
		params->harvest         = d.harvest;
		params->harvest_samples = d.harvest_samples;
		params->harvest_iters   = d.harvest_iters;
		return DMK_OK;
	});
}

int dmk_davidson_ground_d(dmk_handle* handle, const double* init, double* z,
                          const dmk_davidson_params* params, dmk_mlp* mlp,
                          dmk_davidson_result* result)
{
	return guarded([&] {
		dmk::Plan& p = plan_of(handle);
		if (handle->is_complex)
			throw std::invalid_argument("dmk_davidson_ground_d: complex handle");
		if (!params || !result)
			throw std::invalid_argument("dmk_davidson_ground_d: null argument");
		dmk::DavidsonParams dp;
		dp.max_matvecs     = params->max_matvecs;
		dp.max_basis       = params->max_basis;
		dp.tolerance       = params->tolerance;
		dp.precond         = params->precond;
		dp.harvest         = params->harvest;
		dp.harvest_samples = params->harvest_samples;
		dp.harvest_iters   = params->harvest_iters;
		dmk::DavidsonReport r = dmk::davidson_ground(p, init, z, dp, mlp ? mlp->net.get() : nullptr);
		std::memset(result, 0, sizeof(*result));
		result->energy     = r.energy;
		result->residual   = r.residual;
		result->iterations = r.iterations;
		result->matvecs    = r.matvecs;
		result->restarts   = r.restarts;
		result->converged  = r.converged;
		result->samples    = r.samples;
		result->ms_total   = r.ms_total;
		result->ms_matvec  = r.ms_matvec;
		result->ms_precond = r.ms_precond;
		return DMK_OK;
	});
}

int dmk_device_alloc(size_t bytes, void** ptr)
{
	return guarded([&] {
		if (!ptr)
			throw std::invalid_argument("dmk_device_alloc: null pointer");
		*ptr = nullptr;
		*ptr = dmk::pool_device_alloc(bytes ? bytes : 1);
		return DMK_OK;
	});
}

int dmk_device_free(void* ptr)
{
	return guarded([&] {
		if (dmk::pool_device_owns(ptr)) {
			// Work queued on any stream may still read it: wait before it can be handed out again.
			DMK_HIP_CHECK(hipDeviceSynchronize());
			dmk::pool_device_free(ptr);
		} else {
			DMK_HIP_CHECK(hipFree(ptr));
		}
		return DMK_OK;
	});
}

int dmk_memcpy_h2d(void* dst, const void* src, size_t bytes)
{
	return guarded([&] {
		DMK_HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
		return DMK_OK;
	});
}

int dmk_memcpy_d2h(void* dst, const void* src, size_t bytes)
{
	return guarded([&] {
		DMK_HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
		return DMK_OK;
	});
}

int dmk_pool_counters(int host, uint64_t* out)
{
	return guarded([&] {
		if (!out)
			throw std::invalid_argument("dmk_pool_counters: null pointer");
		const dmk::PoolCounters c = host ? dmk::pool_host_counters() : dmk::pool_device_counters();
		out[0] = c.used_bytes;
		out[1] = c.cached_bytes;
		out[2] = c.allocs;
		out[3] = c.frees;
		out[4] = c.reuses;
		out[5] = c.bad_releases;
		return DMK_OK;
	});
}

int dmk_device_count(int* count)
{
	return guarded([&] {
		if (!count)
			throw std::invalid_argument("dmk_device_count: null pointer");
		*count = 0;
		hipError_t e = hipGetDeviceCount(count);
		if (e != hipSuccess)
			*count = 0;
		return DMK_OK;
	});
}

int dmk_device_copy_d(double* dst, const double* src, size_t n)
{
	return guarded([&] {
		if (!dst || !src)
			throw std::invalid_argument("dmk_device_copy_d: null pointer");
		dmk::launch_copy(dst, src, int64_t(n), nullptr);
		DMK_HIP_CHECK(hipGetLastError());
		DMK_HIP_CHECK(hipDeviceSynchronize());
		return DMK_OK;
	});
}

const char* dmk_last_error(void) { return dmk::last_error(); }

} // extern "C"
