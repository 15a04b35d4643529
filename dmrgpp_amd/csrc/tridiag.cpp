#include "tridiag.h"

#include <algorithm>
#include <cmath>
#include <limits>

namespace dmk {

namespace {
constexpr double kEps = std::numeric_limits<double>::epsilon();
}

int sturm_count(const double* a, const double* b, int m, double x)
{
	// LDL^T of T - x I: d_0 = a_0 - x, d_i = a_i - x - b_i^2 / d_{i-1}; count negative pivots.
	double scale = 0;
	for (int i = 0; i < m; ++i)
		scale = std::max(scale, std::fabs(a[i]) + (i > 0 ? std::fabs(b[i]) : 0.0));
	const double pivmin = std::max(std::numeric_limits<double>::min(), kEps * kEps * scale * scale);
	int          count  = 0;
	double       d      = 1.0;
	for (int i = 0; i < m; ++i) {
		const double bb = (i > 0) ? b[i] * b[i] : 0.0;
		d               = a[i] - x - ((i > 0) ? bb / d : 0.0);
		if (std::fabs(d) < pivmin)
			d = -pivmin;
		if (d < 0)
			++count;
	}
	return count;
}

double tridiag_eigenvalue(const double* a, const double* b, int m, int k)
{
	// Gershgorin interval.
	double lo = std::numeric_limits<double>::max(), hi = -lo;
	for (int i = 0; i < m; ++i) {
		const double r = (i > 0 ? std::fabs(b[i]) : 0.0) + (i + 1 < m ? std::fabs(b[i + 1]) : 0.0);
		lo             = std::min(lo, a[i] - r);
		hi             = std::max(hi, a[i] + r);
	}
	const double w = std::max(std::fabs(lo), std::fabs(hi));
	lo -= 2 * kEps * w + std::numeric_limits<double>::min();
	hi += 2 * kEps * w + std::numeric_limits<double>::min();
	for (int it = 0; it < 200; ++it) {
		const double mid = 0.5 * (lo + hi);
		if (mid <= lo || mid >= hi)
			break;
		if (sturm_count(a, b, m, mid) > k)
			hi = mid;
		else
			lo = mid;
	}
	return 0.5 * (lo + hi);
}

std::vector<double> tridiag_eigenvector(const double* a, const double* b, int m, double lambda)
{
	std::vector<double> x(static_cast<size_t>(m), 0.0);
	if (m == 1) {
		x[0] = 1.0;
		return x;
	}
	double scale = 0;
	for (int i = 0; i < m; ++i)
		scale = std::max(scale, std::fabs(a[i]) + (i > 0 ? std::fabs(b[i]) : 0.0));
	const double tiny = std::max(kEps * scale, std::numeric_limits<double>::min());
	// Gaussian elimination with partial pivoting of M = T - lambda I (the dgttrf scheme). Before
	// step i, row i holds (d_i, u_i) in columns (i, i+1) and row i+1 is untouched. A row swap
	// gives the pivot row a second superdiagonal entry v_i.
	const size_t        ms = static_cast<size_t>(m);
	std::vector<double> d(ms, 0.0), u(ms, 0.0), v(ms, 0.0), mult(ms, 0.0);
	std::vector<char>   swapped(ms, 0);
	for (int i = 0; i < m; ++i) {
		d[i] = a[i] - lambda;
		if (i + 1 < m)
			u[i] = b[i + 1];
	}
	for (int i = 0; i + 1 < m; ++i) {
		const double sub = b[i + 1]; // M(i+1, i)
		if (std::fabs(sub) > std::fabs(d[i])) {
			// Pivot row becomes old row i+1 = (sub, d_{i+1}, u_{i+1}); row i+1 = old row i - f * pivot.
			const double f      = d[i] / sub;
			const double old_ui = u[i];
			const double old_d1 = d[i + 1];
			const double old_u1 = u[i + 1];
			swapped[i]          = 1;
			mult[i]             = f;
			d[i]                = sub;
			u[i]                = old_d1;
			v[i]                = old_u1;
			d[i + 1]            = old_ui - f * old_d1;
			u[i + 1]            = -f * old_u1;
		} else {
			if (d[i] == 0.0)
				d[i] = tiny;
			const double f = sub / d[i];
			mult[i]        = f;
			d[i + 1] -= f * u[i];
		}
	}
	if (d[m - 1] == 0.0)
		d[m - 1] = tiny;
	for (int i = 0; i < m; ++i)
		if (std::fabs(d[i]) < tiny)
			d[i] = (d[i] < 0 ? -tiny : tiny);
	// Inverse iteration: solve U x_new = (L^-1 x) with the factors; start from all ones.
	std::vector<double> rhs(ms, 1.0);
	for (int iter = 0; iter < 4; ++iter) {
		// Forward: apply the row operations recorded in mult/swapped.
		std::vector<double> y = rhs;
		for (int i = 0; i + 1 < m; ++i) {
			if (swapped[i]) {
				const double t = y[i];
				y[i]           = y[i + 1];
				y[i + 1]       = t - mult[i] * y[i];
			} else {
				y[i + 1] -= mult[i] * y[i];
			}
		}
		// Backward solve with the upper factor (d, u, v).
		for (int i = m - 1; i >= 0; --i) {
			double s = y[i];
			if (i + 1 < m)
				s -= u[i] * x[i + 1];
			if (i + 2 < m)
				s -= v[i] * x[i + 2];
			x[i] = s / d[i];
		}
		double nrm = 0;
		for (double t : x)
			nrm += t * t;
		nrm = std::sqrt(nrm);
		for (double& t : x)
			t /= nrm;
		rhs = x;
	}
	return x;
}

} // namespace dmk
