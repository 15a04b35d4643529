#include "tridiag.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <stdexcept>

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

void tridiag_eigen_all(const double* a, const double* b, int m, double* evals, double* evecs)
{
	if (m < 1)
		return;
	std::vector<double> d(a, a + m);
	// e[i] couples rows i and i+1; e[m-1] = 0.
	std::vector<double> e(static_cast<size_t>(m), 0.0);
	for (int i = 0; i + 1 < m; ++i)
		e[static_cast<size_t>(i)] = b[i + 1];
	std::vector<double> z(static_cast<size_t>(m) * static_cast<size_t>(m), 0.0);
	for (int i = 0; i < m; ++i)
		z[static_cast<size_t>(i) * static_cast<size_t>(m + 1)] = 1.0;
	auto Z = [&](int row, int col) -> double& {
		return z[static_cast<size_t>(row) + static_cast<size_t>(col) * static_cast<size_t>(m)];
	};
	for (int l = 0; l < m; ++l) {
		for (int iter = 0;; ++iter) {
			// Smallest mm >= l with a negligible coupling below it.
			int mm = l;
			for (; mm + 1 < m; ++mm) {
				const double dd = std::fabs(d[size_t(mm)]) + std::fabs(d[size_t(mm + 1)]);
				if (std::fabs(e[size_t(mm)]) <= kEps * dd)
					break;
			}
			if (mm == l)
				break;
			if (iter >= 100)
				throw std::runtime_error("tridiag_eigen_all: no convergence");
			// Wilkinson shift from the leading 2x2 block of the unreduced part.
			double g = (d[size_t(l + 1)] - d[size_t(l)]) / (2.0 * e[size_t(l)]);
			double r = std::hypot(g, 1.0);
			g        = d[size_t(mm)] - d[size_t(l)] + e[size_t(l)] / (g + std::copysign(r, g));
			double s = 1.0, c = 1.0, p = 0.0;
			int    i        = mm - 1;
			bool   deflated = false;
			for (; i >= l; --i) {
				const double f  = s * e[size_t(i)];
				const double bb = c * e[size_t(i)];
				r               = std::hypot(f, g);
				e[size_t(i + 1)] = r;
				if (r == 0.0) {
					// Underflow: the bulge vanished, split here.
					d[size_t(i + 1)] -= p;
					e[size_t(mm)] = 0.0;
					deflated      = true;
					break;
				}
				s = f / r;
				c = g / r;
				g = d[size_t(i + 1)] - p;
				r = (d[size_t(i)] - g) * s + 2.0 * c * bb;
				p = s * r;
				d[size_t(i + 1)] = g + p;
				g                = c * r - bb;
				for (int k = 0; k < m; ++k) {
					const double zf = Z(k, i + 1);
					Z(k, i + 1)     = s * Z(k, i) + c * zf;
					Z(k, i)         = c * Z(k, i) - s * zf;
				}
			}
			if (deflated)
				continue;
			d[size_t(l)] -= p;
			e[size_t(l)]  = g;
			e[size_t(mm)] = 0.0;
		}
	}
	// Ascending order.
	std::vector<int> order(static_cast<size_t>(m));
	for (int i = 0; i < m; ++i)
		order[size_t(i)] = i;
	std::sort(order.begin(), order.end(), [&](int x, int y) { return d[size_t(x)] < d[size_t(y)]; });
	for (int k = 0; k < m; ++k) {
		const int src = order[size_t(k)];
		evals[k]      = d[size_t(src)];
		for (int row = 0; row < m; ++row)
			evecs[size_t(row) + size_t(k) * size_t(m)] = Z(row, src);
	}
}

} // namespace dmk
