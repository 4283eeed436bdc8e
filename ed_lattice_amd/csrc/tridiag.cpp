// Symmetric tridiagonal eigensolver on the host, FP64.
//
// Implicit QL iteration with Wilkinson-type shifts (the method of Bowdler,
// Martin, Reinsch and Wilkinson, Numer. Math. 11 (1968) 293; see also
// Golub & Van Loan, "Matrix Computations", sec. 8.3).  It stands in for the
// dense `eigh` the reference calls on the Lanczos tridiagonal
// (GF_ELECTRONS.f90:445-447, CHI_SPIN_ELECTRONS.f90:328-330,
// DIAG.f90 via SciFortran sp_lanc_eigh).  Only the first row of the
// eigenvector matrix is needed for pole weights Z(1,j)^2, so rotations can
// be applied to that row alone (O(n^2)); the ground-state reconstruction
// asks for full eigenvectors.
#include "tridiag.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <stdexcept>

namespace edl {

// d: diagonal [n]; e: off-diagonal, e[i] couples i and i+1 (e[n-1] ignored).
// On return w holds ascending eigenvalues; Z (n x n, column-major, column j
// = eigenvector j) if non-null; z1 (first components) if non-null.
void tridiag_eig(int n, const double* d_in, const double* e_in, double* w, double* Z, double* z1) {
  if (n <= 0) return;
  std::vector<double> d(d_in, d_in + n), e(n, 0.0);
  for (int i = 0; i + 1 < n; ++i) e[i] = e_in[i];
  const bool full = Z != nullptr;
  const int rows = full ? n : 1;
  std::vector<double> V((size_t)rows * n, 0.0);  // row-major rows x n
  if (full)
    for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1.0;
  else
    V[0] = 1.0;
  const double eps = std::numeric_limits<double>::epsilon();
  for (int l = 0; l < n; ++l) {
    int iter = 0;
    for (;;) {
      int m = l;
      for (; m < n - 1; ++m) {
        const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= eps * dd) break;
      }
      if (m == l) break;
      if (++iter > 200) throw std::runtime_error("tridiag_eig: no convergence");
      // Wilkinson-type shift from the leading 2x2 block
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = std::hypot(g, 1.0);
      g = d[m] - d[l] + e[l] / (g + std::copysign(r, g));
      double s = 1.0, c = 1.0, p = 0.0;
      int i = m - 1;
      bool deflated = false;
      for (; i >= l; --i) {
        const double f = s * e[i], b = c * e[i];
        r = std::hypot(f, g);
        e[i + 1] = r;
        if (r == 0.0) {  // underflow: split the matrix here
          d[i + 1] -= p;
          e[m] = 0.0;
          deflated = true;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - p;
        r = (d[i] - g) * s + 2.0 * c * b;
        p = s * r;
        d[i + 1] = g + p;
        g = c * r - b;
        for (int k = 0; k < rows; ++k) {  // rotate columns i, i+1
          double* row = &V[(size_t)k * n];
          const double t = row[i + 1];
          row[i + 1] = s * row[i] + c * t;
          row[i] = c * row[i] - s * t;
        }
      }
      if (deflated) continue;
      d[l] -= p;
      e[l] = g;
      e[m] = 0.0;
    }
  }
  std::vector<int> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return d[a] < d[b]; });
  for (int j = 0; j < n; ++j) w[j] = d[ord[j]];
  if (z1)
    for (int j = 0; j < n; ++j) z1[j] = V[ord[j]];  // row 0
  if (full)
    for (int j = 0; j < n; ++j)
      for (int k = 0; k < n; ++k) Z[(size_t)j * n + k] = V[(size_t)k * n + ord[j]];
}

}  // namespace edl
