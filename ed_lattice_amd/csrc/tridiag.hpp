#pragma once
#include <limits>
#include <vector>

namespace edl {
void tridiag_eig(int n, const double* d, const double* e, double* w, double* Z, double* z1);
}
