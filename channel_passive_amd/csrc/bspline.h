// bspline.h -- host-side B-spline collocation operators (see bspline.cpp).
#pragma once

#include <vector>

namespace pb {

struct Collocation {
  int n = 0, order = 0;
  std::vector<double> brk;    // breakpoints
  std::vector<double> y;      // Greville collocation points
  std::vector<double> D[3];   // row-major n x n: value, d/dy, d2/dy2 of basis c at point r
  std::vector<double> integ;  // integration weights of the coefficients
};

std::vector<double> breakpoints(int nbreak, int knot_map, double stretch, double ly);
int build_collocation(int order, int n, int knot_map, double stretch, double ly, Collocation &out);
int basis_row(const Collocation &c, double y, int &first, std::vector<double> &vals);
int eval_bandwidth(const Collocation &c);
void to_band(const std::vector<double> &full, int n, int bw, std::vector<double> &band);
// r with r . v = B(y) . LU(D0)^-1 v for every column v: the transposed
// solve (LU)^T r = b of the packed Crout factor (L with its diagonal, unit
// U), b = the basis values at y in their columns
int detection_row(const Collocation &c, const std::vector<double> &d0lu, int bw, double y, std::vector<double> &r);

}  // namespace pb
