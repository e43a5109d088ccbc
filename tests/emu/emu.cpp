// CPU build of the product's set-up code and per-system solver
// (channel_passive_amd/csrc/bspline.cpp, solver.h), used only by tests to
// check the kernel arithmetic bit-for-bit against the oracle without a GPU.
// The product never loads this library.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../channel_passive_amd/csrc/bspline.h"
#include "../../channel_passive_amd/csrc/solver.h"
#include "../../channel_passive_amd/csrc/solver_adj.h"
#include "../../include/pb_capi.h"

namespace {
struct Setup {
  int n, bw, W, order;
  pb::Collocation col;
  std::vector<double> d0, d1, d2, d0lu;
};
int make_setup(const pb_config *cfg, Setup &S) {
  S.order = cfg->labs_k + 1;
  S.n = (int)cfg->ny;
  if (pb::build_collocation(S.order, S.n, cfg->knot_map, cfg->stretch, 2.0 * cfg->lmfy, S.col)) return -1;
  S.bw = pb::eval_bandwidth(S.col);
  S.W = 2 * S.bw + 1;
  pb::to_band(S.col.D[0], S.n, S.bw, S.d0);
  pb::to_band(S.col.D[1], S.n, S.bw, S.d1);
  pb::to_band(S.col.D[2], S.n, S.bw, S.d2);
  S.d0lu = S.d0;
  pb::OwnBand ob{S.d0lu.data(), 1, S.W};
  pb::band_lu(pb::BandGeom{S.n, S.bw}, ob);
  return 0;
}
pb::LineParams params(const pb_config *cfg, const Setup &S, int s, double dt) {
  pb::LineParams P;
  P.g = pb::BandGeom{S.n, S.bw};
  P.D0 = pb::SharedBand{S.d0.data(), S.W};
  P.D1 = pb::SharedBand{S.d1.data(), S.W};
  P.D2 = pb::SharedBand{S.d2.data(), S.W};
  P.D0LU = pb::SharedBand{S.d0lu.data(), S.W};
  P.integ = S.col.integ.data();
  pb::RK3 rk = pb::rk3_coefficients();
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 3; ++b) P.rk[a][b] = rk.c[a][b];
  P.dt = dt; P.nu = cfg->nu; P.pr = cfg->pr; P.bulk_u = cfg->bulk_u; P.bulk_w = cfg->bulk_w;
  P.bulk_t = cfg->bulk_t; P.u_gal = cfg->u_gal; P.s = s; P.S = cfg->rk_step;
  return P;
}
}  // namespace

extern "C" {
int emu_setup(const pb_config *cfg, int *bw, double *y, double *d0, double *d1, double *d2, double *d0lu,
              double *integ) {
  Setup S;
  if (make_setup(cfg, S)) return -1;
  *bw = S.bw;
  const size_t nb = (size_t)S.n * S.W;
  if (y) std::memcpy(y, S.col.y.data(), S.n * 8);
  if (d0) std::memcpy(d0, S.d0.data(), nb * 8);
  if (d1) std::memcpy(d1, S.d1.data(), nb * 8);
  if (d2) std::memcpy(d2, S.d2.data(), nb * 8);
  if (d0lu) std::memcpy(d0lu, S.d0lu.data(), nb * 8);
  if (integ) std::memcpy(integ, S.col.integ.data(), S.n * 8);
  return 0;
}

// U: 9 slots, RK: 3 slots, device layout [slot][y][kz][kx]; vd: 2 planes or
// NULL; usrc: u of the substep's input state (one field) under the mixed-type
// scalar condition, else NULL; act: the wall actuation (one plane) or NULL
int emu_wavespace_act(const pb_config *cfg, int s, double dt, double *U, double *RK, const double *vd,
                      const double *usrc, const double *act, double *mean4) {
  Setup S;
  if (make_setup(cfg, S)) return -1;
  pb::LineParams P = params(cfg, S, s, dt);
  const int nxh = (int)cfg->nx / 2, nz = (int)cfg->nz;
  const size_t plane = (size_t)nz * nxh, field = plane * S.n;
  std::vector<pb::cplx> buf(2 * field);
  std::vector<double> dbuf((size_t)S.W * field);
  pb::WsArgs A;
  A.U = (pb::cplx *)U; A.RK = (pb::cplx *)RK; A.buf = buf.data(); A.dbuf = dbuf.data(); A.rbuf = nullptr;
  A.xcol = nullptr;
  A.vd = (const pb::cplx *)vd; A.plane = plane; A.field = field; A.nxh = nxh; A.nz = nz; A.nh = nz / 2;
  A.lmfxr = 1.0 / cfg->lmfx; A.lmfzr = 1.0 / cfg->lmfz; A.opp_alpha = cfg->opp_alpha; A.mean_out = mean4;
  A.usrc = cfg->scalar_bc == PB_SCALAR_MIXED ? (const pb::cplx *)usrc : nullptr;
  A.src_scale = cfg->nu / cfg->pr / cfg->bulk_u;
  A.act = (const pb::cplx *)act;
  for (size_t sys = 0; sys < plane; ++sys) pb::wavespace_line(P, A, sys);
  return 0;
}

int emu_wavespace(const pb_config *cfg, int s, double dt, double *U, double *RK, const double *vd,
                  const double *usrc, double *mean4) {
  return emu_wavespace_act(cfg, s, dt, U, RK, vd, usrc, nullptr, mean4);
}

// Adjoint of the wall-normal stage (solver_adj.h), device layout.  On entry
// AU slots 0,1,2,6 hold the adjoints of u, v, w, theta out and ARK those of
// the new stores; on return AU slots 0-4, 6-8 the adjoints of the products,
// slot 5 that of the mixed-type source u, ARK those of the old stores.
// vd: the forward's detection-plane v (2 planes) or NULL; avd (2 planes),
// aact (1 plane, accumulated) and aalpha (1 plane) may be NULL.
int emu_adj_wavespace(const pb_config *cfg, int s, double dt, double *AU, double *ARK, const double *vd, double *avd,
                      double *aact, double *aalpha, double tau_bar, double q_bar) {
  Setup S;
  if (make_setup(cfg, S)) return -1;
  pb::LineParams P = params(cfg, S, s, dt);
  const int nxh = (int)cfg->nx / 2, nz = (int)cfg->nz;
  const size_t plane = (size_t)nz * nxh, field = plane * S.n;
  if (S.bw < 3 || S.bw > 5) return -2;
  std::vector<pb::cplx> abuf(pb::ADJ_SCRATCH * field);
  std::vector<double> fac((size_t)(3 * S.W + 4) * field);
  pb::AdjFactors F;
  F.luh = fac.data(); F.lut = F.luh + (size_t)S.W * field; F.lup = F.lut + (size_t)S.W * field;
  F.hom = F.lup + (size_t)S.W * field; F.plane = plane; F.field = field; F.W = S.W;
  for (size_t sys = 0; sys < plane; ++sys)
    pb::adj_factor_line(P, F, sys, nxh, nz / 2, 1.0 / cfg->lmfx, 1.0 / cfg->lmfz, pb::SV{abuf.data() + sys, plane},
                        pb::SV{abuf.data() + field + sys, plane});
  pb::AdjArgs A;
  A.AU = (pb::cplx *)AU; A.ARK = (pb::cplx *)ARK; A.abuf = abuf.data(); A.F = F;
  A.vd = (const pb::cplx *)vd; A.avd = (pb::cplx *)avd; A.aact = (pb::cplx *)aact; A.aalpha = aalpha;
  A.mixed = cfg->scalar_bc == PB_SCALAR_MIXED; A.ks = cfg->nu / cfg->pr / cfg->bulk_u;
  A.tau_bar = tau_bar; A.q_bar = q_bar;
  A.plane = plane; A.field = field; A.nxh = nxh; A.nz = nz; A.nh = nz / 2;
  A.lmfxr = 1.0 / cfg->lmfx; A.lmfzr = 1.0 / cfg->lmfz; A.opp_alpha = cfg->opp_alpha;
  for (size_t sys = 0; sys < plane; ++sys) {
    if (S.bw == 3) pb::adj_wavespace_line<3>(P, A, sys);
    if (S.bw == 4) pb::adj_wavespace_line<4>(P, A, sys);
    if (S.bw == 5) pb::adj_wavespace_line<5>(P, A, sys);
  }
  return 0;
}
}

#include "../../channel_passive_amd/csrc/solver_fused.h"
template <int BW>
static void run_id(int id, const pb::LineParams &P, const pb::WsArgs &A, double *dbuf2, size_t sys) {
  switch (id) {
    case pb::SW_A1: pb::run_sweep<BW, pb::SW_A1>(P, A, dbuf2, sys); break;
    case pb::SW_BG1: pb::run_sweep<BW, pb::SW_BG1>(P, A, dbuf2, sys); break;
    case pb::SW_BT1: pb::run_sweep<BW, pb::SW_BT1>(P, A, dbuf2, sys); break;
    case pb::SW_A: pb::run_sweep<BW, pb::SW_A>(P, A, dbuf2, sys); break;
    case pb::SW_AT: pb::run_sweep<BW, pb::SW_AT>(P, A, dbuf2, sys); break;
    case pb::SW_BT: pb::run_sweep<BW, pb::SW_BT>(P, A, dbuf2, sys); break;
    case pb::SW_CT: pb::run_sweep<BW, pb::SW_CT>(P, A, dbuf2, sys); break;
    case pb::SW_DT: pb::run_sweep<BW, pb::SW_DT>(P, A, dbuf2, sys); break;
    case pb::SW_BG: pb::run_sweep<BW, pb::SW_BG>(P, A, dbuf2, sys); break;
    case pb::SW_BV: pb::run_sweep<BW, pb::SW_BV>(P, A, dbuf2, sys); break;
    case pb::SW_CNU: pb::run_sweep<BW, pb::SW_CNU>(P, A, dbuf2, sys); break;
    case pb::SW_DNU: pb::run_sweep<BW, pb::SW_DNU>(P, A, dbuf2, sys); break;
    case pb::SW_E: pb::run_sweep<BW, pb::SW_E>(P, A, dbuf2, sys); break;
    case pb::SW_F: pb::run_sweep<BW, pb::SW_F>(P, A, dbuf2, sys); break;
    case pb::SW_G: pb::run_sweep<BW, pb::SW_G>(P, A, dbuf2, sys); break;
  }
}

extern "C" int emu_wavespace_fused_act(const pb_config *cfg, int s, double dt, double *U, double *RK,
                                       const double *vd, const double *usrc, const double *act, double *mean4) {
  Setup S;
  if (make_setup(cfg, S)) return -1;
  pb::LineParams P = params(cfg, S, s, dt);
  const int nxh = (int)cfg->nx / 2, nz = (int)cfg->nz;
  const size_t plane = (size_t)nz * nxh, field = plane * S.n;
  std::vector<pb::cplx> buf(2 * field);
  const size_t nfac = std::max((size_t)S.W * field, pb::ublk_elems(plane, S.n, S.W));
  std::vector<double> dbuf(nfac), dbuf2(nfac), rbuf(field);
  std::vector<pb::cplx> xcol(field);
  pb::WsArgs A;
  A.U = (pb::cplx *)U; A.RK = (pb::cplx *)RK; A.buf = buf.data(); A.dbuf = dbuf.data(); A.rbuf = rbuf.data();
  A.xcol = xcol.data();
  A.vd = (const pb::cplx *)vd; A.plane = plane; A.field = field; A.nxh = nxh; A.nz = nz; A.nh = nz / 2;
  A.lmfxr = 1.0 / cfg->lmfx; A.lmfzr = 1.0 / cfg->lmfz; A.opp_alpha = cfg->opp_alpha; A.mean_out = mean4;
  A.usrc = cfg->scalar_bc == PB_SCALAR_MIXED ? (const pb::cplx *)usrc : nullptr;
  A.src_scale = cfg->nu / cfg->pr / cfg->bulk_u;
  A.act = (const pb::cplx *)act;
  if (S.bw < 3 || S.bw > 5) return -2;
  pb::wavespace_line(P, A, 0);  // (0,0): mean modes
  // sweep-major, as the device runs them: each sweep over all systems
  int ids[pb::SW_COUNT];
  const int nid = pb::sweep_sequence(s, ids);
  for (int q = 0; q < nid; ++q)
    for (size_t idx = 0; idx < pb::sweep_threads(ids[q], A); ++idx) {
      if (S.bw == 3) run_id<3>(ids[q], P, A, dbuf2.data(), idx);
      if (S.bw == 4) run_id<4>(ids[q], P, A, dbuf2.data(), idx);
      if (S.bw == 5) run_id<5>(ids[q], P, A, dbuf2.data(), idx);
    }
  return 0;
}
extern "C" int emu_wavespace_fused(const pb_config *cfg, int s, double dt, double *U, double *RK, const double *vd,
                                   const double *usrc, double *mean4) {
  return emu_wavespace_fused_act(cfg, s, dt, U, RK, vd, usrc, nullptr, mean4);
}

// Opposition-control detection planes with the device's per-system code.
// V: the v field in the device layout [y][kz][kx]; out: 2 planes [kz][kx].
extern "C" int emu_opp_planes(const pb_config *cfg, double *V, double *out) {
  Setup S;
  if (make_setup(cfg, S)) return -1;
  std::vector<double> lo, hi;
  if (pb::detection_row(S.col, S.d0lu, S.bw, -1.0 + cfg->opp_ydist, lo) ||
      pb::detection_row(S.col, S.d0lu, S.bw, 1.0 - cfg->opp_ydist, hi))
    return -3;
  lo.insert(lo.end(), hi.begin(), hi.end());
  const size_t plane = (size_t)(cfg->nz) * (cfg->nx / 2);
  pb::cplx *Vc = (pb::cplx *)V, *o = (pb::cplx *)out;
  for (size_t sys = 0; sys < plane; ++sys)
    pb::opp_dots(pb::SV{Vc + sys, plane}, lo.data(), S.n, o + sys, o + plane + sys);
  return 0;
}

// Statistics-time pressure with the product's per-system code, device
// layout: V = v field [y][kz][kx]; U = 9 product slots (slot 5: v*v in, p
// out); vv = <vv>(y).
extern "C" int emu_pressure(const pb_config *cfg, double *V, double *U, double *vv) {
  Setup S;
  if (make_setup(cfg, S)) return -1;
  pb::LineParams P = params(cfg, S, 1, 0.0);
  const int nxh = (int)cfg->nx / 2, nz = (int)cfg->nz, nh = nz / 2;
  const size_t plane = (size_t)nz * nxh, field = plane * S.n;
  pb::cplx *Vc = (pb::cplx *)V, *Uc = (pb::cplx *)U;
  for (int y = 0; y < S.n; ++y) vv[y] = pb::mean_vv_row(Vc + (size_t)y * plane, nxh, plane);
  std::vector<pb::cplx> b1(field), b2(field);
  std::vector<double> db((size_t)S.W * field);
  for (size_t sys = 0; sys < plane; ++sys) {
    const int k = (int)(sys % nxh), j = (int)(sys / nxh);
    const int kzi = (j <= nh) ? j : j + 1 - 2 * nh - 1;
    auto col = [&](int slot) { return pb::SV{Uc + (size_t)slot * field + sys, plane}; };
    const pb::SV t1{b1.data() + sys, plane}, t2{b2.data() + sys, plane};
    if (k == 0 && kzi == 0) {
      for (int y = 0; y < S.n; ++y) col(5)[y] = pb::mk(-vv[y], 0.0);
      continue;
    }
    pb::cplx bc1, bc2;
    pb::ppe_wall_values(P, pb::SV{Vc + sys, plane}, t1, cfg->nu, &bc1, &bc2);
    const pb::OwnBand Db{db.data() + sys, plane, S.W};
    pb::solve_ppe(P, Db, col(0), col(1), col(2), col(3), col(4), col(5), bc1, bc2, t1, t2, k * (1.0 / cfg->lmfx),
                  kzi * (1.0 / cfg->lmfz));
  }
  return 0;
}

// get_statistics_1 profiles with the product's code, device layout U
// (9 slots; u, v, w, theta in slots 0, 1, 2, 6); out: 16*ny + 2 doubles.
extern "C" int emu_statistics(const pb_config *cfg, double *U, double *out) {
  Setup S;
  if (make_setup(cfg, S)) return -1;
  pb::LineParams P = params(cfg, S, 1, 0.0);
  const int nxh = (int)cfg->nx / 2, n = S.n;
  const size_t plane = (size_t)cfg->nz * nxh, field = plane * n;
  pb::cplx *Uc = (pb::cplx *)U;
  const int pairs[10][2] = {{0, 0}, {1, 1}, {2, 2}, {6, 6}, {0, 1}, {0, 2}, {1, 2}, {0, 6}, {6, 2}, {1, 6}};
  for (int q = 0; q < 10; ++q)
    for (int y = 0; y < n; ++y)
      out[(size_t)(6 + q) * n + y] = pb::mean_uv_row(Uc + pairs[q][0] * field + (size_t)y * plane,
                                                      Uc + pairs[q][1] * field + (size_t)y * plane, nxh, plane);
  std::vector<double> cen(n, 0.0), vals;
  int first = 0;
  if (pb::basis_row(S.col, 0.0, first, vals)) return -3;
  for (size_t j = 0; j < vals.size(); ++j) cen[first + j] = vals[j];
  std::vector<pb::cplx> t(2 * field);
  const pb::SV m[3] = {pb::SV{Uc, plane}, pb::SV{Uc + 2 * field, plane}, pb::SV{Uc + 6 * field, plane}};
  pb::mean_line_stats(P, m, pb::SV{t.data(), plane}, pb::SV{t.data() + field, plane}, cen.data(), cfg->u_gal, out,
                      out + (size_t)16 * n);
  return 0;
}
