// capi.hip -- implementation of include/pb_capi.h on one gfx950 GPU.
//
// Replaces the body of time_march (SOURCE_PASSOPB/time_march.f90:36-128):
// the MPI pencil transposes collapse into device-layout choices, every
// stage runs as kernels on one HIP stream and the state never leaves HBM
// between pb_set_state and pb_get_state.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <chrono>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pb_capi.h"
#include "bspline.h"
#include "kernels.hip"

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(call)                                                                          \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess) return fail(PB_E_HIP, "%s: %s", #call, hipGetErrorString(e_));      \
  } while (0)

pb::FftPlan make_fft_plan(int N) {
  pb::FftPlan p{};
  p.N = N;
  p.np = 0;
  int m = N;
  for (int r : {8, 4, 2, 3, 5})
    while (m % r == 0 && p.np < 16) { p.r[p.np++] = r; m /= r; }
  if (m != 1) p.np = -1;
  return p;
}

std::vector<double2> twiddles(int N) {
  std::vector<double2> t(N);
  const long double pi = 3.14159265358979323846264338327950288L;
  for (int j = 0; j < N; ++j) {
    const long double a = 2.0L * pi * (long double)j / (long double)N;
    t[j] = make_double2((double)cosl(a), -(double)sinl(a));
  }
  return t;
}

}  // namespace

struct pb_ctx {
  pb_config cfg;
  int nx, ny, nz, nxh, nx32, nz32, nh, bw, W, order;
  size_t plane, field, pplane, pfield, nsys;
  double Lx, Lz, dxr, dzr, cflx, rescale, lmfxr, lmfzr;
  pb::Collocation col;
  std::vector<double> d0, d1, d2, d0lu, dyr;
  std::vector<double> opp_rows;  // the two detection rows, ny doubles each
  pb::FftPlan pz, px;
  int Bz;
  hipStream_t stream = nullptr;
  hipStream_t side = nullptr;           // the (0,0) mean line runs here, beside the sweeps
  hipStream_t side2 = nullptr;          // the theta chain of the sweeps
  hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_fork2 = nullptr, ev_join2 = nullptr;
  // device memory
  double2 *U = nullptr, *RK = nullptr, *pad = nullptr, *buf = nullptr, *vd = nullptr;
  double2 *usrc = nullptr;  // mixed-type scalar condition: u of the substep's input state
  hipGraph_t step_graph = nullptr;         // pb_step_graph: one step, captured on the first call
  hipGraphExec_t step_exec = nullptr;
  // enqueue everything on `stream` (no side-stream forks): for graphs that
  // must run beside other graphs (a graph's own branches share the
  // runtime's internal streams with every other graph's)
  bool linear = false;
  // get_statistics_2 (allocated on its first use): copies of u, v, w of the
  // statistics substep's input, the moment partials of the x pass (set in
  // mom_active while that pass runs), their sums, the Parseval sums, the 29
  // output lines and two ny-long complex scratch columns
  double2 *stat_uvw = nullptr;
  double *dmom = nullptr, *dmsum = nullptr, *dpv = nullptr, *dstat2 = nullptr, *mom_active = nullptr;
  double2 *dstat2_tmp = nullptr;
  // wall actuation (pb_set_control): one complex value per system, added to
  // the wall v of both walls; used when act_on
  double2 *act = nullptr;
  bool act_on = false;
  // discrete adjoint (allocated on first use): adjoints of the 9 slots, of
  // the stores, 6 scratch columns per system, the padded product / state
  // adjoints (8 fields), of vd (2 planes), of the actuation (accumulated)
  // and per-system d/d alpha terms (accumulated); adj_states: what the
  // gradient driver saves (checkpoints, one segment's substep inputs),
  // adj_nstates double2 elements
  double2 *adj_U = nullptr, *adj_RK = nullptr, *adj_buf = nullptr, *adj_pad = nullptr, *adj_vd = nullptr,
          *adj_act = nullptr, *adj_states = nullptr;
  double *adj_alpha = nullptr;
  size_t adj_nstates = 0;
  // the transposed substeps' matrices for s = 1..3 at dt adj_fac_dt:
  // [luh_s, lut_s (W fields each), hom_s (4 fields)] x 3, then lup (W fields)
  double *adj_fac = nullptr;
  double adj_fac_dt = 0.0;
  double *dbuf = nullptr, *dbuf2 = nullptr, *rbuf = nullptr, *vpart = nullptr, *scal = nullptr;
  double2 *ppe_bc = nullptr;  // statistics time: pressure wall values [2][kz][kx]
  double *dcenter = nullptr;  // B_k(0): centre-line values of the basis (get_statistics.f90:566-570)
  double *dstat = nullptr;    // 16 profiles of ny + 2 centre values
  double *dvv = nullptr;      // statistics time: <vv>(y)
  size_t vpart_n = 0;         // per-wave CFL maxima written by the last x-physical launch
  size_t mean_lds = 0;        // dynamic LDS of mean_kernel (0: one-thread wave00_kernel)
  double *mean_db = nullptr;  // mean_kernel band scratch when it does not fit in LDS
  double *dD0 = nullptr, *dD1 = nullptr, *dD2 = nullptr, *dD0LU = nullptr, *dinteg = nullptr, *ddyr = nullptr,
         *drows = nullptr;
  double2 *twz = nullptr, *twx = nullptr;
  int *slotmap = nullptr;
  double *hscal = nullptr;  // pinned host mirror of scal
  size_t bytes = 0;
  bool have_state = false;
  bool fused = true;  // fused-sweep wall-normal solver (bit-identical to the statement-by-statement port)
  bool ct_fft = true;  // fixed-size transforms where one exists for the grid
  bool ct_z_ok = false, ct_x_ok = false;
  bool zs_on = true;   // streamed (double-buffered, one block per CU) z transforms where they apply
  bool zs_ok = false;
  bool zs_inv = true, zs_fwd = true;
  bool xs_on = true;   // streamed x pass (one block per CU, next line copied to LDS) where it applies
  bool xs_ok = false;
  int ncu = 0;         // compute units: grid of the streamed transforms
  pb::RK3 rk;
  // per-stage event timing
  bool timing = false;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  std::vector<std::pair<int, size_t>> ev_marks;  // (stage, index of start event)

  pb::LineParams line_params(int s, double dt) const {
    pb::LineParams P;
    P.g = pb::BandGeom{ny, bw};
    P.D0 = pb::SharedBand{dD0, W};
    P.D1 = pb::SharedBand{dD1, W};
    P.D2 = pb::SharedBand{dD2, W};
    P.D0LU = pb::SharedBand{dD0LU, W};
    P.integ = dinteg;
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 3; ++b) P.rk[a][b] = rk.c[a][b];
    P.dt = dt;
    P.nu = cfg.nu;
    P.pr = cfg.pr;
    P.bulk_u = cfg.bulk_u;
    P.bulk_w = cfg.bulk_w;
    P.bulk_t = cfg.bulk_t;
    P.u_gal = cfg.u_gal;
    P.s = s;
    P.S = cfg.rk_step;
    return P;
  }
};

using pb::SCAL_N;
// padded z-pencil fields: 4 state fields in, 8 products out, + v*v at statistics time
static constexpr int kPadFields = 9;

extern "C" {

const char *pb_last_error(void) { return g_err.c_str(); }

static int ctx_free(pb_ctx *c) {
  if (!c) return 0;
  void *ptrs[] = {c->U, c->RK, c->pad, c->buf, c->vd, c->dbuf, c->dbuf2, c->rbuf, c->vpart, c->scal, c->dD0, c->dD1, c->dD2,
                  c->dD0LU, c->dinteg, c->ddyr, c->drows, c->twz, c->twx, c->slotmap, c->mean_db,
                  c->ppe_bc, c->dvv, c->dcenter, c->dstat, c->usrc, c->stat_uvw, c->dmom, c->dmsum, c->dpv,
                  c->dstat2, c->dstat2_tmp, c->act, c->adj_U, c->adj_RK, c->adj_buf, c->adj_pad, c->adj_vd,
                  c->adj_act, c->adj_states, c->adj_alpha, c->adj_fac};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (c->hscal) (void)hipHostFree(c->hscal);
  if (c->step_exec) (void)hipGraphExecDestroy(c->step_exec);
  if (c->step_graph) (void)hipGraphDestroy(c->step_graph);
  for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  if (c->ev_fork2) (void)hipEventDestroy(c->ev_fork2);
  if (c->ev_join2) (void)hipEventDestroy(c->ev_join2);
  if (c->side2) (void)hipStreamDestroy(c->side2);
  if (c->side) (void)hipStreamDestroy(c->side);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

int pb_destroy(pb_ctx *ctx) { return ctx_free(ctx); }

int pb_create(const pb_config *cfg, pb_ctx **out) {
  if (!cfg || !out) return fail(PB_E_ARG, "null argument");
  *out = nullptr;
  if (cfg->nx < 4 || cfg->nz < 4 || cfg->nx % 2 || cfg->nz % 2 || cfg->ny < 8)
    return fail(PB_E_ARG, "bad grid %lld x %lld x %lld", (long long)cfg->nx, (long long)cfg->ny, (long long)cfg->nz);
  if ((cfg->nx * 3) % 2 || (cfg->nz * 3) % 2) return fail(PB_E_ARG, "nx, nz must give integer 3/2 sizes");
  if (cfg->rk_step != 3) return fail(PB_E_ARG, "only the 3-substep scheme is implemented (rk_step=%d)", cfg->rk_step);
  pb_ctx *c = new pb_ctx();
  c->cfg = *cfg;
  c->nx = (int)cfg->nx; c->ny = (int)cfg->ny; c->nz = (int)cfg->nz;
  c->nxh = c->nx / 2; c->nx32 = c->nx * 3 / 2; c->nz32 = c->nz * 3 / 2; c->nh = c->nz / 2;
  c->order = cfg->labs_k + 1;
  const double pi = 3.14159265358979323;  // module_global_variable.f90:18
  c->Lx = cfg->lmfx * 2 * pi;
  c->Lz = cfg->lmfz * 2 * pi;
  c->rescale = 1.0 / (double)(c->nx32 * c->nz32);
  c->dxr = (double)c->nx / c->Lx;
  c->dzr = (double)c->nz / c->Lz;
  c->cflx = cfg->cfl / pi;
  c->lmfxr = 1.0 / cfg->lmfx;
  c->lmfzr = 1.0 / cfg->lmfz;
  c->rk = pb::rk3_coefficients();
  if (const char *e = getenv("PB_WAVESPACE_LITERAL")) c->fused = (e[0] == '0');
  if (const char *e = getenv("PB_FFT_GENERIC")) c->ct_fft = (e[0] == '0');
  if (const char *e = getenv("PB_Z_STREAM")) {
    c->zs_on = (e[0] != '0');
    c->zs_inv = (e[0] != 'f');
    c->zs_fwd = (e[0] != 'i');
  }
  if (const char *e = getenv("PB_X_STREAM")) c->xs_on = (e[0] != '0');
  int r = pb::build_collocation(c->order, c->ny, cfg->knot_map, cfg->stretch, 2.0 * cfg->lmfy, c->col);
  if (r) { ctx_free(c); return fail(PB_E_ARG, "B-spline set-up failed (%d): check ny, labs_k, knots", r); }
  c->bw = pb::eval_bandwidth(c->col);
  if (c->bw < 1) { ctx_free(c); return fail(PB_E_ARG, "no band structure found"); }
  c->W = 2 * c->bw + 1;
  pb::to_band(c->col.D[0], c->ny, c->bw, c->d0);
  pb::to_band(c->col.D[1], c->ny, c->bw, c->d1);
  pb::to_band(c->col.D[2], c->ny, c->bw, c->d2);
  c->d0lu = c->d0;
  {
    pb::OwnBand ob{c->d0lu.data(), 1, c->W};
    pb::band_lu(pb::BandGeom{c->ny, c->bw}, ob);
  }
  c->dyr.resize(c->ny);
  const std::vector<double> &y = c->col.y;
  for (int i = 0; i < c->ny; ++i) {
    double d;
    if (i == 0) d = std::fabs(y[1] - y[0]);
    else if (i == c->ny - 1) d = std::fabs(y[c->ny - 1] - y[c->ny - 2]);
    else d = std::min(std::fabs(y[i] - y[i - 1]), std::fabs(y[i] - y[i + 1]));
    c->dyr[i] = 1.0 / (cfg->dy_factor * d);
  }
  if (cfg->scalar_bc != PB_SCALAR_SOURCE && cfg->scalar_bc != PB_SCALAR_MIXED) {
    ctx_free(c);
    return fail(PB_E_ARG, "scalar_bc %d is neither PB_SCALAR_SOURCE nor PB_SCALAR_MIXED", cfg->scalar_bc);
  }
  // detection rows r with v(y_d) = r . v (= B(y_d) . LU(D0)^-1 v): needed
  // with control; also set up without it when the plane lies inside the
  // channel (pb_set_control may switch control on, the adjoint's d/d alpha
  // needs v(y_d))
  if (cfg->opp_alpha != 0.0 || (cfg->opp_ydist > 0.0 && cfg->opp_ydist < 1.0)) {
    std::vector<double> lo, hi;
    if (cfg->opp_ydist <= 0.0 || cfg->opp_ydist >= 1.0 ||
        pb::detection_row(c->col, c->d0lu, c->bw, -1.0 + cfg->opp_ydist, lo) ||
        pb::detection_row(c->col, c->d0lu, c->bw, 1.0 - cfg->opp_ydist, hi)) {
      if (cfg->opp_alpha != 0.0) {
        ctx_free(c);
        return fail(PB_E_ARG, "opposition-control plane %g outside the channel", cfg->opp_ydist);
      }
    } else {
      c->opp_rows = lo;
      c->opp_rows.insert(c->opp_rows.end(), hi.begin(), hi.end());
    }
  }
  c->pz = make_fft_plan(c->nz32);
  c->px = make_fft_plan(c->nx32);
  if (c->pz.np < 0 || c->px.np < 0) {
    ctx_free(c);
    return fail(PB_E_ARG, "3/2 sizes %d, %d must factor into 2,3,5", c->nx32, c->nz32);
  }
  if (2 * c->nx32 > pb::FFT_MAX_ELEMS || c->nz32 * 1 > pb::FFT_MAX_ELEMS) {
    ctx_free(c);
    return fail(PB_E_ARG, "transform lengths exceed the LDS tile limit");
  }
  int dev = 0, lds_max = 65536;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) {
    ctx_free(c);
    return fail(PB_E_HIP, "no HIP device");
  }
  const size_t lds_cap = std::min<size_t>((size_t)lds_max, 96 * 1024);
  if ((size_t)c->nz32 * sizeof(double2) > lds_cap || (size_t)2 * c->nx32 * sizeof(double2) > lds_cap) {
    ctx_free(c);
    return fail(PB_E_ARG, "transform lengths exceed the LDS limit (%d bytes)", lds_max);
  }
  // z tiles: B consecutive kx columns per block (B*16 bytes contiguous per z row)
  c->Bz = 1;
  while (c->Bz * 2 <= 16 && c->nxh % (c->Bz * 2) == 0 && (size_t)c->nz32 * c->Bz * 2 * sizeof(double2) <= lds_cap &&
         c->nz32 * c->Bz * 2 <= 4608)
    c->Bz *= 2;
  // fixed-size z transforms bring their own tile width and LDS size
  {
    using FZ = pb::FftZ1152;
    const int lz = FZ::N * FZ::B * (int)sizeof(double2);
    c->ct_z_ok = c->ct_fft && c->nz32 == FZ::N && c->nxh % FZ::B == 0 && lz <= lds_max &&
                 hipFuncSetAttribute(reinterpret_cast<const void *>(&pb::zinv_kernel<FZ>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, lz) == hipSuccess &&
                 hipFuncSetAttribute(reinterpret_cast<const void *>(&pb::zfwd_kernel<FZ>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, lz) == hipSuccess;
    if (c->ct_z_ok) c->Bz = FZ::B;
    // the streamed z transforms: same passes and twiddles on 4-column tiles,
    // two tile buffers in static LDS
    const int lzs = (2 * pb::zs::NE + pb::zs::NTW) * (int)sizeof(double2);
    if (hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) c->ncu = 0;
    c->zs_ok = c->zs_on && c->ct_z_ok && c->nxh % pb::zs::B == 0 && lzs <= lds_max && c->ncu > 0;
    using FX = pb::FftX1536;
    c->ct_x_ok = c->ct_fft && c->nx32 == FX::N && FX::L::SIZE * (int)sizeof(double2) <= lds_max;
    // the streamed x pass: 4 transform lines, the 4 input rows and the twiddles in static LDS
    const int lxs = (pb::xs::L4::SIZE + 4 * pb::xs::NH + pb::xs::NTW) * (int)sizeof(double2);
    c->xs_ok = c->xs_on && c->ct_x_ok && c->nxh == pb::xs::NH && lxs <= lds_max && c->ncu > 0 &&
               (size_t)c->ncu <= (size_t)c->ny * c->nz32 * 16;  // one vpart entry per block
  }
  c->plane = (size_t)c->nz * c->nxh;
  c->field = c->plane * c->ny;
  c->pplane = (size_t)c->nz32 * c->nxh;
  c->pfield = c->pplane * c->ny;
  c->nsys = c->plane;

  auto dalloc = [&](void **p, size_t bytes) -> int {
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return fail(PB_E_HIP, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    c->bytes += bytes;
    return 0;
  };
  const size_t bandb = (size_t)c->ny * c->W * sizeof(double);
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&c->side2, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_fork2, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_join2, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess) {
    ctx_free(c);
    return fail(PB_E_HIP, "no HIP device / stream");
  }
  // Placement of a large context: while the context allocates, a block of up
  // to 160 GiB (as much as leaves room for the context) is held and then
  // released, so that the arrays lie above the device's first free region.
  // Measured on fresh devices at 1024x385x768, the first process's
  // wall-normal stage: 36.0-36.1 ms without the block (five of five), 33.6-
  // 33.7 ms with 100 GiB (three of three), 32.1 and 32.7 ms with 160 and 180
  // GiB; later processes 31.7-35.5 ms.  The sweeps walk columns whose rows
  // are one plane apart, every row a new memory page, and are bound by
  // address translation as much as by bandwidth, so the physical placement
  // matters.  PB_RESERVE=<GiB> caps the block (0: off).
  void *reserve = nullptr;
  {
    const size_t est = (size_t)40 * c->field * sizeof(double2), GiB = (size_t)1 << 30;
    const char *e = getenv("PB_RESERVE");
    const size_t cap = (e ? (size_t)std::strtoull(e, nullptr, 10) : 160) * GiB;
    size_t fr = 0, tot = 0;
    if (est >= 32 * GiB && cap > 0 && hipMemGetInfo(&fr, &tot) == hipSuccess && fr > est + 72 * GiB) {
      const size_t want = std::min(cap, (fr - est - 8 * GiB) / GiB * GiB);
      if (hipMalloc(&reserve, want) != hipSuccess) {
        reserve = nullptr;
        (void)hipGetLastError();
      }
    }
  }
  int rc = 0;
  rc = rc ? rc : dalloc((void **)&c->U, 9 * c->field * sizeof(double2));
  rc = rc ? rc : dalloc((void **)&c->RK, 3 * c->field * sizeof(double2));
  rc = rc ? rc : dalloc((void **)&c->pad, kPadFields * c->pfield * sizeof(double2));
  rc = rc ? rc : dalloc((void **)&c->buf, 2 * c->field * sizeof(double2));
  rc = rc ? rc : dalloc((void **)&c->vd, 2 * c->plane * sizeof(double2));
  if (cfg->scalar_bc == PB_SCALAR_MIXED) rc = rc ? rc : dalloc((void **)&c->usrc, c->field * sizeof(double2));
  // U factors: W per row and system, in the plane layout (statement-by-statement
  // solver) or in blocks of 64 systems (fused sweeps), whichever is larger
  const size_t nfac = std::max((size_t)c->W * c->field, pb::ublk_elems(c->plane, c->ny, c->W));
  rc = rc ? rc : dalloc((void **)&c->dbuf, nfac * sizeof(double));
  rc = rc ? rc : dalloc((void **)&c->dbuf2, nfac * sizeof(double));
  rc = rc ? rc : dalloc((void **)&c->rbuf, c->field * sizeof(double));
  rc = rc ? rc : dalloc((void **)&c->vpart, (size_t)c->ny * c->nz32 * 16 * sizeof(double));  // <= 16 waves/block
  rc = rc ? rc : dalloc((void **)&c->scal, SCAL_N * sizeof(double));
  rc = rc ? rc : dalloc((void **)&c->dD0, bandb);
  rc = rc ? rc : dalloc((void **)&c->dD1, bandb);
  rc = rc ? rc : dalloc((void **)&c->dD2, bandb);
  rc = rc ? rc : dalloc((void **)&c->dD0LU, bandb);
  rc = rc ? rc : dalloc((void **)&c->dinteg, c->ny * sizeof(double));
  rc = rc ? rc : dalloc((void **)&c->ddyr, c->ny * sizeof(double));
  rc = rc ? rc : dalloc((void **)&c->drows, std::max<size_t>(1, c->opp_rows.size()) * sizeof(double));
  rc = rc ? rc : dalloc((void **)&c->twz, c->nz32 * sizeof(double2));
  rc = rc ? rc : dalloc((void **)&c->twx, c->nx32 * sizeof(double2));
  rc = rc ? rc : dalloc((void **)&c->slotmap, 16 * sizeof(int));
  rc = rc ? rc : dalloc((void **)&c->ppe_bc, 2 * c->plane * sizeof(double2));
  rc = rc ? rc : dalloc((void **)&c->dvv, c->ny * sizeof(double));
  rc = rc ? rc : dalloc((void **)&c->dcenter, c->ny * sizeof(double));
  rc = rc ? rc : dalloc((void **)&c->dstat, ((size_t)16 * c->ny + 2) * sizeof(double));
  if (reserve) (void)hipFree(reserve);
  // mean line: its columns (and, when they fit, its band) in LDS
  {
    const size_t cols = (size_t)pb::MEAN_COLS * c->ny * sizeof(double2), db = bandb;
    if (cols + db <= (size_t)lds_max) {
      c->mean_lds = cols + db;
    } else if (cols <= (size_t)lds_max) {
      c->mean_lds = cols;
      rc = rc ? rc : dalloc((void **)&c->mean_db, 3 * bandb);
    }
    if (!rc && c->mean_lds > 65536 &&
        (hipFuncSetAttribute(reinterpret_cast<const void *>(&pb::mean_kernel<3>),
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)c->mean_lds) != hipSuccess ||
         hipFuncSetAttribute(reinterpret_cast<const void *>(&pb::mean_kernel<4>),
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)c->mean_lds) != hipSuccess ||
         hipFuncSetAttribute(reinterpret_cast<const void *>(&pb::mean_kernel<5>),
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)c->mean_lds) != hipSuccess))
      c->mean_lds = 0;
  }
  if (rc) { ctx_free(c); return rc; }
  if (hipHostMalloc((void **)&c->hscal, SCAL_N * sizeof(double), hipHostMallocDefault) != hipSuccess) {
    ctx_free(c);
    return fail(PB_E_HIP, "hipHostMalloc failed");
  }
  std::vector<double2> tz = twiddles(c->nz32), tx = twiddles(c->nx32);
#define UP(dst, src, n) \
  if (hipMemcpy(dst, src, n, hipMemcpyHostToDevice) != hipSuccess) { ctx_free(c); return fail(PB_E_HIP, "upload " #dst); }
  UP(c->dD0, c->d0.data(), bandb);
  UP(c->dD1, c->d1.data(), bandb);
  UP(c->dD2, c->d2.data(), bandb);
  UP(c->dD0LU, c->d0lu.data(), bandb);
  UP(c->dinteg, c->col.integ.data(), c->ny * sizeof(double));
  UP(c->ddyr, c->dyr.data(), c->ny * sizeof(double));
  if (!c->opp_rows.empty()) UP(c->drows, c->opp_rows.data(), c->opp_rows.size() * sizeof(double));
  {
    std::vector<double> cen(c->ny, 0.0), vals;
    int first = 0;
    if (pb::basis_row(c->col, 0.0, first, vals)) { ctx_free(c); return fail(PB_E_ARG, "no basis row at y = 0"); }
    for (size_t j = 0; j < vals.size(); ++j) cen[first + j] = vals[j];
    UP(c->dcenter, cen.data(), c->ny * sizeof(double));
  }
  UP(c->twz, tz.data(), c->nz32 * sizeof(double2));
  UP(c->twx, tx.data(), c->nx32 * sizeof(double2));
#undef UP
  if (hipMemset(c->scal, 0, SCAL_N * sizeof(double)) != hipSuccess) { ctx_free(c); return fail(PB_E_HIP, "memset"); }
  *out = c;
  return PB_OK;
}

int pb_get_info(const pb_ctx *c, int32_t *bw, int64_t *bytes) {
  if (!c) return fail(PB_E_ARG, "null context");
  if (bw) *bw = c->bw;
  if (bytes) *bytes = (int64_t)c->bytes;
  return PB_OK;
}

int pb_get_labs(const pb_ctx *c, double *y, double *d0, double *d1, double *d2, double *d0lu, double *integ) {
  if (!c) return fail(PB_E_ARG, "null context");
  const size_t nb = (size_t)c->ny * c->W;
  if (y) std::copy(c->col.y.begin(), c->col.y.end(), y);
  if (d0) std::copy(c->d0.begin(), c->d0.begin() + nb, d0);
  if (d1) std::copy(c->d1.begin(), c->d1.begin() + nb, d1);
  if (d2) std::copy(c->d2.begin(), c->d2.begin() + nb, d2);
  if (d0lu) std::copy(c->d0lu.begin(), c->d0lu.begin() + nb, d0lu);
  if (integ) std::copy(c->col.integ.begin(), c->col.integ.end(), integ);
  return PB_OK;
}

// copy nslot reference-layout slots from host into device slots map[]
static int upload_ref(pb_ctx *c, const double *host, int nslot, const int *map, double2 *dev_base) {
  const size_t bytes = (size_t)nslot * c->field * sizeof(double2);
  if (bytes > kPadFields * c->pfield * sizeof(double2)) return fail(PB_E_ARG, "staging too small");
  HIPCHK(hipMemcpyAsync(c->slotmap, map, nslot * sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->pad, host, bytes, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(pb::ref_to_dev_kernel, dim3(c->nz, nslot), dim3(256), 0, c->stream, c->pad, dev_base, c->ny,
                     c->nxh, c->nz, c->slotmap);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return PB_OK;
}

static int download_ref(pb_ctx *c, double *host, int nslot, const int *map, const double2 *dev_base) {
  const size_t bytes = (size_t)nslot * c->field * sizeof(double2);
  HIPCHK(hipMemcpyAsync(c->slotmap, map, nslot * sizeof(int), hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(pb::dev_to_ref_kernel, dim3(c->nz, nslot), dim3(256), 0, c->stream, dev_base, c->pad, c->ny,
                     c->nxh, c->nz, c->slotmap);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(host, c->pad, bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return PB_OK;
}

static const int kStateMap[4] = {0, 1, 2, 6};
static const int kRKMap[3] = {0, 1, 2};

int pb_set_state(pb_ctx *c, const double *U, const double *rk) {
  if (!c || !U) return fail(PB_E_ARG, "null argument");
  int r = upload_ref(c, U, 4, kStateMap, c->U);
  if (r) return r;
  if (rk) {
    r = upload_ref(c, rk, 3, kRKMap, c->RK);
    if (r) return r;
  }
  c->have_state = true;
  return PB_OK;
}

int pb_get_state(pb_ctx *c, double *U, double *rk) {
  if (!c) return fail(PB_E_ARG, "null context");
  if (!c->have_state) return fail(PB_E_STATE, "no state set");
  if (U) { int r = download_ref(c, U, 4, kStateMap, c->U); if (r) return r; }
  if (rk) { int r = download_ref(c, rk, 3, kRKMap, c->RK); if (r) return r; }
  return PB_OK;
}

int pb_set_state_slots(pb_ctx *c, const double *u, const double *v, const double *w, const double *theta,
                       const double *rk) {
  if (!c || !u || !v || !w || !theta) return fail(PB_E_ARG, "null argument");
  const double *f[4] = {u, v, w, theta};
  for (int i = 0; i < 4; ++i) {
    int r = upload_ref(c, f[i], 1, &kStateMap[i], c->U);
    if (r) return r;
  }
  if (rk) {
    int r = upload_ref(c, rk, 3, kRKMap, c->RK);
    if (r) return r;
  }
  c->have_state = true;
  return PB_OK;
}

int pb_get_state_slots(pb_ctx *c, double *u, double *v, double *w, double *theta, double *rk) {
  if (!c) return fail(PB_E_ARG, "null context");
  if (!c->have_state) return fail(PB_E_STATE, "no state set");
  double *f[4] = {u, v, w, theta};
  for (int i = 0; i < 4; ++i)
    if (f[i]) {
      int r = download_ref(c, f[i], 1, &kStateMap[i], c->U);
      if (r) return r;
    }
  if (rk) {
    int r = download_ref(c, rk, 3, kRKMap, c->RK);
    if (r) return r;
  }
  return PB_OK;
}

static unsigned blocks_for(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

// Record an event pair around a launch when timing is on.
static hipEvent_t next_event(pb_ctx *c) {
  if (c->ev_used == c->ev_pool.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    c->ev_pool.push_back(e);
  }
  return c->ev_pool[c->ev_used++];
}
struct StageTimer {
  pb_ctx *c;
  int stage;
  size_t idx = 0;
  bool on = false;
  StageTimer(pb_ctx *cc, int st) : c(cc), stage(st) {
    if (!c->timing) return;
    idx = c->ev_used;
    hipEvent_t a = next_event(c), b = next_event(c);
    if (!a || !b) return;
    on = (hipEventRecord(a, c->stream) == hipSuccess);
  }
  ~StageTimer() {
    if (!on) return;
    if (hipEventRecord(c->ev_pool[idx + 1], c->stream) == hipSuccess) c->ev_marks.push_back({stage, idx});
  }
};

int pb_init_stores(pb_ctx *c) {
  if (!c || !c->have_state) return fail(PB_E_STATE, "pb_init_stores needs a state");
  pb::LineParams P = c->line_params(1, 0.0);
  hipLaunchKernelGGL(pb::init_stores_kernel, dim3(blocks_for(c->nsys, 64)), dim3(64), 0, c->stream, P,
                     (pb::cplx *)c->U, (pb::cplx *)c->RK, (pb::cplx *)c->buf, c->plane, c->nsys, c->nxh, c->nh,
                     c->lmfxr, c->lmfzr);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return PB_OK;
}

// ---- restart paths and the checkpoint (initial_step.f90:20-45, 237-513;
// data_store.f90:92-118) --------------------------------------------------
static const int kSlot0[1] = {0}, kSlot1[1] = {1}, kSlot2[1] = {2}, kSlot6[1] = {6};

// the three mean lines (ny doubles each) into the (0,0) line of RK 0, 1, 2,
// u_mean shifted by -u_gal (data_read_vg / data_read_gphi)
static int upload_means(pb_ctx *c, const double *u_mean, const double *w_mean, const double *t_mean) {
  std::vector<double2> col(c->ny);
  const double *src[3] = {u_mean, w_mean, t_mean};
  for (int f = 0; f < 3; ++f) {
    for (int i = 0; i < c->ny; ++i) col[i] = make_double2(src[f][i] - (f == 0 ? c->cfg.u_gal : 0.0), 0.0);
    HIPCHK(hipMemcpy2D(c->RK + (size_t)f * c->field, c->plane * sizeof(double2), col.data(), sizeof(double2),
                       sizeof(double2), c->ny, hipMemcpyHostToDevice));
  }
  return PB_OK;
}

int pb_set_state_vg(pb_ctx *c, const double *v, const double *g, const double *theta, const double *u_mean,
                    const double *w_mean, const double *t_mean) {
  if (!c || !v || !g || !theta || !u_mean || !w_mean || !t_mean) return fail(PB_E_ARG, "null argument");
  int r;
  if ((r = upload_ref(c, v, 1, kSlot1, c->U))) return r;
  if ((r = upload_ref(c, g, 1, kSlot0, c->RK))) return r;
  if ((r = upload_ref(c, theta, 1, kSlot2, c->RK))) return r;
  if ((r = upload_means(c, u_mean, w_mean, t_mean))) return r;
  pb::LineParams P = c->line_params(1, 0.0);
  hipLaunchKernelGGL(pb::init_vg_kernel, dim3(blocks_for(c->nsys, 64)), dim3(64), 0, c->stream, P, (pb::cplx *)c->U,
                     (pb::cplx *)c->RK, (pb::cplx *)c->buf, c->plane, c->nsys, c->nxh, c->nh, c->lmfxr, c->lmfzr);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  c->have_state = true;
  return PB_OK;
}

int pb_set_state_gphi(pb_ctx *c, const double *g, const double *phi, const double *theta, const double *u_mean,
                      const double *w_mean, const double *t_mean) {
  if (!c || !g || !phi || !theta || !u_mean || !w_mean || !t_mean) return fail(PB_E_ARG, "null argument");
  int r;
  if ((r = upload_ref(c, g, 1, kSlot0, c->RK))) return r;
  if ((r = upload_ref(c, phi, 1, kSlot1, c->RK))) return r;
  if ((r = upload_ref(c, theta, 1, kSlot2, c->RK))) return r;
  if ((r = upload_means(c, u_mean, w_mean, t_mean))) return r;
  pb::LineParams P = c->line_params(1, 0.0);
  hipLaunchKernelGGL(pb::init_gphi_kernel, dim3(blocks_for(c->nsys, 64)), dim3(64), 0, c->stream, P,
                     (pb::cplx *)c->U, (const pb::cplx *)c->RK, (pb::cplx *)c->buf, c->dbuf, c->plane, c->nsys, c->nxh,
                     c->nh, c->lmfxr, c->lmfzr);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  c->have_state = true;
  return PB_OK;
}

int pb_get_checkpoint(pb_ctx *c, double *u, double *v, double *w, double *theta, double *g, double *phi,
                      double *u_mean, double *w_mean, double *t_mean) {
  if (!c) return fail(PB_E_ARG, "null context");
  if (!c->have_state) return fail(PB_E_STATE, "no state set");
  int r;
  double *fld[6] = {u, v, w, theta, g, phi};
  const int *map[6] = {kSlot0, kSlot1, kSlot2, kSlot6, kSlot0, kSlot1};
  for (int q = 0; q < 6; ++q)
    if (fld[q] && (r = download_ref(c, fld[q], 1, map[q], q < 4 ? c->U : c->RK))) return r;
  std::vector<double2> col(c->ny);
  double *dst[3] = {u_mean, w_mean, t_mean};
  const int slot[3] = {0, 2, 6};
  for (int f = 0; f < 3; ++f) {
    if (!dst[f]) continue;
    HIPCHK(hipMemcpy2D(col.data(), sizeof(double2), c->U + (size_t)slot[f] * c->field, c->plane * sizeof(double2),
                       sizeof(double2), c->ny, hipMemcpyDeviceToHost));
    for (int i = 0; i < c->ny; ++i) dst[f][i] = col[i].x + (f == 0 ? c->cfg.u_gal : 0.0);
  }
  return PB_OK;
}

// ---- stage launchers ------------------------------------------------------
static int launch_oppvd(pb_ctx *c, const double2 *state) {
  StageTimer tm(c, PB_STAGE_OPPVD);
  hipLaunchKernelGGL(pb::opp_kernel, dim3(blocks_for(c->nsys, 64)), dim3(64), 0, c->stream,
                     (const pb::cplx *)state, (pb::cplx *)c->vd, c->plane, c->nsys, c->ny, c->drows);
  HIPCHK(hipGetLastError());
  return PB_OK;
}

// Smallest elements-per-thread instantiation that holds n elements.
static int pick_ept(size_t n) {
  const int ept[] = {4, 8, 12, 16, 24, 32};
  for (int e : ept)
    if ((size_t)e * pb::FFT_THREADS >= n) return e;
  return -1;
}

extern "C++" {
template <class F>
static int launch_zinv_t(pb_ctx *c) {
  const size_t ldsz = (size_t)c->nz32 * c->Bz * sizeof(double2);
  StageTimer tm(c, PB_STAGE_ZINV);
  hipLaunchKernelGGL(pb::zinv_kernel<F>, dim3(c->nxh / c->Bz, c->ny, 4), dim3(F::T), ldsz, c->stream, c->U,
                     c->pad, c->pz, c->twz, c->nxh, c->nz, c->ny, c->Bz, c->field, c->pfield, pb::MAP_STATE, 1.0);
  HIPCHK(hipGetLastError());
  return PB_OK;
}
template <class F>
static int launch_xphys_t(pb_ctx *c, int want_vv) {
  const size_t ldsx = F::lds_elems(c->nx32, 2) * sizeof(double2);
  StageTimer tm(c, PB_STAGE_XPHYS);
  hipLaunchKernelGGL(pb::xphys_kernel<F>, dim3(c->nz32, c->ny), dim3(F::T), ldsx, c->stream, c->pad, c->px, c->twx,
                     c->nxh, c->nz32, c->pfield, c->ddyr, c->dxr, c->dzr, c->vpart, want_vv ? 3 : 2, c->mom_active);
  c->vpart_n = (size_t)c->ny * c->nz32 * ((F::T + 63) / 64);
  HIPCHK(hipGetLastError());
  return PB_OK;
}
template <class F>
static int launch_zfwd_t(pb_ctx *c, int want_vv, double scale) {
  const size_t ldsz = (size_t)c->nz32 * c->Bz * sizeof(double2);
  StageTimer tm(c, PB_STAGE_ZFWD);
  hipLaunchKernelGGL(pb::zfwd_kernel<F>, dim3(c->nxh / c->Bz, c->ny, want_vv ? 9 : 8), dim3(F::T), ldsz, c->stream,
                     c->pad, c->U, c->pz, c->twz, c->nxh, c->nz, c->ny, c->Bz, c->field, c->pfield, scale,
                     pb::MAP_PROD);
  HIPCHK(hipGetLastError());
  return PB_OK;
}
}  // extern "C++"

#define PB_EPT_DISPATCH(ept, call)                                      \
  switch (ept) {                                                        \
    case 4: { using F = pb::RtFft<4>; call; } break;                    \
    case 8: { using F = pb::RtFft<8>; call; } break;                    \
    case 12: { using F = pb::RtFft<12>; call; } break;                  \
    case 16: { using F = pb::RtFft<16>; call; } break;                  \
    case 24: { using F = pb::RtFft<24>; call; } break;                  \
    case 32: { using F = pb::RtFft<32>; call; } break;                  \
    default: return fail(PB_E_ARG, "transform too large for one block"); \
  }

// fixed-size transforms for the sizes they were built for (PB_FFT_GENERIC=1
// forces the run-time passes everywhere)
static bool ct_z(const pb_ctx *c) { return c->ct_z_ok; }
static bool ct_x(const pb_ctx *c) { return c->ct_x_ok; }

static int launch_zinv_stream(pb_ctx *c) {
  StageTimer tm(c, PB_STAGE_ZINV);
  hipLaunchKernelGGL(pb::zinv_stream_kernel, dim3(c->ncu), dim3(pb::zs::T), 0, c->stream, c->U, c->pad, c->twz, c->nxh,
                     c->nz, c->ny, c->field, c->pfield);
  HIPCHK(hipGetLastError());
  return PB_OK;
}
static int launch_zfwd_stream(pb_ctx *c, int want_vv, double scale) {
  StageTimer tm(c, PB_STAGE_ZFWD);
  hipLaunchKernelGGL(pb::zfwd_stream_kernel, dim3(c->ncu), dim3(pb::zs::T), 0, c->stream, c->pad, c->U, c->twz,
                     c->nxh, c->nz, c->ny, c->field, c->pfield, scale, want_vv ? 9 : 8);
  HIPCHK(hipGetLastError());
  return PB_OK;
}

#ifndef PB_XS_THREADS
#define PB_XS_THREADS 512
#endif
static int launch_xphys_stream(pb_ctx *c) {
  constexpr int T = PB_XS_THREADS;
  StageTimer tm(c, PB_STAGE_XPHYS);
  hipLaunchKernelGGL(pb::xphys_stream_kernel<T>, dim3(c->ncu), dim3(T), 0, c->stream, c->pad, c->twx, c->nz32, c->ny,
                     c->pfield, c->ddyr, c->dxr, c->dzr, c->vpart);
  c->vpart_n = (size_t)c->ncu;  // one partial maximum per block
  HIPCHK(hipGetLastError());
  return PB_OK;
}

// y_to_x of the four state fields (device slots 0, 1, 2, 6 -> pad fields 0-3)
static int launch_zinv_any(pb_ctx *c) {
  int r = PB_OK;
  const int ez = pick_ept((size_t)c->nz32 * c->Bz);
  if (c->zs_ok && c->zs_inv) r = launch_zinv_stream(c);
  else if (ct_z(c)) r = launch_zinv_t<pb::FftZ1152>(c);
  else PB_EPT_DISPATCH(ez, r = launch_zinv_t<F>(c));
  return r;
}
// dft_x_realspace: pad fields 0-3 -> products in pad fields 0-7 (8 = v*v),
// per-block CFL partials in vpart
static int launch_xphys_any(pb_ctx *c, int want_vv) {
  int r = PB_OK;
  const int ex = pick_ept((size_t)c->nx32 * 2);
  if (c->xs_ok && !want_vv) r = launch_xphys_stream(c);
  else if (ct_x(c)) r = launch_xphys_t<pb::FftX1536>(c, want_vv);
  else PB_EPT_DISPATCH(ex, r = launch_xphys_t<F>(c, want_vv));
  return r;
}
// x_to_y (+ rescale_data when scale = 1/(nx32 nz32)): pad fields 0-7 (8)
// -> device slots 0, 1, 2, 3, 4, 6, 7, 8 (5)
static int launch_zfwd_any(pb_ctx *c, int want_vv, double scale) {
  int r = PB_OK;
  const int ez = pick_ept((size_t)c->nz32 * c->Bz);
  if (c->zs_ok && c->zs_fwd) r = launch_zfwd_stream(c, want_vv, scale);
  else if (ct_z(c)) r = launch_zfwd_t<pb::FftZ1152>(c, want_vv, scale);
  else PB_EPT_DISPATCH(ez, r = launch_zfwd_t<F>(c, want_vv, scale));
  return r;
}
// update: velmax_kernel's VM_STORE / VM_DT bits
static int launch_velmax(pb_ctx *c, int update) {
  StageTimer tm(c, PB_STAGE_VELMAX);
  hipLaunchKernelGGL(pb::velmax_kernel, dim3(1), dim3(pb::VELMAX_THREADS), 0, c->stream, c->vpart, c->vpart_n,
                     c->scal, update, c->cfg.dt_factor, c->cflx);
  HIPCHK(hipGetLastError());
  return PB_OK;
}

static int launch_nonlinear(pb_ctx *c, int update, int want_vv) {
  int r;
  if ((r = launch_zinv_any(c))) return r;
  if ((r = launch_xphys_any(c, want_vv))) return r;
  if ((r = launch_velmax(c, update))) return r;
  return launch_zfwd_any(c, want_vv, c->rescale);
}

extern "C++" {
template <int BW>
static int launch_sweeps(pb_ctx *c, const pb::LineParams &P, const pb::WsArgs &A, int s) {
  const dim3 block(64);
#define SWEEP(id, strm)                                                                                     \
  {                                                                                                         \
    const size_t nt = pb::sweep_threads(id, A);                                                             \
    hipLaunchKernelGGL((pb::sweep_kernel<BW, id>), dim3(blocks_for(nt, 64)), block, 0, strm, P, A, c->dbuf2, \
                       nt, c->scal);                                                                        \
    HIPCHK(hipGetLastError());                                                                              \
  }
  const bool s1 = s == 1;
  if (s1) SWEEP(pb::SW_A1, c->stream);
  // theta chain (slots 5-8, RK3, dbuf2) beside all other sweeps (slots 0-4,
  // the extra column, buffers 0-1, RK1, RK2, dbuf); joined at the end
  hipStream_t th = c->linear ? c->stream : c->side2;
  if (!c->linear) {
    HIPCHK(hipEventRecord(c->ev_fork2, c->stream));
    HIPCHK(hipStreamWaitEvent(c->side2, c->ev_fork2, 0));
  }
  SWEEP(pb::SW_AT, th);
  if (s1) SWEEP(pb::SW_BT1, th) else SWEEP(pb::SW_BT, th);
  SWEEP(pb::SW_CT, th);
  SWEEP(pb::SW_DT, th);
  if (!c->linear) HIPCHK(hipEventRecord(c->ev_join2, c->side2));
  SWEEP(pb::SW_A, c->stream);
  if (s1) SWEEP(pb::SW_BG1, c->stream) else SWEEP(pb::SW_BG, c->stream);
  SWEEP(pb::SW_BV, c->stream);
  SWEEP(pb::SW_CNU, c->stream);
  SWEEP(pb::SW_DNU, c->stream);
  SWEEP(pb::SW_E, c->stream);
  SWEEP(pb::SW_F, c->stream);
  SWEEP(pb::SW_G, c->stream);
  if (!c->linear) HIPCHK(hipStreamWaitEvent(c->stream, c->ev_join2, 0));
#undef SWEEP
  return PB_OK;
}
}  // extern "C++"

static pb::WsArgs ws_args(pb_ctx *c, int s) {
  pb::WsArgs A;
  A.U = (pb::cplx *)c->U;
  A.RK = (pb::cplx *)c->RK;
  A.buf = (pb::cplx *)c->buf;
  A.dbuf = c->dbuf;
  A.rbuf = c->rbuf;
  // the padded pencils are dead during the wall-normal stage: their first
  // field holds the sweeps' extra column
  A.xcol = (pb::cplx *)c->pad;
  A.vd = c->cfg.opp_alpha != 0.0 ? (const pb::cplx *)c->vd : nullptr;
  A.usrc = c->cfg.scalar_bc == PB_SCALAR_MIXED ? (const pb::cplx *)c->usrc : nullptr;
  A.src_scale = c->cfg.nu / c->cfg.pr / c->cfg.bulk_u;
  A.act = c->act_on ? (const pb::cplx *)c->act : nullptr;
  A.plane = c->plane;
  A.field = c->field;
  A.nxh = c->nxh;
  A.nz = c->nz;
  A.nh = c->nh;
  A.lmfxr = c->lmfxr;
  A.lmfzr = c->lmfzr;
  A.opp_alpha = c->cfg.opp_alpha;
  A.mean_out = c->scal + pb::SCAL_MEAN0 + 4 * (s - 1);
  return A;
}

static int launch_wavespace(pb_ctx *c, int s) {
  pb::LineParams P = c->line_params(s, 0.0);
  const pb::WsArgs A = ws_args(c, s);
  StageTimer tm(c, PB_STAGE_WAVESPACE);
  if (c->fused && (c->bw >= 3 && c->bw <= 5)) {
    // mean_kernel touches only the columns of system 0 and its own LDS: fork
    // it onto the side stream, run the sweeps of all other systems, join.
    // The one-thread wave00_kernel also uses the scratch columns, which the
    // sweeps use too, so it runs after them.
    if (c->mean_lds) {
      if (!c->linear) {
        HIPCHK(hipEventRecord(c->ev_fork, c->stream));
        HIPCHK(hipStreamWaitEvent(c->side, c->ev_fork, 0));
      }
      hipStream_t ms = c->linear ? c->stream : c->side;
      const dim3 g3(3), bt(pb::MEAN_THREADS);
      if (c->bw == 3) hipLaunchKernelGGL(pb::mean_kernel<3>, g3, bt, c->mean_lds, ms, P, A, c->scal, c->mean_db);
      if (c->bw == 4) hipLaunchKernelGGL(pb::mean_kernel<4>, g3, bt, c->mean_lds, ms, P, A, c->scal, c->mean_db);
      if (c->bw == 5) hipLaunchKernelGGL(pb::mean_kernel<5>, g3, bt, c->mean_lds, ms, P, A, c->scal, c->mean_db);
      HIPCHK(hipGetLastError());
      if (!c->linear) HIPCHK(hipEventRecord(c->ev_join, c->side));
    }
    int r = PB_OK;
    switch (c->bw) {
      case 3: r = launch_sweeps<3>(c, P, A, s); break;
      case 4: r = launch_sweeps<4>(c, P, A, s); break;
      case 5: r = launch_sweeps<5>(c, P, A, s); break;
    }
    if (c->mean_lds) {
      if (!c->linear) HIPCHK(hipStreamWaitEvent(c->stream, c->ev_join, 0));
    } else {
      hipLaunchKernelGGL(pb::wave00_kernel, dim3(1), dim3(64), 0, c->stream, P, A, c->scal);
      HIPCHK(hipGetLastError());
    }
    return r;
  }
  hipLaunchKernelGGL(pb::wavespace_kernel, dim3(blocks_for(c->nsys, 64)), dim3(64), 0, c->stream, P, A, c->nsys,
                     c->scal);
  HIPCHK(hipGetLastError());
  return PB_OK;
}

static int check_dt(double dt) {
  if (!(dt > 1e-14) || dt > 1e5) return fail(PB_E_NUMERIC, "dt out of range: %g", dt);
  return PB_OK;
}

// Statistics time, before the transforms (get_statistics_1: <vv> and the
// pressure wall values from the substep's input v).
static int launch_stats_pre(pb_ctx *c) {
  pb::LineParams P = c->line_params(1, 0.0);
  hipLaunchKernelGGL(pb::vv_kernel, dim3(blocks_for(c->ny, 64)), dim3(64), 0, c->stream, (const pb::cplx *)c->U,
                     c->plane, c->ny, c->nxh, c->dvv);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(pb::ppe_bc_kernel, dim3(blocks_for(c->nsys, 64)), dim3(64), 0, c->stream, P,
                     (const pb::cplx *)c->U, (pb::cplx *)c->buf, (pb::cplx *)c->ppe_bc, c->plane, c->nsys, c->cfg.nu);
  HIPCHK(hipGetLastError());
  return PB_OK;
}
// Statistics time, after the transforms: p into product slot 5 (get_pressure.f90).
static int launch_pressure(pb_ctx *c) {
  pb::LineParams P = c->line_params(1, 0.0);
  const pb::WsArgs A = ws_args(c, 1);
  hipLaunchKernelGGL(pb::ppe_kernel, dim3(blocks_for(c->nsys, 64)), dim3(64), 0, c->stream, P, A,
                     (const pb::cplx *)c->ppe_bc, c->dvv);
  HIPCHK(hipGetLastError());
  return PB_OK;
}

// transforms of substep s (with the statistics-time pressure when stats)
static int substep_front(pb_ctx *c, int s, bool stats) {
  int r;
  if (stats && (r = launch_stats_pre(c))) return r;
  if (c->cfg.opp_alpha != 0.0 && (r = launch_oppvd(c, c->U))) return r;
  // mixed-type scalar condition: the source needs u after the transforms
  // have overwritten slot 0 with P1
  if (c->usrc)
    HIPCHK(hipMemcpyAsync(c->usrc, c->U, c->field * sizeof(double2), hipMemcpyDeviceToDevice, c->stream));
  if ((r = launch_nonlinear(c, s == 1 ? (pb::VM_STORE | pb::VM_DT) : 0, stats))) return r;
  if (stats && (r = launch_pressure(c))) return r;
  return PB_OK;
}

static int substep_async(pb_ctx *c, int s) {
  int r;
  if ((r = substep_front(c, s, false))) return r;
  return launch_wavespace(c, s);
}

// dt into the device scalars, the non-finite flag cleared (through the
// pinned mirror; the previous read_outputs has synchronised, so the mirror is
// free)
static int upload_dt(pb_ctx *c, double dt) {
  static_assert(pb::SCAL_DT == 0 && pb::SCAL_NONFINITE == 1, "dt and the flag are uploaded together");
  c->hscal[pb::SCAL_DT] = dt;
  c->hscal[pb::SCAL_NONFINITE] = 0.0;
  HIPCHK(hipMemcpyAsync(c->scal, c->hscal, 2 * sizeof(double), hipMemcpyHostToDevice, c->stream));
  return PB_OK;
}

static int parse_outputs(pb_ctx *c, int s0, int s1, pb_substep_out *out, double *dt_inout);
static int read_outputs(pb_ctx *c, int s0, int s1, pb_substep_out *out, double *dt_inout) {
  HIPCHK(hipMemcpyAsync(c->hscal, c->scal, SCAL_N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return parse_outputs(c, s0, s1, out, dt_inout);
}
// outputs of substeps s0..s1 from the host mirror of the device scalars
static int parse_outputs(pb_ctx *c, int s0, int s1, pb_substep_out *out, double *dt_inout) {
  const double *h = c->hscal;
  *dt_inout = h[pb::SCAL_DT];
  if (out)
    for (int s = s0; s <= s1; ++s) {
      pb_substep_out &o = out[s - s0];
      const double *m = h + pb::SCAL_MEAN0 + 4 * (s - 1);
      o.dt = h[pb::SCAL_DT];
      o.velmax = h[pb::SCAL_VELMAX];
      o.phi_u = m[0];
      o.phi_t = m[1];
      o.tau_w = m[2];
      o.q_w = m[3];
    }
  // a non-finite velocity seen by any CFL reduction of these substeps, or
  // non-finite mean-line outputs: the state has blown up
  bool bad = h[pb::SCAL_NONFINITE] != 0.0 || !std::isfinite(h[pb::SCAL_VELMAX]);
  for (int s = s0; s <= s1; ++s)
    for (int q = 0; q < 4; ++q) bad = bad || !std::isfinite(h[pb::SCAL_MEAN0 + 4 * (s - 1) + q]);
  if (bad)
    return fail(PB_E_NONFINITE, "state not finite in substeps %d..%d (velmax = %g, phi_u = %g)", s0, s1,
                h[pb::SCAL_VELMAX], h[pb::SCAL_MEAN0 + 4 * (s1 - 1)]);
  return check_dt(*dt_inout);
}

int pb_substep(pb_ctx *c, int32_t s, double *dt_inout, pb_substep_out *out) {
  if (!c || !dt_inout) return fail(PB_E_ARG, "null argument");
  if (!c->have_state) return fail(PB_E_STATE, "pb_substep before pb_set_state");
  if (s < 1 || s > c->cfg.rk_step) return fail(PB_E_ARG, "substep %d out of range", s);
  int r = check_dt(*dt_inout);
  if (r && !(s == 1 && c->cfg.dt_factor != 1.0)) return r;
  if ((r = upload_dt(c, *dt_inout))) return r;
  if ((r = substep_async(c, s))) return r;
  return read_outputs(c, s, s, out, dt_inout);
}

static const int kSlot5[1] = {5};

int pb_substep_stats(pb_ctx *c, int32_t s, double *dt_inout, pb_substep_out *out, double *p_out, double *vv_out) {
  if (s != 1) return pb_substep(c, s, dt_inout, out);
  if (!c || !dt_inout) return fail(PB_E_ARG, "null argument");
  if (!c->have_state) return fail(PB_E_STATE, "pb_substep_stats before pb_set_state");
  int r = check_dt(*dt_inout);
  if (r && c->cfg.dt_factor == 1.0) return r;
  if ((r = upload_dt(c, *dt_inout))) return r;
  if ((r = substep_front(c, s, true))) return r;
  // p leaves through the staging buffer before the solver reuses slot 5
  if (p_out && (r = download_ref(c, p_out, 1, kSlot5, c->U))) return r;
  if (vv_out) {
    HIPCHK(hipMemcpyAsync(vv_out, c->dvv, c->ny * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  if ((r = launch_wavespace(c, s))) return r;
  return read_outputs(c, s, s, out, dt_inout);
}

static constexpr int kMaxWavesPerLine = 16;  // the x pass's blocks have at most 16 waves

int pb_substep_stats2(pb_ctx *c, int32_t s, double *dt_inout, pb_substep_out *out, double *p_out, double *vv_out,
                      double *stat2) {
  if (!stat2) return pb_substep_stats(c, s, dt_inout, out, p_out, vv_out);
  if (!c || !dt_inout) return fail(PB_E_ARG, "null argument");
  if (s != 1) return fail(PB_E_ARG, "pb_substep_stats2: statistics are taken on substep 1, not %d", s);
  if (!c->have_state) return fail(PB_E_STATE, "pb_substep_stats2 before pb_set_state");
  if (c->bw < 3 || c->bw > 5) return fail(PB_E_ARG, "pb_substep_stats2: bandwidth %d outside 3..5", c->bw);
  int r = check_dt(*dt_inout);
  if (r && c->cfg.dt_factor == 1.0) return r;
  if (!c->stat_uvw) {
    auto alloc = [&](void **p, size_t bytes) -> int {
      const hipError_t e = hipMalloc(p, bytes);
      if (e != hipSuccess) return fail(PB_E_HIP, "hipMalloc(%zu) for statistics: %s", bytes, hipGetErrorString(e));
      c->bytes += bytes;
      return PB_OK;
    };
    if ((r = alloc((void **)&c->stat_uvw, 3 * c->field * sizeof(double2))) ||
        (r = alloc((void **)&c->dmom, (size_t)pb::NMOM * c->ny * c->nz32 * kMaxWavesPerLine * sizeof(double))) ||
        (r = alloc((void **)&c->dmsum, (size_t)pb::NMOM * c->ny * sizeof(double))) ||
        (r = alloc((void **)&c->dpv, (size_t)8 * c->ny * sizeof(double))) ||
        (r = alloc((void **)&c->dstat2, (size_t)PB_NSTAT2_LINES * c->ny * sizeof(double))) ||
        (r = alloc((void **)&c->dstat2_tmp, (size_t)2 * pb::STAT2_DDY * c->ny * sizeof(double2))))
      return r;
  }
  if ((r = upload_dt(c, *dt_inout))) return r;
  // the substep's input u, v, w, kept for the correlations with p
  HIPCHK(hipMemcpyAsync(c->stat_uvw, c->U, 3 * c->field * sizeof(double2), hipMemcpyDeviceToDevice, c->stream));
  c->mom_active = c->dmom;
  r = substep_front(c, s, true);
  c->mom_active = nullptr;
  if (r) return r;
  if (c->vpart_n > (size_t)c->ny * c->nz32 * kMaxWavesPerLine) return fail(PB_E_ARG, "x pass: too many waves");
  const pb::LineParams P = c->line_params(1, 0.0);
  const dim3 b64(64), grid(blocks_for(c->plane, 64));
  pb::cplx *buf = (pb::cplx *)c->buf;
  switch (c->bw) {
    case 3: hipLaunchKernelGGL(pb::stats2_deriv_kernel<3>, grid, b64, 0, c->stream, P, (const pb::cplx *)c->stat_uvw, buf, c->field, c->plane); break;
    case 4: hipLaunchKernelGGL(pb::stats2_deriv_kernel<4>, grid, b64, 0, c->stream, P, (const pb::cplx *)c->stat_uvw, buf, c->field, c->plane); break;
    case 5: hipLaunchKernelGGL(pb::stats2_deriv_kernel<5>, grid, b64, 0, c->stream, P, (const pb::cplx *)c->stat_uvw, buf, c->field, c->plane); break;
  }
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(pb::stats2_mom_kernel, dim3(c->ny, pb::NMOM), dim3(256), 0, c->stream, (const double *)c->dmom,
                     c->vpart_n / c->ny, c->dmsum);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(pb::stats2_pv_kernel, dim3(c->ny), dim3(pb::PV_THREADS), 0, c->stream,
                     (const pb::cplx *)c->stat_uvw, (const pb::cplx *)(c->U + 5 * c->field), (const pb::cplx *)buf,
                     c->plane, c->ny, c->nxh, c->nh, c->lmfxr, c->lmfzr, c->dpv);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(pb::stats2_final_kernel, dim3(1), dim3(64), 0, c->stream, P, (const double *)c->dmsum,
                     (const double *)c->dpv, c->rescale, (pb::cplx *)c->dstat2_tmp, c->dstat2);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(stat2, c->dstat2, (size_t)PB_NSTAT2_LINES * c->ny * sizeof(double), hipMemcpyDeviceToHost,
                        c->stream));
  if (p_out && (r = download_ref(c, p_out, 1, kSlot5, c->U))) return r;
  if (vv_out) HIPCHK(hipMemcpyAsync(vv_out, c->dvv, c->ny * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if ((r = launch_wavespace(c, s))) return r;
  return read_outputs(c, s, s, out, dt_inout);
}

int pb_statistics(pb_ctx *c, double *out) {
  if (!c || !out) return fail(PB_E_ARG, "null argument");
  if (!c->have_state) return fail(PB_E_STATE, "pb_statistics before pb_set_state");
  pb::LineParams P = c->line_params(1, 0.0);
  hipLaunchKernelGGL(pb::stats_uv_kernel, dim3(blocks_for((size_t)10 * c->ny, 64)), dim3(64), 0, c->stream,
                     (const pb::cplx *)c->U, c->plane, c->ny, c->nxh, c->dstat);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(pb::stats_mean_kernel, dim3(1), dim3(64), 0, c->stream, P, (const pb::cplx *)c->U,
                     (pb::cplx *)c->buf, c->plane, c->dcenter, c->cfg.u_gal, c->dstat);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, c->dstat, ((size_t)16 * c->ny + 2) * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return PB_OK;
}

int pb_spectra(pb_ctx *c, double *ex, double *ez) {
  if (!c || (!ex && !ez)) return fail(PB_E_ARG, "null argument");
  if (!c->have_state) return fail(PB_E_STATE, "pb_spectra before pb_set_state");
  if (c->bw < 3 || c->bw > 5) return fail(PB_E_ARG, "pb_spectra: bandwidth %d outside 3..5", c->bw);
  const pb::LineParams P = c->line_params(1, 0.0);
  pb::cplx *U = (pb::cplx *)c->U, *buf = (pb::cplx *)c->buf;
  const dim3 b64(64), grid(blocks_for(c->plane, 64));
  switch (c->bw) {
    case 3: hipLaunchKernelGGL(pb::deriv_kernel<3>, grid, b64, 0, c->stream, P, U, buf, c->field, c->plane); break;
    case 4: hipLaunchKernelGGL(pb::deriv_kernel<4>, grid, b64, 0, c->stream, P, U, buf, c->field, c->plane); break;
    case 5: hipLaunchKernelGGL(pb::deriv_kernel<5>, grid, b64, 0, c->stream, P, U, buf, c->field, c->plane); break;
  }
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(pb::vort_kernel, dim3(blocks_for(c->field, 256)), dim3(256), 0, c->stream, U,
                     (const pb::cplx *)buf, c->field, c->plane, c->nxh, c->nh, c->lmfxr, c->lmfzr, c->field);
  HIPCHK(hipGetLastError());
  // outputs staged in the padded-field buffer (free between substeps)
  double *dex = reinterpret_cast<double *>(c->pad), *dez = dex + (size_t)PB_NSPECTRA * c->nxh * c->ny;
  const size_t nx_out = (size_t)PB_NSPECTRA * c->nxh * c->ny, nz_out = (size_t)PB_NSPECTRA * c->nz * c->ny;
  if (ex) {
    hipLaunchKernelGGL(pb::spectra_x_kernel, dim3(blocks_for(nx_out, 256)), dim3(256), 0, c->stream,
                       (const pb::cplx *)U, c->field, c->plane, c->ny, c->nxh, c->nz, dex);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(ex, dex, nx_out * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  }
  if (ez) {
    hipLaunchKernelGGL(pb::spectra_z_kernel, dim3(blocks_for(nz_out, 256)), dim3(256), 0, c->stream,
                       (const pb::cplx *)U, c->field, c->plane, c->ny, c->nxh, c->nz, dez);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(ez, dez, nz_out * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return PB_OK;
}

int pb_step(pb_ctx *c, double *dt_inout, pb_substep_out *out) {
  if (!c || !dt_inout) return fail(PB_E_ARG, "null argument");
  if (!c->have_state) return fail(PB_E_STATE, "pb_step before pb_set_state");
  int r = check_dt(*dt_inout);
  if (r && c->cfg.dt_factor == 1.0) return r;
  if ((r = upload_dt(c, *dt_inout))) return r;
  for (int s = 1; s <= c->cfg.rk_step; ++s)
    if ((r = substep_async(c, s))) return r;
  return read_outputs(c, 1, c->cfg.rk_step, out, dt_inout);
}

// ---- captured step ----------------------------------------------------------
// pb_step's device work -- dt into the device scalars, the rk_step substeps,
// the scalars back to the pinned host mirror -- enqueued on the context's
// stream (the wall-normal stage forks onto the two side streams and joins
// them).  The kernels read dt from device memory, so the same work serves
// every dt.  No memset node: the device scalars are overwritten by copies.
static int enqueue_step(pb_ctx *c) {
  HIPCHK(hipMemcpyAsync(c->scal, c->hscal, 2 * sizeof(double), hipMemcpyHostToDevice, c->stream));
  for (int s = 1; s <= c->cfg.rk_step; ++s) {
    const int r = substep_async(c, s);
    if (r) return r;
  }
  HIPCHK(hipMemcpyAsync(c->hscal, c->scal, SCAL_N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  return PB_OK;
}

static void drop_graph(pb_ctx *c) {
  if (c->step_exec) (void)hipGraphExecDestroy(c->step_exec);
  if (c->step_graph) (void)hipGraphDestroy(c->step_graph);
  c->step_exec = nullptr;
  c->step_graph = nullptr;
}

// Capture `work` on stream `st` into *g and instantiate it into *x; the
// capture is always ended, also when `work` fails.
extern "C++" template <class W>
static int capture(hipStream_t st, W &&work, hipGraph_t *g, hipGraphExec_t *x) {
  HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  const int r = work();
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(st, &graph);
  if (r) {
    if (graph) (void)hipGraphDestroy(graph);
    return r;
  }
  if (e != hipSuccess || !graph) return fail(PB_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
  const hipError_t ei = hipGraphInstantiate(x, graph, nullptr, nullptr, 0);
  if (ei != hipSuccess) {
    (void)hipGraphDestroy(graph);
    return fail(PB_E_HIP, "hipGraphInstantiate: %s", hipGetErrorString(ei));
  }
  *g = graph;
  return PB_OK;
}

static int step_checks(pb_ctx *c, const double *dt_inout) {
  if (!c || !dt_inout) return fail(PB_E_ARG, "null argument");
  if (!c->have_state) return fail(PB_E_STATE, "step before pb_set_state");
  if (c->timing) return fail(PB_E_STATE, "a captured step cannot record the stage timing events; pb_timing(ctx, 0)");
  const int r = check_dt(*dt_inout);
  if (r && c->cfg.dt_factor == 1.0) return r;
  return PB_OK;
}

int pb_step_graph(pb_ctx *c, double *dt_inout, pb_substep_out *out) {
  int r;
  if ((r = step_checks(c, dt_inout))) return r;
  if (!c->step_exec && (r = capture(c->stream, [&] { return enqueue_step(c); }, &c->step_graph, &c->step_exec)))
    return r;
  c->hscal[pb::SCAL_DT] = *dt_inout;
  c->hscal[pb::SCAL_NONFINITE] = 0.0;
  HIPCHK(hipGraphLaunch(c->step_exec, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return parse_outputs(c, 1, c->cfg.rk_step, out, dt_inout);
}

// ---- ensemble of independent channels ---------------------------------------
struct pb_ensemble {
  std::vector<pb_ctx *> ch;
  std::vector<int> order;  // capture / launch order of the channels (tuner)
  // 0: one graph holding every channel's step; 1: each channel's own step
  // captured without side-stream forks and launched on its own stream; 2: each
  // channel's step enqueued directly (no graph) on its own streams
  int mode = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev_fork = nullptr;
  std::vector<hipEvent_t> ev_join;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

static void ens_drop_graph(pb_ensemble *e) {
  if (e->exec) (void)hipGraphExecDestroy(e->exec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  e->exec = nullptr;
  e->graph = nullptr;
}

int pb_ensemble_destroy(pb_ensemble *e) {
  if (!e) return PB_OK;
  ens_drop_graph(e);
  for (pb_ctx *c : e->ch) pb_destroy(c);
  for (hipEvent_t v : e->ev_join)
    if (v) (void)hipEventDestroy(v);
  if (e->ev_fork) (void)hipEventDestroy(e->ev_fork);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
  return PB_OK;
}

int pb_ensemble_create(int32_t n, const pb_config *cfgs, pb_ensemble **out) {
  if (!cfgs || !out || n < 1) return fail(PB_E_ARG, "pb_ensemble_create: %d channels", n);
  *out = nullptr;
  pb_ensemble *e = new pb_ensemble;
  e->ch.assign(n, nullptr);
  e->ev_join.assign(n, nullptr);
  for (int i = 0; i < n; ++i) {
    e->order.push_back(i);
    const int r = pb_create(&cfgs[i], &e->ch[i]);
    if (r) {
      pb_ensemble_destroy(e);
      return r;
    }
  }
  if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&e->ev_fork, hipEventDisableTiming) != hipSuccess) {
    pb_ensemble_destroy(e);
    return fail(PB_E_HIP, "ensemble stream / event");
  }
  for (int i = 0; i < n; ++i)
    if (hipEventCreateWithFlags(&e->ev_join[i], hipEventDisableTiming) != hipSuccess) {
      pb_ensemble_destroy(e);
      return fail(PB_E_HIP, "ensemble event");
    }
  *out = e;
  return PB_OK;
}

int pb_ensemble_channel(pb_ensemble *e, int32_t i, pb_ctx **ctx) {
  if (!e || !ctx || i < 0 || i >= (int)e->ch.size()) return fail(PB_E_ARG, "pb_ensemble_channel: index %d", i);
  *ctx = e->ch[i];
  return PB_OK;
}

// every channel's step forked from the ensemble stream onto the channel's
// own streams and joined back: one graph, one launch per ensemble step
static int ens_enqueue(pb_ensemble *e) {
  HIPCHK(hipEventRecord(e->ev_fork, e->stream));
  for (int i : e->order) {
    pb_ctx *c = e->ch[i];
    HIPCHK(hipStreamWaitEvent(c->stream, e->ev_fork, 0));
    const int r = enqueue_step(c);
    if (r) return r;
    HIPCHK(hipEventRecord(e->ev_join[i], c->stream));
    HIPCHK(hipStreamWaitEvent(e->stream, e->ev_join[i], 0));
  }
  return PB_OK;
}

// capture what the mode needs (once)
static int ens_prepare(pb_ensemble *e) {
  int r;
  if (e->mode == 0) {
    if (!e->exec && (r = capture(e->stream, [&] { return ens_enqueue(e); }, &e->graph, &e->exec))) return r;
  } else if (e->mode == 1) {
    for (pb_ctx *c : e->ch)
      if (!c->step_exec) {
        c->linear = true;
        r = capture(c->stream, [&] { return enqueue_step(c); }, &c->step_graph, &c->step_exec);
        c->linear = false;
        if (r) return r;
      }
  }
  return PB_OK;
}
// launch one ensemble step and wait for it
static int ens_launch(pb_ensemble *e) {
  if (e->mode == 0) {
    HIPCHK(hipGraphLaunch(e->exec, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
  } else if (e->mode == 1) {
    for (int i : e->order) HIPCHK(hipGraphLaunch(e->ch[i]->step_exec, e->ch[i]->stream));
    for (int i : e->order) HIPCHK(hipStreamSynchronize(e->ch[i]->stream));
  } else {
    for (int i : e->order) {
      const int r = enqueue_step(e->ch[i]);
      if (r) return r;
    }
    for (int i : e->order) HIPCHK(hipStreamSynchronize(e->ch[i]->stream));
  }
  return PB_OK;
}

int pb_ensemble_step(pb_ensemble *e, double *dt, pb_substep_out *out) {
  if (!e || !dt) return fail(PB_E_ARG, "null argument");
  int r;
  for (size_t i = 0; i < e->ch.size(); ++i)
    if ((r = step_checks(e->ch[i], &dt[i]))) return r;
  if ((r = ens_prepare(e))) return r;
  for (size_t i = 0; i < e->ch.size(); ++i) {
    e->ch[i]->hscal[pb::SCAL_DT] = dt[i];
    e->ch[i]->hscal[pb::SCAL_NONFINITE] = 0.0;
  }
  if ((r = ens_launch(e))) return r;
  int first = PB_OK;
  size_t off = 0;
  for (size_t i = 0; i < e->ch.size(); ++i) {
    pb_ctx *c = e->ch[i];
    r = parse_outputs(c, 1, c->cfg.rk_step, out ? out + off : nullptr, &dt[i]);
    off += c->cfg.rk_step;
    if (r && !first) first = r;
  }
  return first;
}

// The tuner: per channel, the z-transform tile width (columns per block of
// the per-block z kernels: fewer, longer rows per block or more blocks); for
// the ensemble, the capture order of the channels.  Candidates
// are timed as captured steps with HIP events, `reps` launches each, on the
// fields the channels hold (the zero fields of a fresh ensemble: call it
// before pb_set_state and pb_init_stores, which it leaves to the caller).
// None of the knobs changes any arithmetic.  ms_out[i] (may be NULL)
// receives the tuned step time of channel i, ms_out[n] that of the ensemble.
static int time_graph(hipGraphExec_t x, hipStream_t st, int reps, float *ms) {
  hipEvent_t a = nullptr, b = nullptr;
  if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
    if (a) (void)hipEventDestroy(a);
    return fail(PB_E_HIP, "tuner events");
  }
  hipError_t err = hipGraphLaunch(x, st);  // warm-up
  if (err == hipSuccess) err = hipEventRecord(a, st);
  for (int k = 0; k < reps && err == hipSuccess; ++k) err = hipGraphLaunch(x, st);
  if (err == hipSuccess) err = hipEventRecord(b, st);
  if (err == hipSuccess) err = hipEventSynchronize(b);
  if (err == hipSuccess) err = hipEventElapsedTime(ms, a, b);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (err != hipSuccess) return fail(PB_E_HIP, "tuner timing: %s", hipGetErrorString(err));
  *ms /= (float)reps;
  return PB_OK;
}

int pb_ensemble_tune(pb_ensemble *e, int32_t reps, double *ms_out) {
  if (!e || reps < 1) return fail(PB_E_ARG, "pb_ensemble_tune: reps %d", reps);
  const int n = (int)e->ch.size();
  int r;
  for (int i = 0; i < n; ++i) {
    pb_ctx *c = e->ch[i];
    if (c->timing) return fail(PB_E_STATE, "pb_ensemble_tune: stage timing is on");
    HIPCHK(hipMemsetAsync(c->U, 0, 9 * c->field * sizeof(double2), c->stream));
    HIPCHK(hipMemsetAsync(c->RK, 0, 3 * c->field * sizeof(double2), c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->have_state = true;
    c->hscal[pb::SCAL_DT] = 1e-3;
    c->hscal[pb::SCAL_NONFINITE] = 0.0;
    // candidate tile widths: the default and the halvings down to 1 (the
    // streamed z transforms of 3nz/2 = 1152 have their own fixed tiles)
    std::vector<int> bzs;
    for (int b = c->Bz; b >= 1 && bzs.size() < 3; b /= 2) bzs.push_back(b);
    if (c->ct_z_ok) bzs.assign(1, c->Bz);
    float best = 0.f;
    int best_bz = c->Bz;
    for (int bz : bzs) {
      drop_graph(c);
      c->Bz = bz;
      if ((r = capture(c->stream, [&] { return enqueue_step(c); }, &c->step_graph, &c->step_exec))) return r;
      float ms = 0.f;
      if ((r = time_graph(c->step_exec, c->stream, reps, &ms))) return r;
      if (best == 0.f || ms < best) { best = ms; best_bz = bz; }
    }
    drop_graph(c);
    c->Bz = best_bz;
    if (ms_out) ms_out[i] = best;
  }
  // capture order: as given, largest channel first, smallest first
  std::vector<int> by_size(n);
  for (int i = 0; i < n; ++i) by_size[i] = i;
  std::stable_sort(by_size.begin(), by_size.end(),
                   [&](int a, int b) { return e->ch[a]->field > e->ch[b]->field; });
  std::vector<std::vector<int>> orders = {e->order, by_size, std::vector<int>(by_size.rbegin(), by_size.rend())};
  float best = 0.f;
  std::vector<int> best_order = e->order;
  int best_mode = 0;
  for (int mode = 0; mode < 3; ++mode)
    for (const auto &o : orders) {
      ens_drop_graph(e);
      for (pb_ctx *c : e->ch) drop_graph(c);
      e->mode = mode;
      e->order = o;
      if ((r = ens_prepare(e))) return r;
      if ((r = ens_launch(e))) return r;  // warm-up
      const auto t0 = std::chrono::steady_clock::now();
      for (int k = 0; k < reps; ++k)
        if ((r = ens_launch(e))) return r;
      const float ms =
          std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count() / (float)reps;
      if (best == 0.f || ms < best) { best = ms; best_order = o; best_mode = mode; }
    }
  ens_drop_graph(e);
  for (pb_ctx *c : e->ch) drop_graph(c);
  e->order = best_order;
  e->mode = best_mode;
  for (pb_ctx *c : e->ch) c->have_state = false;
  if (ms_out) ms_out[n] = best;
  return PB_OK;
}

int pb_sync(pb_ctx *c) {
  if (!c) return fail(PB_E_ARG, "null context");
  HIPCHK(hipStreamSynchronize(c->stream));
  return PB_OK;
}

int pb_timing(pb_ctx *c, int32_t enable) {
  if (!c) return fail(PB_E_ARG, "null context");
  HIPCHK(hipStreamSynchronize(c->stream));
  c->timing = enable != 0;
  c->ev_used = 0;
  c->ev_marks.clear();
  return PB_OK;
}

int pb_timing_read(pb_ctx *c, double *ms, int64_t *launches) {
  if (!c || !ms || !launches) return fail(PB_E_ARG, "null argument");
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int i = 0; i < PB_NSTAGE; ++i) { ms[i] = 0.0; launches[i] = 0; }
  for (const auto &m : c->ev_marks) {
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, c->ev_pool[m.second], c->ev_pool[m.second + 1]));
    ms[m.first] += t;
    launches[m.first] += 1;
  }
  return PB_OK;
}

// ---- host-array drop-ins ---------------------------------------------------
int pb_nonlinear_host(pb_ctx *c, const double *U, double *P, int32_t want_vv, double *velmax) {
  if (!c || !U || !P) return fail(PB_E_ARG, "null argument");
  int r = upload_ref(c, U, 4, kStateMap, c->U);
  if (r) return r;
  c->have_state = true;
  if ((r = launch_nonlinear(c, pb::VM_STORE, want_vv != 0))) return r;
  static const int map8[9] = {0, 1, 2, 3, 4, 5, 6, 7, 8};
  // reference slots 0-8 of P (slot 5, v*v, only meaningful with want_vv)
  r = download_ref(c, P, 9, map8, c->U);
  if (r) return r;
  if (velmax) {
    HIPCHK(hipMemcpy(c->hscal, c->scal, SCAL_N * sizeof(double), hipMemcpyDeviceToHost));
    *velmax = c->hscal[pb::SCAL_VELMAX];
  }
  return PB_OK;
}

int pb_wavespace_host(pb_ctx *c, int32_t s, double dt, double *U, double *rk, const double *v_old,
                      pb_substep_out *out) {
  if (!c || !U || !rk) return fail(PB_E_ARG, "null argument");
  if (s < 1 || s > c->cfg.rk_step) return fail(PB_E_ARG, "substep %d out of range", s);
  static const int map9[9] = {0, 1, 2, 3, 4, 5, 6, 7, 8};
  int r;
  if (c->cfg.opp_alpha != 0.0) {
    if (!v_old) return fail(PB_E_ARG, "opposition control needs v_old");
    // stage v_old in U slot 1 and evaluate the detection planes
    static const int map1[1] = {1};
    if ((r = upload_ref(c, v_old, 1, map1, c->U))) return r;
    if ((r = launch_oppvd(c, c->U))) return r;
  }
  if (c->usrc) {
    // mixed-type scalar condition: u of the substep's input state in slot 9
    static const int map0[1] = {0};
    if ((r = upload_ref(c, U + 9 * 2 * c->field, 1, map0, c->usrc))) return r;
  }
  if ((r = upload_ref(c, U, 9, map9, c->U))) return r;
  if ((r = upload_ref(c, rk, 3, kRKMap, c->RK))) return r;
  c->have_state = true;
  if ((r = upload_dt(c, dt))) return r;
  if ((r = launch_wavespace(c, s))) return r;
  if ((r = download_ref(c, U, 9, map9, c->U))) return r;
  if ((r = download_ref(c, rk, 3, kRKMap, c->RK))) return r;
  double d = dt;
  return read_outputs(c, s, s, out, &d);
}

int pb_y_to_x_host(pb_ctx *c, const double *U, double *Upad) {
  if (!c || !U || !Upad) return fail(PB_E_ARG, "null argument");
  int r;
  if ((r = upload_ref(c, U, 4, kStateMap, c->U))) return r;
  c->have_state = true;
  if ((r = launch_zinv_any(c))) return r;
  HIPCHK(hipMemcpyAsync(Upad, c->pad, 4 * c->pfield * sizeof(double2), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return PB_OK;
}

int pb_realspace_host(pb_ctx *c, const double *Upad_in, double *Upad_out, int32_t want_vv, double *velmax) {
  if (!c || !Upad_in || !Upad_out) return fail(PB_E_ARG, "null argument");
  const size_t pf = c->pfield;
  HIPCHK(hipMemcpyAsync(c->pad, Upad_in, 4 * pf * sizeof(double2), hipMemcpyHostToDevice, c->stream));
  int r;
  if ((r = launch_xphys_any(c, want_vv != 0))) return r;
  if ((r = launch_velmax(c, pb::VM_STORE))) return r;
  // pad fields P1..P5, P7..P9 (, v*v) -> reference slots 0-4, 6-8 (, 5)
  double2 *out = reinterpret_cast<double2 *>(Upad_out);
  static const int dst[9] = {0, 1, 2, 3, 4, 6, 7, 8, 5};
  for (int g = 0; g < (want_vv ? 9 : 8); ++g)
    HIPCHK(hipMemcpyAsync(out + dst[g] * pf, c->pad + g * pf, pf * sizeof(double2), hipMemcpyDeviceToHost,
                          c->stream));
  HIPCHK(hipMemcpyAsync(c->hscal, c->scal, SCAL_N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (!want_vv) std::memset(out + 5 * pf, 0, pf * sizeof(double2));
  // slot 10 of the reference keeps the substep's v (dft_x_realspace.f90:123)
  std::memcpy(out + 9 * pf, reinterpret_cast<const double2 *>(Upad_in) + pf, pf * sizeof(double2));
  if (velmax) *velmax = c->hscal[pb::SCAL_VELMAX];
  return PB_OK;
}

int pb_x_to_y_host(pb_ctx *c, int32_t nfield, const double *Upad, double *U, double scale) {
  if (!c || !Upad || !U) return fail(PB_E_ARG, "null argument");
  if (nfield < 1 || nfield > 9) return fail(PB_E_ARG, "pb_x_to_y_host: nfield %d outside 1..9", nfield);
  const size_t pf = c->pfield;
  HIPCHK(hipMemcpyAsync(c->pad, Upad, nfield * pf * sizeof(double2), hipMemcpyHostToDevice, c->stream));
  if (nfield < 8) HIPCHK(hipMemsetAsync(c->pad + nfield * pf, 0, (8 - nfield) * pf * sizeof(double2), c->stream));
  int r;
  if ((r = launch_zfwd_any(c, nfield == 9, scale))) return r;
  static const int slot_of_pad[9] = {0, 1, 2, 3, 4, 6, 7, 8, 5};
  if ((r = download_ref(c, U, nfield, slot_of_pad, c->U))) return r;
  c->have_state = false;  // the device slots now hold transformed fields, not a state
  return PB_OK;
}

int pb_pressure_host(pb_ctx *c, const double *U, const double *P, double *p_out, double *vv_out) {
  if (!c || !U || !P || !p_out) return fail(PB_E_ARG, "null argument");
  int r;
  if ((r = upload_ref(c, U, 4, kStateMap, c->U))) return r;
  if ((r = launch_stats_pre(c))) return r;
  static const int map9[9] = {0, 1, 2, 3, 4, 5, 6, 7, 8};
  if ((r = upload_ref(c, P, 9, map9, c->U))) return r;
  if ((r = launch_pressure(c))) return r;
  if ((r = download_ref(c, p_out, 1, kSlot5, c->U))) return r;
  if (vv_out) {
    HIPCHK(hipMemcpyAsync(vv_out, c->dvv, c->ny * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  c->have_state = false;
  return PB_OK;
}

// ---- LABS band operations on host arrays (device kernels, one thread per
// matrix / column; the null stream, synchronous)
static int labs_check(int32_t n, int32_t bw, int32_t count) {
  if (bw < 1 || n < 2 * bw + 1 || count < 1) return fail(PB_E_ARG, "labs: n %d, bw %d, count %d", n, bw, count);
  return PB_OK;
}
struct DevBuf {
  void *p = nullptr;
  ~DevBuf() { if (p) (void)hipFree(p); }
};

int pb_labs_lu(int32_t n, int32_t bw, int32_t nmat, double *A) {
  int r;
  if (!A) return fail(PB_E_ARG, "null argument");
  if ((r = labs_check(n, bw, nmat))) return r;
  const size_t bytes = (size_t)nmat * n * (2 * bw + 1) * sizeof(double);
  DevBuf d;
  HIPCHK(hipMalloc(&d.p, bytes));
  HIPCHK(hipMemcpy(d.p, A, bytes, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(pb::labs_lu_kernel, dim3(blocks_for(nmat, 64)), dim3(64), 0, 0, n, bw, nmat, (double *)d.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpy(A, d.p, bytes, hipMemcpyDeviceToHost));
  return PB_OK;
}

int pb_labs_bs(int32_t n, int32_t bw, const double *A_lu, int32_t nrhs, double *b) {
  int r;
  if (!A_lu || !b) return fail(PB_E_ARG, "null argument");
  if ((r = labs_check(n, bw, nrhs))) return r;
  const size_t ab = (size_t)n * (2 * bw + 1) * sizeof(double), bb = (size_t)nrhs * n * sizeof(double2);
  DevBuf da, db;
  HIPCHK(hipMalloc(&da.p, ab));
  HIPCHK(hipMalloc(&db.p, bb));
  HIPCHK(hipMemcpy(da.p, A_lu, ab, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(db.p, b, bb, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(pb::labs_bs_kernel, dim3(blocks_for(nrhs, 64)), dim3(64), 0, 0, n, bw, (const double *)da.p,
                     nrhs, (pb::cplx *)db.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpy(b, db.p, bb, hipMemcpyDeviceToHost));
  return PB_OK;
}

// ---- pb_labs_bs without allocation: the band factor and the right-hand
// sides are staged through `scratch`, file-scope device memory of
// PB_NOALLOC_SCRATCH_BYTES that every GPU holds (its own copy) until the
// process ends.  Host code uses the per-device address from
// hipGetSymbolAddress, never the array's name.
static __device__ __align__(256) double scratch[PB_NOALLOC_SCRATCH_BYTES / sizeof(double)];
static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[kMaxDevices];

// Address of `scratch` on the current device, looked up on the first call
// on that device.
static int scratch_ptr_current(double **out) {
  int dev = -1;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(PB_E_HIP, "device %d outside 0..%d", dev, kMaxDevices - 1);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(PB_E_HIP, "hipGetSymbolAddress(scratch) returned null on device %d", dev);
    g_scratch_ptr[dev] = static_cast<double *>(p);
  }
  *out = g_scratch_ptr[dev];
  return PB_OK;
}

int pb_labs_bs_noalloc(int32_t n, int32_t bw, const double *A_lu, int32_t nrhs, double *b) {
  int r;
  if (!A_lu || !b) return fail(PB_E_ARG, "null argument");
  if ((r = labs_check(n, bw, nrhs))) return r;
  constexpr size_t cap = PB_NOALLOC_SCRATCH_BYTES / sizeof(double);
  const size_t na = (size_t)n * (2 * bw + 1), nb = (size_t)nrhs * n * 2;
  const size_t off_b = (na + 31) & ~(size_t)31;  // right-hand sides 256-byte aligned after the factor
  if (off_b + nb > cap)
    return fail(PB_E_ARG, "pb_labs_bs_noalloc: factor and %d right-hand sides of %d need %zu bytes, above the %zu-byte "
                "scratch", nrhs, n, (off_b + nb) * sizeof(double), (size_t)PB_NOALLOC_SCRATCH_BYTES);
  double *ptr = nullptr;
  if ((r = scratch_ptr_current(&ptr))) return r;
  HIPCHK(hipMemcpy(ptr, A_lu, na * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(ptr + off_b, b, nb * sizeof(double), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(pb::labs_bs_kernel, dim3(blocks_for(nrhs, 64)), dim3(64), 0, 0, n, bw, (const double *)ptr, nrhs,
                     (pb::cplx *)(ptr + off_b));
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpy(b, ptr + off_b, nb * sizeof(double), hipMemcpyDeviceToHost));
  return PB_OK;
}

// ---- wall actuation and discrete adjoint --------------------------------------
int pb_set_control(pb_ctx *c, double alpha, const double *act) {
  if (!c) return fail(PB_E_ARG, "null context");
  if (alpha != 0.0 && c->opp_rows.empty())
    return fail(PB_E_ARG, "opposition control needs a detection plane inside the channel (opp_ydist in (0,1))");
  if (act) {
    if (!c->act) {
      const hipError_t e = hipMalloc((void **)&c->act, c->plane * sizeof(double2));
      if (e != hipSuccess) return fail(PB_E_HIP, "hipMalloc for the actuation: %s", hipGetErrorString(e));
      c->bytes += c->plane * sizeof(double2);
    }
    HIPCHK(hipMemcpy(c->act, act, c->plane * sizeof(double2), hipMemcpyHostToDevice));
  }
  c->act_on = act != nullptr;
  c->cfg.opp_alpha = alpha;
  drop_graph(c);  // the captured step holds the old control
  return PB_OK;
}

static int ensure_adj(pb_ctx *c) {
  if (c->adj_U) return PB_OK;
  auto alloc = [&](void **p, size_t bytes) -> int {
    const hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return fail(PB_E_HIP, "hipMalloc(%zu) for the adjoint: %s", bytes, hipGetErrorString(e));
    c->bytes += bytes;
    return PB_OK;
  };
  int r;
  if ((r = alloc((void **)&c->adj_U, 9 * c->field * sizeof(double2))) ||
      (r = alloc((void **)&c->adj_RK, 3 * c->field * sizeof(double2))) ||
      (r = alloc((void **)&c->adj_buf, pb::ADJ_SCRATCH * c->field * sizeof(double2))) ||
      (r = alloc((void **)&c->adj_fac, (size_t)(3 * (2 * c->W + 4) + c->W) * c->field * sizeof(double))) ||
      (r = alloc((void **)&c->adj_pad, 8 * c->pfield * sizeof(double2))) ||
      (r = alloc((void **)&c->adj_vd, 2 * c->plane * sizeof(double2))) ||
      (r = alloc((void **)&c->adj_act, c->plane * sizeof(double2))) ||
      (r = alloc((void **)&c->adj_alpha, c->plane * sizeof(double))))
    return r;
  return PB_OK;
}

extern "C++" {
template <class F>
static int launch_zmap_inv_t(pb_ctx *c, const double2 *src, double2 *dst, int nf, uint64_t map, double scale) {
  const size_t ldsz = (size_t)c->nz32 * c->Bz * sizeof(double2);
  hipLaunchKernelGGL(pb::zinv_kernel<F>, dim3(c->nxh / c->Bz, c->ny, nf), dim3(F::T), ldsz, c->stream, src, dst,
                     c->pz, c->twz, c->nxh, c->nz, c->ny, c->Bz, c->field, c->pfield, map, scale);
  HIPCHK(hipGetLastError());
  return PB_OK;
}
template <class F>
static int launch_zmap_fwd_t(pb_ctx *c, const double2 *src, double2 *dst, int nf, uint64_t map, double scale) {
  const size_t ldsz = (size_t)c->nz32 * c->Bz * sizeof(double2);
  hipLaunchKernelGGL(pb::zfwd_kernel<F>, dim3(c->nxh / c->Bz, c->ny, nf), dim3(F::T), ldsz, c->stream, src, dst,
                     c->pz, c->twz, c->nxh, c->nz, c->ny, c->Bz, c->field, c->pfield, scale, map);
  HIPCHK(hipGetLastError());
  return PB_OK;
}
template <class F>
static int launch_xadj_t(pb_ctx *c, const double2 *fpad) {
  const size_t ldsx = F::lds_elems(c->nx32, 2) * sizeof(double2);
  hipLaunchKernelGGL(pb::xadj_kernel<F>, dim3(c->nz32, c->ny), dim3(F::T), ldsx, c->stream, fpad, c->adj_pad,
                     c->px, c->twx, c->nxh, c->nz32, c->pfield);
  HIPCHK(hipGetLastError());
  return PB_OK;
}
}  // extern "C++"

// spectral slots -> padded fields (z inverse, times scale) / padded fields ->
// spectral slots (z forward, truncation, times scale)
static int launch_zmap_inv(pb_ctx *c, const double2 *src, double2 *dst, int nf, uint64_t map, double scale) {
  int r = PB_OK;
  const int ez = pick_ept((size_t)c->nz32 * c->Bz);
  if (ct_z(c)) r = launch_zmap_inv_t<pb::FftZ1152>(c, src, dst, nf, map, scale);
  else PB_EPT_DISPATCH(ez, r = launch_zmap_inv_t<F>(c, src, dst, nf, map, scale));
  return r;
}
static int launch_zmap_fwd(pb_ctx *c, const double2 *src, double2 *dst, int nf, uint64_t map, double scale) {
  int r = PB_OK;
  const int ez = pick_ept((size_t)c->nz32 * c->Bz);
  if (ct_z(c)) r = launch_zmap_fwd_t<pb::FftZ1152>(c, src, dst, nf, map, scale);
  else PB_EPT_DISPATCH(ez, r = launch_zmap_fwd_t<F>(c, src, dst, nf, map, scale));
  return r;
}
static int launch_xadj(pb_ctx *c, const double2 *fpad) {
  int r = PB_OK;
  const int ex = pick_ept((size_t)c->nx32 * 2);
  if (ct_x(c)) r = launch_xadj_t<pb::FftX1536>(c, fpad);
  else PB_EPT_DISPATCH(ex, r = launch_xadj_t<F>(c, fpad));
  return r;
}

// Transpose of the nonlinear stage around a state: the product adjoints in
// adj_U slots 0-4, 6-8 -> the state adjoints in adj_U slots 0,1,2,6.  fpad:
// the state's z inverse (4 padded fields) saved by the forward, or null:
// then it is formed here from U slots 0,1,2,6 into pad fields 0-3.
static int adj_nonlinear_dev(pb_ctx *c, const double2 *fpad) {
  int r;
  if (!fpad) {
    if ((r = launch_zinv_any(c))) return r;
    fpad = c->pad;
  }
  if ((r = launch_zmap_inv(c, c->adj_U, c->adj_pad, 8, pb::MAP_PROD & 0xFFFFFFFFull, c->rescale))) return r;
  if ((r = launch_xadj(c, fpad))) return r;
  return launch_zmap_fwd(c, c->adj_pad, c->adj_U, 4, pb::MAP_STATE, 1.0);
}

static pb::AdjFactors adj_factors(pb_ctx *c, int s) {
  pb::AdjFactors F;
  const size_t f = c->field, per = (size_t)(2 * c->W + 4) * f;
  F.luh = c->adj_fac + (size_t)(s - 1) * per;
  F.lut = F.luh + (size_t)c->W * f;
  F.hom = F.lut + (size_t)c->W * f;
  F.lup = c->adj_fac + 3 * per;
  F.plane = c->plane;
  F.field = f;
  F.W = c->W;
  return F;
}
// the factors of the three substeps at dt (formed again when dt changes)
static int adj_prepare(pb_ctx *c, double dt) {
  if (c->adj_fac_dt == dt) return PB_OK;
  for (int s = 1; s <= c->cfg.rk_step; ++s) {
    const pb::LineParams P = c->line_params(s, dt);
    hipLaunchKernelGGL(pb::adj_factor_kernel, dim3(blocks_for(c->nsys, 64)), dim3(64), 0, c->stream, P,
                       adj_factors(c, s), c->nsys, c->nxh, c->nh, c->lmfxr, c->lmfzr, (pb::cplx *)c->adj_buf);
    HIPCHK(hipGetLastError());
  }
  c->adj_fac_dt = dt;
  return PB_OK;
}
// The transposed wall-normal stage: the three chains of every system, the
// four parts of the transposed hvhg, their elementwise combination.  Built
// for bandwidth 4 (the order-7 B-splines of every benchmark configuration).
static int launch_adj_ws(pb_ctx *c, const pb::LineParams &P, const pb::AdjArgs &A) {
  if (c->bw != 4) return fail(PB_E_ARG, "the adjoint is built for bandwidth 4 (order-7 B-splines), not %d", c->bw);
  const unsigned nb = (unsigned)blocks_for(c->nsys, 64);
  hipLaunchKernelGGL(pb::adj_chain_kernel<4>, dim3(nb, 3), dim3(64), 0, c->stream, P, A, c->nsys);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(pb::adj_hvhg_part_kernel<4>, dim3(nb, 4), dim3(64), 0, c->stream, P, A, c->nsys);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(pb::adj_hvhg_combine_kernel, dim3(blocks_for(c->field, 256)), dim3(256), 0, c->stream, A,
                     c->field);
  HIPCHK(hipGetLastError());
  return PB_OK;
}

static pb::AdjArgs adj_args(pb_ctx *c, int s, const double2 *vd, double tau_bar, double q_bar) {
  pb::AdjArgs A;
  A.AU = (pb::cplx *)c->adj_U;
  A.ARK = (pb::cplx *)c->adj_RK;
  A.abuf = (pb::cplx *)c->adj_buf;
  A.F = adj_factors(c, s);
  A.vd = (const pb::cplx *)vd;
  A.avd = (pb::cplx *)c->adj_vd;
  A.aact = (pb::cplx *)c->adj_act;
  A.aalpha = c->adj_alpha;
  A.mixed = c->cfg.scalar_bc == PB_SCALAR_MIXED;
  A.ks = c->cfg.nu / c->cfg.pr / c->cfg.bulk_u;
  A.tau_bar = tau_bar;
  A.q_bar = q_bar;
  A.plane = c->plane;
  A.field = c->field;
  A.nxh = c->nxh;
  A.nz = c->nz;
  A.nh = c->nh;
  A.lmfxr = c->lmfxr;
  A.lmfzr = c->lmfzr;
  A.opp_alpha = c->cfg.opp_alpha;
  return A;
}

// Transpose of substep s (fixed dt) at the input state in U slots 0,1,2,6:
// adj_U slots 0,1,2,6 and adj_RK hold the adjoints of the substep's output
// state and stores on entry, of its input state and stores on return;
// adj_act and adj_alpha accumulate.  tau_bar, q_bar: adjoints of tau_w, q_w
// (used on the last substep).
// fpad, fvd: the input state's z inverse and detection planes saved by the
// forward, or null: then they are formed from U slots 0,1,2,6.
static int adj_substep_dev(pb_ctx *c, int s, double dt, double tau_bar, double q_bar, const double2 *fpad = nullptr,
                           const double2 *fvd = nullptr) {
  int r;
  const bool rows = !c->opp_rows.empty();
  if (rows && !fvd) {
    if ((r = launch_oppvd(c, c->U))) return r;
    fvd = c->vd;
  }
  if ((r = adj_prepare(c, dt))) return r;
  const pb::LineParams P = c->line_params(s, dt);
  if ((r = launch_adj_ws(c, P, adj_args(c, s, rows ? fvd : nullptr, tau_bar, q_bar)))) return r;
  if ((r = adj_nonlinear_dev(c, fpad))) return r;
  const bool opp = c->cfg.opp_alpha != 0.0, mixed = c->cfg.scalar_bc == PB_SCALAR_MIXED;
  if (opp || mixed) {
    hipLaunchKernelGGL(pb::adj_finish_kernel, dim3(blocks_for(c->field, 256)), dim3(256), 0, c->stream,
                       (pb::cplx *)c->adj_U, opp ? (const pb::cplx *)c->adj_vd : nullptr, c->drows, c->plane,
                       c->field, c->ny, mixed ? 1 : 0);
    HIPCHK(hipGetLastError());
  }
  return PB_OK;
}

static int adj_clear_controls(pb_ctx *c, const double *act_bar) {
  HIPCHK(hipMemsetAsync(c->adj_alpha, 0, c->plane * sizeof(double), c->stream));
  if (act_bar) HIPCHK(hipMemcpyAsync(c->adj_act, act_bar, c->plane * sizeof(double2), hipMemcpyHostToDevice, c->stream));
  else HIPCHK(hipMemsetAsync(c->adj_act, 0, c->plane * sizeof(double2), c->stream));
  return PB_OK;
}
// sum of the per-system d/d alpha terms (ascending system order) and the
// actuation adjoint, to the host
static int adj_read_controls(pb_ctx *c, double *alpha_sum, double *alpha_plane, double *act_bar) {
  std::vector<double> a(c->plane);
  HIPCHK(hipMemcpyAsync(a.data(), c->adj_alpha, c->plane * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (act_bar) HIPCHK(hipMemcpyAsync(act_bar, c->adj_act, c->plane * sizeof(double2), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (alpha_plane) std::memcpy(alpha_plane, a.data(), c->plane * sizeof(double));
  if (alpha_sum) {
    double t = 0.0;
    for (double x : a) t += x;
    *alpha_sum += t;
  }
  return PB_OK;
}

int pb_adj_wavespace_host(pb_ctx *c, int32_t s, double dt, double *AU, double *ark, const double *vd, double *avd,
                          double *act_bar, double *alpha_bar, double tau_bar, double q_bar) {
  if (!c || !AU || !ark) return fail(PB_E_ARG, "null argument");
  if (s < 1 || s > c->cfg.rk_step) return fail(PB_E_ARG, "substep %d out of range", s);
  int r;
  if ((r = check_dt(dt)) || (r = ensure_adj(c))) return r;
  static const int map9[9] = {0, 1, 2, 3, 4, 5, 6, 7, 8};
  if ((r = upload_ref(c, AU, 9, map9, c->adj_U)) || (r = upload_ref(c, ark, 3, kRKMap, c->adj_RK))) return r;
  if (vd) HIPCHK(hipMemcpyAsync(c->vd, vd, 2 * c->plane * sizeof(double2), hipMemcpyHostToDevice, c->stream));
  if ((r = adj_clear_controls(c, act_bar)) || (r = adj_prepare(c, dt))) return r;
  const pb::LineParams P = c->line_params(s, dt);
  if ((r = launch_adj_ws(c, P, adj_args(c, s, vd ? c->vd : nullptr, tau_bar, q_bar)))) return r;
  if ((r = download_ref(c, AU, 9, map9, c->adj_U)) || (r = download_ref(c, ark, 3, kRKMap, c->adj_RK))) return r;
  if (avd) HIPCHK(hipMemcpy(avd, c->adj_vd, 2 * c->plane * sizeof(double2), hipMemcpyDeviceToHost));
  return adj_read_controls(c, nullptr, alpha_bar, act_bar);
}

int pb_adj_nonlinear_host(pb_ctx *c, const double *U, const double *aprod, double *astate) {
  if (!c || !U || !aprod || !astate) return fail(PB_E_ARG, "null argument");
  int r;
  if ((r = ensure_adj(c))) return r;
  static const int map9[9] = {0, 1, 2, 3, 4, 5, 6, 7, 8};
  if ((r = upload_ref(c, U, 4, kStateMap, c->U)) || (r = upload_ref(c, aprod, 9, map9, c->adj_U))) return r;
  c->have_state = true;
  if ((r = adj_nonlinear_dev(c, nullptr))) return r;
  return download_ref(c, astate, 4, kStateMap, c->adj_U);
}

int pb_adj_substep(pb_ctx *c, int32_t s, double dt, const double *state, double *astate, double *ark,
                   double tau_bar, double q_bar, double *alpha_bar, double *act_bar) {
  if (!c || !state || !astate || !ark) return fail(PB_E_ARG, "null argument");
  if (s < 1 || s > c->cfg.rk_step) return fail(PB_E_ARG, "substep %d out of range", s);
  int r;
  if ((r = check_dt(dt)) || (r = ensure_adj(c))) return r;
  if ((r = upload_ref(c, state, 4, kStateMap, c->U))) return r;
  c->have_state = true;
  if ((r = upload_ref(c, astate, 4, kStateMap, c->adj_U)) || (r = upload_ref(c, ark, 3, kRKMap, c->adj_RK))) return r;
  if ((r = adj_clear_controls(c, act_bar))) return r;
  if ((r = adj_substep_dev(c, s, dt, tau_bar, q_bar))) return r;
  if ((r = download_ref(c, astate, 4, kStateMap, c->adj_U)) || (r = download_ref(c, ark, 3, kRKMap, c->adj_RK)))
    return r;
  return adj_read_controls(c, alpha_bar, nullptr, act_bar);
}

// saved states of the gradient driver: checkpoint q = U slots 0,1,2,6 and the
// three stores (7 fields), substep inputs = U slots 0,1,2,6 (4 fields)
static int copy_slots(pb_ctx *c, double2 *dst, const double2 *src, const int *slots, int n, bool to_saved) {
  for (int i = 0; i < n; ++i) {
    double2 *d = to_saved ? dst + (size_t)i * c->field : dst + (size_t)slots[i] * c->field;
    const double2 *s = to_saved ? src + (size_t)slots[i] * c->field : src + (size_t)i * c->field;
    HIPCHK(hipMemcpyAsync(d, s, c->field * sizeof(double2), hipMemcpyDeviceToDevice, c->stream));
  }
  return PB_OK;
}
static int save_state(pb_ctx *c, double2 *dst, bool with_rk) {
  static const int st[4] = {0, 1, 2, 6}, rk[3] = {0, 1, 2};
  int r;
  if ((r = copy_slots(c, dst, c->U, st, 4, true))) return r;
  return with_rk ? copy_slots(c, dst + 4 * c->field, c->RK, rk, 3, true) : PB_OK;
}
static int restore_state(pb_ctx *c, const double2 *src, bool with_rk) {
  static const int st[4] = {0, 1, 2, 6}, rk[3] = {0, 1, 2};
  int r;
  if ((r = copy_slots(c, c->U, src, st, 4, false))) return r;
  return with_rk ? copy_slots(c, c->RK, src + 4 * c->field, rk, 3, false) : PB_OK;
}

// substep s of the forward, keeping what its transpose needs: the input
// state's z inverse (pad fields 0-3 after the z inverse) into pad_dst and
// its detection planes into vd_dst (when the channel has them)
static int substep_saving(pb_ctx *c, int s, double2 *pad_dst, double2 *vd_dst) {
  int r;
  if (!c->opp_rows.empty()) {
    if ((r = launch_oppvd(c, c->U))) return r;
    HIPCHK(hipMemcpyAsync(vd_dst, c->vd, 2 * c->plane * sizeof(double2), hipMemcpyDeviceToDevice, c->stream));
  }
  if (c->usrc)
    HIPCHK(hipMemcpyAsync(c->usrc, c->U, c->field * sizeof(double2), hipMemcpyDeviceToDevice, c->stream));
  if ((r = launch_zinv_any(c))) return r;
  HIPCHK(hipMemcpyAsync(pad_dst, c->pad, 4 * c->pfield * sizeof(double2), hipMemcpyDeviceToDevice, c->stream));
  if ((r = launch_xphys_any(c, 0)) || (r = launch_velmax(c, s == 1 ? (pb::VM_STORE | pb::VM_DT) : 0)) ||
      (r = launch_zfwd_any(c, 0, c->rescale)))
    return r;
  return launch_wavespace(c, s);
}

int pb_gradient(pb_ctx *c, int32_t nsteps, int32_t ckpt_every, double dt, double wq, double *J, double *dJ_dalpha,
                double *dJ_da) {
  if (!c) return fail(PB_E_ARG, "null context");
  if (!c->have_state) return fail(PB_E_STATE, "pb_gradient before pb_set_state");
  if (nsteps < 1 || ckpt_every < 1) return fail(PB_E_ARG, "nsteps %d, ckpt_every %d", nsteps, ckpt_every);
  if (c->cfg.dt_factor != 1.0) return fail(PB_E_ARG, "the adjoint needs a fixed dt (dt_factor = 1)");
  int r;
  if ((r = check_dt(dt)) || (r = ensure_adj(c))) return r;
  const int S = c->cfg.rk_step, C = std::min(ckpt_every, nsteps), nck = (nsteps + C - 1) / C;
  // saved data, in double2 units: checkpoints (7 fields each) and the final
  // state (7 fields), then per substep of one segment either the z inverse
  // of its input state and its detection planes (4 padded fields + 2
  // planes), or its input state (4 fields) when that does not fit in a
  // quarter of the free memory (or PB_ADJ_SAVE=state)
  const size_t fixed = (size_t)(nck + 1) * 7 * c->field;
  const size_t per_pad = 4 * c->pfield + 2 * c->plane, per_state = 4 * c->field;
  size_t free_b = 0, total_b = 0;
  (void)hipMemGetInfo(&free_b, &total_b);
  const char *mode = std::getenv("PB_ADJ_SAVE");
  const bool force_state = mode && std::strcmp(mode, "state") == 0;
  const size_t held = c->adj_nstates * sizeof(double2);
  const bool pad_path = !force_state && (fixed + (size_t)C * S * per_pad) * sizeof(double2) <= (free_b + held) / 4;
  const size_t need = fixed + (size_t)C * S * (pad_path ? per_pad : per_state);
  if (c->adj_nstates < need) {
    if (c->adj_states) {
      (void)hipFree(c->adj_states);
      c->bytes -= c->adj_nstates * sizeof(double2);
      c->adj_states = nullptr;
      c->adj_nstates = 0;
    }
    const hipError_t e = hipMalloc((void **)&c->adj_states, need * sizeof(double2));
    if (e != hipSuccess)
      return fail(PB_E_HIP, "hipMalloc(%zu) for the saved states: %s", need * sizeof(double2), hipGetErrorString(e));
    c->adj_nstates = need;
    c->bytes += need * sizeof(double2);
  }
  double2 *ckpt = c->adj_states, *fin = ckpt + (size_t)nck * 7 * c->field, *seg = fin + 7 * c->field;
  std::vector<pb_substep_out> out(S);
  // forward: J and the checkpoints; the last segment keeps what its
  // reverse pass needs right away (it is not recomputed)
  double jsum = 0.0;
  const int klast = (nck - 1) * C;
  for (int k = 0; k < nsteps; ++k) {
    if (k % C == 0 && (r = save_state(c, ckpt + (size_t)(k / C) * 7 * c->field, true))) return r;
    double dtv = dt;
    if ((r = upload_dt(c, dt))) return r;
    for (int s = 1; s <= S; ++s) {
      if (k < klast) {
        r = substep_async(c, s);
      } else {
        const size_t i = (size_t)(k - klast) * S + s - 1;
        if (pad_path) {
          double2 *p = seg + i * per_pad;
          r = substep_saving(c, s, p, p + 4 * c->pfield);
        } else if (!(r = save_state(c, seg + i * per_state, false))) {
          r = substep_async(c, s);
        }
      }
      if (r) return r;
    }
    if ((r = read_outputs(c, 1, S, out.data(), &dtv))) return r;
    jsum += out[S - 1].tau_w + wq * out[S - 1].q_w;
  }
  if ((r = save_state(c, fin, true))) return r;
  if (J) *J = jsum / nsteps;
  // reverse: segments last to first, each recomputed from its checkpoint
  HIPCHK(hipMemsetAsync(c->adj_U, 0, 9 * c->field * sizeof(double2), c->stream));
  HIPCHK(hipMemsetAsync(c->adj_RK, 0, 3 * c->field * sizeof(double2), c->stream));
  if ((r = adj_clear_controls(c, nullptr))) return r;
  const double tb = 1.0 / nsteps, qb = wq / nsteps;
  for (int q = nck - 1; q >= 0; --q) {
    const int k0 = q * C, k1 = std::min(nsteps, k0 + C);
    if (q != nck - 1 && (r = restore_state(c, ckpt + (size_t)q * 7 * c->field, true))) return r;
    for (int k = k0; k < k1 && q != nck - 1; ++k) {
      if ((r = upload_dt(c, dt))) return r;
      for (int s = 1; s <= S; ++s) {
        const size_t i = (size_t)(k - k0) * S + s - 1;
        if (pad_path) {
          double2 *p = seg + i * per_pad;
          if (k == k1 - 1 && s == S) {  // the last substep: only its input's z inverse and planes
            if (!c->opp_rows.empty()) {
              if ((r = launch_oppvd(c, c->U))) return r;
              HIPCHK(hipMemcpyAsync(p + 4 * c->pfield, c->vd, 2 * c->plane * sizeof(double2),
                                    hipMemcpyDeviceToDevice, c->stream));
            }
            if ((r = launch_zinv_any(c))) return r;
            HIPCHK(hipMemcpyAsync(p, c->pad, 4 * c->pfield * sizeof(double2), hipMemcpyDeviceToDevice, c->stream));
            break;
          }
          if ((r = substep_saving(c, s, p, p + 4 * c->pfield))) return r;
        } else {
          if ((r = save_state(c, seg + i * per_state, false))) return r;
          if (k == k1 - 1 && s == S) break;  // the last substep's output is not needed
          if ((r = substep_async(c, s))) return r;
        }
      }
    }
    for (int k = k1 - 1; k >= k0; --k)
      for (int s = S; s >= 1; --s) {
        const size_t i = (size_t)(k - k0) * S + s - 1;
        const double a = s == S ? tb : 0.0, b = s == S ? qb : 0.0;
        if (pad_path) {
          const double2 *p = seg + i * per_pad;
          if ((r = adj_substep_dev(c, s, dt, a, b, p, c->opp_rows.empty() ? nullptr : p + 4 * c->pfield))) return r;
        } else {
          if ((r = restore_state(c, seg + i * per_state, false))) return r;
          if ((r = adj_substep_dev(c, s, dt, a, b))) return r;
        }
      }
  }
  double asum = 0.0;
  if ((r = adj_read_controls(c, &asum, nullptr, dJ_da))) return r;
  if (dJ_dalpha) *dJ_dalpha = asum;
  // leave the context at the state after the window, as a forward run would
  if ((r = restore_state(c, fin, true))) return r;
  HIPCHK(hipStreamSynchronize(c->stream));
  return PB_OK;
}

int pb_labs_mat_mul(int32_t n, int32_t bw, const double *A, int32_t nrhs, const double *x, double *b) {
  int r;
  if (!A || !x || !b) return fail(PB_E_ARG, "null argument");
  if ((r = labs_check(n, bw, nrhs))) return r;
  const size_t ab = (size_t)n * (2 * bw + 1) * sizeof(double), vb = (size_t)nrhs * n * sizeof(double2);
  DevBuf da, dx, db;
  HIPCHK(hipMalloc(&da.p, ab));
  HIPCHK(hipMalloc(&dx.p, vb));
  HIPCHK(hipMalloc(&db.p, vb));
  HIPCHK(hipMemcpy(da.p, A, ab, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dx.p, x, vb, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(pb::labs_mv_kernel, dim3(blocks_for(nrhs, 64)), dim3(64), 0, 0, n, bw, (const double *)da.p,
                     nrhs, (const pb::cplx *)dx.p, (pb::cplx *)db.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpy(b, db.p, vb, hipMemcpyDeviceToHost));
  return PB_OK;
}

}  // extern "C"
