// lgame_capi.hip -- C-ABI of the MI355X HEVI step (include/lgame_gpu.h): handle, device memory,
// layout conversion between the reference's Fortran arrays and the device layout, the step
// sequence, budget sums and the finite check.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/lgame_gpu.h"
#include "lgame_ops.h"
#include "lgame_step.h"
#include "lgame_batch.h"
#include "lgame_grid.h"

using namespace lgame;

struct lgame_handle {
  lgame_dims dims;
  lgame_cfg cfg;
  Ctx c;                         // plain form; its grid (FP64) also serves the budget sums
  CtxT<float> cf;                // mixed form: FP32 grid copy and work arrays, FP64 grid in g64
  int mixed = 0;
  Tiles tiles;                   // block shapes of the stencil kernels (lgame_autotune)
  size_t state_bytes = 0;        // the two state slots: the first bytes of the slab
  size_t work_off = 0, work_bytes = 0;  // column scratch and work arrays (zero at the first step)
  StateDev slot[2];
  char *slab = nullptr;          // all fields, one allocation
  size_t slab_bytes = 0;
  double *staging = nullptr;     // device staging for host<->device transposes
  size_t staging_elems = 0;
  double *red = nullptr;         // reduction partials
  unsigned long long *counter = nullptr;  // [0] lgame_check_finite's count, [1] the step's non-finite flag
  double *trsk = nullptr;        // TRSK weights, FP64 then FP32
  double *soil_z = nullptr;      // soil grid: interfaces, centres, z_t_const (with the prognostic soil)
  hipStream_t stream = nullptr;
  int have_state[2] = {0, 0};
  int stepped = 0;
  // optional event timing of kernel groups (LGAME_PROF_*)
  int prof = 0;
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<int, std::pair<int, int>>> ev_marks;  // group, (start, stop) pool index
  size_t ev_next = 0;
  // radiation: per-column data (one allocation), surrogate and its audit
  double *rad_mem = nullptr;
  float *rad_x = nullptr, *rad_y = nullptr;
  unsigned long long *rad_counts = nullptr;  // audited, fallbacks (device)
  double *rad_sq = nullptr;                  // squared-error sum (device)
  int *rad_list = nullptr;                   // the audit pass's columns (device)
  unsigned *rad_list_count = nullptr;
  int rad_surface = 0;
  lgame_mlp *mlp = nullptr;
  MlpWork *mlp_work = nullptr;  // this handle's inference buffers (made by lgame_rad_attach_mlp)
  double rad_tol = 0.0;
  unsigned rad_thresh = 0;
  unsigned long long rad_seed = 0, rad_calls = 0;
  std::string err;
};

namespace {

thread_local std::string g_init_error;

int fail(lgame_handle *h, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (h)
    h->err = buf;
  else
    g_init_error = buf;
  return code;
}

#define HIPCHK(h, call)                                                                         \
  do {                                                                                          \
    hipError_t e_ = (call);                                                                     \
    if (e_ != hipSuccess) return fail(h, LGAME_EHIP, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

// Fortran array (n1, n2, n3) with n1 = y extent, n2 = x extent, optional leading component
// dimension of size `stride` -> device layout
__global__ void k_from_ref(double *dst, const double *src, int n1, int n2, int n3, int stride, int comp, Geom G) {
  const int jk = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int ji = blockIdx.y * blockDim.y + threadIdx.y + 1;
  const int jl = blockIdx.z + 1;
  if (jk > n2 || ji > n1 || jl > n3) return;
  const size_t s = (size_t)comp + (size_t)stride * ((size_t)(ji - 1) + (size_t)n1 * ((size_t)(jk - 1) + (size_t)n2 * (size_t)(jl - 1)));
  dst[LG_IX(G, ji, jk, jl)] = src[s];
}

template <class T>
__global__ void k_to_ref(double *dst, const T *src, int n1, int n2, int n3, Geom G) {
  const int jk = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int ji = blockIdx.y * blockDim.y + threadIdx.y + 1;
  const int jl = blockIdx.z + 1;
  if (jk > n2 || ji > n1 || jl > n3) return;
  dst[(size_t)(ji - 1) + (size_t)n1 * ((size_t)(jk - 1) + (size_t)n2 * (size_t)(jl - 1))] = (double)src[LG_IX(G, ji, jk, jl)];
}

// c1*a + c2*b of one field, written in the reference layout (linear_combine_two_states,
// mo_linear_combination.F90:27-35); with tb, the field is rhotheta/rho_gas - theta_v_bg of the
// combined rhotheta (a, b) and gas densities (ga, gb) instead
__global__ void k_combine_to_ref(double *dst, const double *a, const double *b, const double *ga, const double *gb, const double *tb,
                                 double c1, double c2, int n1, int n2, int n3, Geom G) {
  const int jk = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int ji = blockIdx.y * blockDim.y + threadIdx.y + 1;
  const int jl = blockIdx.z + 1;
  if (jk > n2 || ji > n1 || jl > n3) return;
  const ix_t i = LG_IX(G, ji, jk, jl);
  double v = c1 * a[i] + c2 * b[i];
  if (tb) v = v / (c1 * ga[i] + c2 * gb[i]) - tb[i];
  dst[(size_t)(ji - 1) + (size_t)n1 * ((size_t)(jk - 1) + (size_t)n2 * (size_t)(jl - 1))] = v;
}

// FP32 copy of a block of FP64 fields (the grid of the mixed form)
__global__ void k_to_float(float *dst, const double *src, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) dst[i] = (float)src[i];
}

// Klemp damping prefactor at w levels, 1 + damping_coeff*dtime (mo_column_solvers.F90:190-196)
double klemp_prefactor(const lgame_cfg &cf, double z) {
  const double start = cf.klemp_begin_rel * cf.toa;
  const double above = z - start;
  double coeff;
  if (above < 0.0 || !cf.lklemp) {
    coeff = 0.0;
  } else {
    const double sn = std::sin(0.5 * MPI * above / (cf.toa - start));
    coeff = cf.klemp_damp_max * (sn * sn);
  }
  return 1.0 + coeff * cf.dtime;
}

// diag%roughness_length = grid%roughness_length (control.F90:456-459): the carried surface
// roughness starts from the grid's, at lgame_init and whenever the work arrays are zeroed again
hipError_t init_surface(lgame_handle *h) {
  if (!h->c.C.lsfc_turb) return hipSuccess;
  const size_t n = (size_t)h->c.G.lvl;
  hipError_t e = h->mixed ? hipMemcpyAsync(h->cf.w.sfc.p, h->cf.g.roughness_length.p, n * sizeof(float), hipMemcpyDeviceToDevice, h->stream)
                          : hipMemcpyAsync(h->c.w.sfc.p, h->c.g.roughness_length.p, n * sizeof(double), hipMemcpyDeviceToDevice, h->stream);
  return e == hipSuccess ? hipStreamSynchronize(h->stream) : e;
}

int upload(lgame_handle *h, double *dst, const double *host, int n1, int n2, int n3, int stride = 1, int ncomp = 1) {
  const size_t n = (size_t)n1 * n2 * n3 * stride;
  if (n > h->staging_elems) return fail(h, LGAME_EINVAL, "internal: staging buffer too small");
  HIPCHK(h, hipMemcpyAsync(h->staging, host, n * sizeof(double), hipMemcpyHostToDevice, h->stream));
  dim3 blk(BX, BY, 1), grd((n2 + BX - 1) / BX, (n1 + BY - 1) / BY, n3);
  for (int comp = 0; comp < ncomp; ++comp)
    hipLaunchKernelGGL(k_from_ref, grd, blk, 0, h->stream, dst + (size_t)comp * h->c.G.fsize, (const double *)h->staging, n1, n2, n3,
                       stride, comp, h->c.G);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

template <class T>
int download(lgame_handle *h, double *host, const T *src, int n1, int n2, int n3) {
  const size_t n = (size_t)n1 * n2 * n3;
  if (n > h->staging_elems) return fail(h, LGAME_EINVAL, "internal: staging buffer too small");
  dim3 blk(BX, BY, 1), grd((n2 + BX - 1) / BX, (n1 + BY - 1) / BY, n3);
  hipLaunchKernelGGL(k_to_ref<T>, grd, blk, 0, h->stream, h->staging, src, n1, n2, n3, h->c.G);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipMemcpyAsync(host, h->staging, n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

// ---- budget sums (write_out_integrals, mo_write_out.F90:226-304), two-sum compensated -------
struct DD {
  double s, c;
};
__device__ __forceinline__ void dd_add(DD &a, double x) {
  const double t = a.s + x;
  const double bp = t - a.s;
  const double err = (a.s - (t - bp)) + (x - bp);
  a.s = t;
  a.c = a.c + err;
}

constexpr int NRED = 4 + MAX_CONST;
constexpr int RED_BLOCK = 256;
constexpr int RED_GRID = 512;

template <class T>
__global__ void __launch_bounds__(RED_BLOCK) k_integrals(const Ctx c, const StateDev s, Fld<const T> temperature, double *partial) {
  DD acc[NRED];
#pragma unroll
  for (int q = 0; q < NRED; ++q) acc[q] = {0.0, 0.0};
  const int nx = c.G.nx, ny = c.G.ny, nl = c.G.nl, ncond = c.G.ncond, nc = c.G.nc;
  const long long ncell = (long long)nx * ny * nl;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < ncell; t += (long long)gridDim.x * blockDim.x) {
    const int jk = (int)(t % nx) + 1;
    const int ji = (int)((t / nx) % ny) + 1;
    const int jl = (int)(t / ((long long)nx * ny)) + 1;
    const ix_t i = LG_IX(c.G, ji, jk, jl);
    const double V = c.g.volume[i];
    const double rg = s.rho[ncond][i];
    dd_add(acc[0], s.rhotheta_v[i] * V);
    dd_add(acc[1], (rg * ke_point(c, s, ji, jk, jl)) * V);
    dd_add(acc[2], (rg * c.g.gravity_potential[i]) * V);
    dd_add(acc[3], (C_D_V * rg * (double)temperature[i]) * V);
    for (int jc = 0; jc < nc; ++jc) dd_add(acc[4 + jc], s.rho[jc][i] * V);
  }
  __shared__ double sh_s[RED_BLOCK], sh_c[RED_BLOCK];
  for (int q = 0; q < 4 + nc; ++q) {
    sh_s[threadIdx.x] = acc[q].s;
    sh_c[threadIdx.x] = acc[q].c;
    __syncthreads();
    for (int st = RED_BLOCK / 2; st > 0; st >>= 1) {
      if ((int)threadIdx.x < st) {
        DD a = {sh_s[threadIdx.x], sh_c[threadIdx.x]};
        dd_add(a, sh_s[threadIdx.x + st]);
        a.c = a.c + sh_c[threadIdx.x + st];
        sh_s[threadIdx.x] = a.s;
        sh_c[threadIdx.x] = a.c;
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      partial[2 * (q * RED_GRID + blockIdx.x)] = sh_s[0];
      partial[2 * (q * RED_GRID + blockIdx.x) + 1] = sh_c[0];
    }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(RED_BLOCK) k_integrals_final(const double *partial, int nq, double *out) {
  __shared__ double sh_s[RED_BLOCK], sh_c[RED_BLOCK];
  for (int q = 0; q < nq; ++q) {
    DD a = {0.0, 0.0};
    for (int b = threadIdx.x; b < RED_GRID; b += RED_BLOCK) {
      dd_add(a, partial[2 * (q * RED_GRID + b)]);
      a.c = a.c + partial[2 * (q * RED_GRID + b) + 1];
    }
    sh_s[threadIdx.x] = a.s;
    sh_c[threadIdx.x] = a.c;
    __syncthreads();
    for (int st = RED_BLOCK / 2; st > 0; st >>= 1) {
      if ((int)threadIdx.x < st) {
        DD x = {sh_s[threadIdx.x], sh_c[threadIdx.x]};
        dd_add(x, sh_s[threadIdx.x + st]);
        x.c = x.c + sh_c[threadIdx.x + st];
        sh_s[threadIdx.x] = x.s;
        sh_c[threadIdx.x] = x.c;
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) out[q] = sh_s[0] + sh_c[0];
    __syncthreads();
  }
}

// number of non-finite values in the prognostic fields of a slot
__global__ void k_count_nonfinite(const Ctx c, const StateDev s, unsigned long long *count) {
  const int jk = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int ji = blockIdx.y * blockDim.y + threadIdx.y + 1;
  const int jl = blockIdx.z + 1;
  const int nx = c.G.nx, ny = c.G.ny, nl = c.G.nl;
  if (jk > nx + 1 || ji > ny + 1 || jl > nl + 1) return;
  const ix_t i = LG_IX(c.G, ji, jk, jl);
  unsigned long long bad = 0;
  if (ji <= ny && jk <= nx && jl <= nl) {
    for (int jc = 0; jc < c.G.nc; ++jc) bad += !isfinite(s.rho[jc][i]);
    bad += !isfinite(s.rhotheta_v[i]) + !isfinite(s.theta_v_pert[i]) + !isfinite(s.exner_pert[i]);
  }
  if (ji <= ny && jl <= nl) bad += !isfinite(s.wind_u[i]);
  if (jk <= nx && jl <= nl) bad += !isfinite(s.wind_v[i]);
  if (ji <= ny && jk <= nx) bad += !isfinite(s.wind_w[i]);
  if (bad) atomicAdd(count, bad);
}

inline dim3 grid3(int nxr, int nyr, int nzr) { return dim3((nxr + BX - 1) / BX, (nyr + BY - 1) / BY, nzr); }

}  // namespace

// ---------------------------------------------------------------------------------------------
// the step sequence, pchevi (mo_manage_pchevi.F90:27-118)
// ---------------------------------------------------------------------------------------------
namespace lgame {

template <class R>
int enqueue_pchevi(const CtxT<R> &c, const StateDev &old, const StateDev &nw, int step, hipStream_t st, ProfHook *prof,
                    const RadCall *rad, const Tiles &tiles) {
  const int nx = c.G.nx, ny = c.G.ny, nl = c.G.nl, nlev = c.G.nlev;
  const bool o3 = c.C.order == 3;
  const bool pk = prof && prof->kernels;
  // a stencil kernel on a (64 x by) block grid covering (ex, ey, ez) points, by from the tile table
#define LG_LAUNCH(KI, kern, ex, ey, ez, ...)                                        \
  do {                                                                              \
    const int by_ = tiles.by[KI];                                                   \
    const dim3 b_(BX, by_, 1), g_(((ex) + BX - 1) / BX, ((ey) + by_ - 1) / by_, (ez)); \
    if (pk) prof->begin(PROF_KERNEL0 + (KI));                                       \
    hipLaunchKernelGGL(kern<R>, g_, b_, 0, st, __VA_ARGS__);                        \
    if (pk) prof->end(PROF_KERNEL0 + (KI));                                         \
  } while (0)

  // the same on a 1-D grid, for kernels that take their tile from tile_zfast
#define LG_LAUNCH_ZFAST(KI, kern, ex, ey, ez, ...)                                           \
  do {                                                                                     \
    const int by_ = tiles.by[KI];                                                          \
    const dim3 b_(BX, by_, 1), g_((((ex) + BX - 1) / BX) * (((ey) + by_ - 1) / by_) * (ez)); \
    if (pk) prof->begin(PROF_KERNEL0 + (KI));                                              \
    hipLaunchKernelGGL(kern<R>, g_, b_, 0, st, __VA_ARGS__);                               \
    if (pk) prof->end(PROF_KERNEL0 + (KI));                                                \
  } while (0)

  LG_LAUNCH(KI_TEMPERATURE, k_temperature, nx, ny, nl, c, old);
  // update_sfc_turb_quantities with state_old (mo_manage_pchevi.F90:52-54)
  if (c.C.lsfc_turb) LG_LAUNCH(KI_DIFF_DUAL, k_sfc_turb, nx, ny, 1, c, old);
  // update_rad_fluxes with state_old (mo_manage_pchevi.F90:56-58)
  if (rad) {
    if (prof) prof->begin(3);
    const int rc = enqueue_radiation(c, old, *rad, st);
    if (prof) prof->end(3);
    if (rc) return rc;
  }
  for (int rk = 1; rk <= 2; ++rk) {
    const StateDev &S = rk == 1 ? old : nw;
    if (rk == 1) {
      // manage_p_grad (mo_p_grad.F90:20-104); theta_v_adv_3rd_order of theta_v and theta_v_pert
      if (o3) {
        LG_LAUNCH(KI_THETA_GRAD, k_theta_grad, nx + 1, ny + 1, nl, c, old, 2);
        LG_LAUNCH(KI_THETA_DIV, k_theta_div, nx, ny, nl, c, 2);
        LG_LAUNCH(KI_THETA_EDGE, k_theta_edge, nx + 1, ny + 1, nl, c, old, 2);
      }
    }
    // vector_tend_expl (mo_vector_tend_expl.F90:30-182)
    bool winds_done = false;  // the wind update was made with the vorticity flux term
    if ((rk == 2 || step == 0) && (!c.C.llinear || c.C.lcorio)) {
      // the potential vorticity, and from the same densities and winds the mass fluxes and the
      // kinetic energy
      if (nx % 2 == 0)
        LG_LAUNCH_ZFAST(KI_POT_VORT, k_pot_vort2, nx / 2, ny, nlev, c, S, c.C.llinear ? 0 : 1);
      else
        LG_LAUNCH_ZFAST(KI_POT_VORT, k_pot_vort, nx, ny, nlev, c, S, c.C.llinear ? 0 : 1);
      // the vorticity flux term and, in the same cells, the gradient of the kinetic energy; at
      // rk 2 with the wind update, which follows it directly
      if (rk == 2) {
        LG_LAUNCH_ZFAST(KI_VORT_FLUX, k_vort_flux_vt, nx, ny, nl, c, old, nw, c.C.llinear ? 0 : 1);
        winds_done = true;
      } else {
        LG_LAUNCH_ZFAST(KI_VORT_FLUX, k_vort_flux, nx, ny, nl, c, c.C.llinear ? 0 : 1);
      }
    }
    // momentum diffusion and dissipation, first substep only (mo_vector_tend_expl.F90:101-126)
    if (rk == 1 && c.C.lmom_h) {
      LG_LAUNCH_ZFAST(KI_DIFF_ZETA, k_diff_zeta, nx + 1, ny + 1, nl, c, old);
      // the coefficients and the divergence take two cells per thread in the FP64 form only: in the
      // FP32 form the pair kernels were slower (0.80 against 1.06 and 0.72 against 0.86 ms per step
      // at 512x512x60)
      if constexpr (sizeof(R) == 8) {
        if (nx % 2 == 0)
          LG_LAUNCH_ZFAST(KI_DIFF_COEF, k_diff_coef2, nx / 2, ny, nl, c, old);
        else
          LG_LAUNCH_ZFAST(KI_DIFF_COEF, k_diff_coef, nx, ny, nl, c, old);
      } else {
        LG_LAUNCH_ZFAST(KI_DIFF_COEF, k_diff_coef, nx, ny, nl, c, old);
      }
      LG_LAUNCH(KI_DIFF_DUAL, k_diff_dual, nx, ny, nl, c, old);
      LG_LAUNCH_ZFAST(KI_DIFF_MOM, k_diff_mom, nx, ny, nl, c, old);
    }
    if (rk == 1 && c.C.lpbl) LG_LAUNCH(KI_DIFF_DUAL, k_pbl_wind, nx, ny, 1, c, old);
    if (rk == 1 && (c.C.lklemp || c.C.lmom_h || c.C.lmom_v || c.C.lpbl))
      LG_LAUNCH(KI_DISSIPATION, k_dissipation, nx, ny, nl, c, old);
    // the first substep's pressure gradient (manage_p_grad) and, from its values, the wind update;
    // no kernel in between reads or writes what the pressure gradient uses or forms
    if (rk == 1)
      LG_LAUNCH_ZFAST(KI_PGRAD, k_pgrad_vt, nx, ny, nl, c, old, nw, step == 0 ? 1 : 0);
    else if (!winds_done)
      LG_LAUNCH(KI_VEC_TEND, k_vec_tend, nx + 1, ny + 1, nl, c, old, nw, rk);
    // temperature and mass diffusion with the coefficients of this substep (mo_scalar_tend_expl.F90:
    // 60-99).  The temperature diffusion of the second substep repeats the first one with the same
    // temperature and coefficients, so its result is kept.
    if (rk == 1 && (c.C.ltemp_h || c.C.lmass_h)) {
      if (nx % 2 == 0)
        LG_LAUNCH_ZFAST(KI_DIFF_FLUX, k_diff_flux2, nx / 2, ny, nl, c, old);
      else
        LG_LAUNCH_ZFAST(KI_DIFF_FLUX, k_diff_flux, nx, ny, nl, c, old);
      if constexpr (sizeof(R) == 8) {
        if (nx % 2 == 0)
          LG_LAUNCH_ZFAST(KI_DIFF_DIV, k_diff_div2, nx / 2, ny, nl, c);
        else
          LG_LAUNCH_ZFAST(KI_DIFF_DIV, k_diff_div, nx, ny, nl, c);
      } else {
        LG_LAUNCH_ZFAST(KI_DIFF_DIV, k_diff_div, nx, ny, nl, c);
      }
    }
    // scalar_tend_expl (mo_scalar_tend_expl.F90:24-181); at rk 1 the theta_v edge values of the
    // pressure gradient are reused (same state, same winds)
    if (o3 && rk == 2) {
      LG_LAUNCH(KI_THETA_GRAD, k_theta_grad, nx + 1, ny + 1, nl, c, nw, 1);
      LG_LAUNCH(KI_THETA_DIV, k_theta_div, nx, ny, nl, c, 1);
      LG_LAUNCH(KI_THETA_EDGE, k_theta_edge, nx + 1, ny + 1, nl, c, nw, 1);
    }
    if (pk) prof->begin(PROF_KERNEL0 + KI_SCALAR_DIV);
    launch_scalar_col(c, S, nw, rk, tiles.by[KI_SCALAR_DIV], tiles.seg, st);
    if (pk) prof->end(PROF_KERNEL0 + KI_SCALAR_DIV);
    // three_band_solver_ver (mo_column_solvers.F90:21-351): the surface part, then the atmospheric
    // columns; w_free_slip is done in k_scalar_col
    if (c.C.lsoil || c.C.lsens) LG_LAUNCH(KI_DIFF_DUAL, k_surface, nx, ny, 1, c, old, nw, rk);
    if (prof) prof->begin(0);
    launch_solve_ver(c, old, nw, rk, tiles.ver_passes, st);
    if (prof) prof->end(0);
    if (c.G.nc > 1) {
      if (prof) prof->begin(1);
      const dim3 g_gen((unsigned)((col_waves(c.G) + GEN_GROUPS - 1) / GEN_GROUPS));
      hipLaunchKernelGGL(k_solve_gen<R>, g_gen, dim3(BCOL, NTR * GEN_GROUPS), 0, st, c, old, nw, rk);
      if (prof) prof->end(1);
    }
  }
#undef LG_LAUNCH
#undef LG_LAUNCH_ZFAST
  return 0;
}

template int enqueue_pchevi<double>(const CtxT<double> &, const StateDev &, const StateDev &, int, hipStream_t, ProfHook *,
                                     const RadCall *, const Tiles &);
template int enqueue_pchevi<float>(const CtxT<float> &, const StateDev &, const StateDev &, int, hipStream_t, ProfHook *,
                                    const RadCall *, const Tiles &);

}  // namespace lgame

namespace {

// Records a pair of events around each timed kernel group on the handle's stream.
struct EventProf : ProfHook {
  lgame_handle *h;
  int open_idx[PROF_GROUPS];
  bool failed = false;
  explicit EventProf(lgame_handle *hh, bool per_kernel = false) : h(hh) {
    kernels = per_kernel;
    for (int &i : open_idx) i = -1;
  }
  int next_event() {
    if (h->ev_next == h->ev_pool.size()) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) {
        failed = true;
        return -1;
      }
      h->ev_pool.push_back(e);
    }
    return (int)h->ev_next++;
  }
  void begin(int group) override {
    const int i = next_event();
    if (i < 0) return;
    if (hipEventRecord(h->ev_pool[i], h->stream) != hipSuccess) failed = true;
    open_idx[group] = i;
  }
  void end(int group) override {
    const int i = next_event();
    if (i < 0 || open_idx[group] < 0) return;
    if (hipEventRecord(h->ev_pool[i], h->stream) != hipSuccess) failed = true;
    h->ev_marks.push_back({group, {open_idx[group], i}});
    open_idx[group] = -1;
  }
};

}  // namespace

namespace {

// column arrays of the radiation data block (one allocation of 6 ncol doubles)
RadDev rad_dev(lgame_handle *h) {
  const size_t ncol = (size_t)h->dims.ny * h->dims.nx;
  RadDev r{};
  r.lat = h->rad_mem;
  r.lon = h->rad_mem + ncol;
  r.albedo = h->rad_mem + 2 * ncol;
  r.tsfc = h->rad_mem + 3 * ncol;
  r.sw_in = h->rad_mem + 4 * ncol;
  r.lw_out = h->rad_mem + 5 * ncol;
  r.x = h->rad_x;
  r.y = h->rad_y;
  r.ypred = h->rad_y;
  r.counts = h->rad_counts;
  r.sq = h->rad_sq;
  r.list = h->rad_list;
  r.list_count = h->rad_list_count;
  r.tol = h->rad_tol;
  r.audit_thresh = h->rad_thresh;
  if (h->mlp) mlp_out_norm(h->mlp, &r.y_mean, &r.y_scale);
  return r;
}

int rad_alloc(lgame_handle *h) {
  if (h->rad_mem) return 0;
  const size_t ncol = (size_t)h->dims.ny * h->dims.nx, nl = h->dims.n_layers;
  HIPCHK(h, hipMalloc(&h->rad_mem, 6 * ncol * sizeof(double)));
  HIPCHK(h, hipMemset(h->rad_mem, 0, 6 * ncol * sizeof(double)));
  // columns of the harvest (5 nl + 4 inputs, nl + 2 outputs) or of the network (its own sizes)
  const size_t nin = 5 * nl + 4 > LGAME_RAD_N_IN ? 5 * nl + 4 : LGAME_RAD_N_IN;
  const size_t nout = nl + 2 > LGAME_RAD_N_OUT ? nl + 2 : LGAME_RAD_N_OUT;
  HIPCHK(h, hipMalloc(&h->rad_x, ncol * nin * sizeof(float)));
  HIPCHK(h, hipMalloc(&h->rad_y, ncol * nout * sizeof(float)));
  HIPCHK(h, hipMalloc(&h->rad_list, ncol * sizeof(int)));
  HIPCHK(h, hipMalloc(&h->rad_list_count, sizeof(unsigned)));
  HIPCHK(h, hipMalloc(&h->rad_counts, 2 * sizeof(unsigned long long)));
  HIPCHK(h, hipMemset(h->rad_counts, 0, 2 * sizeof(unsigned long long)));
  HIPCHK(h, hipMalloc(&h->rad_sq, sizeof(double)));
  HIPCHK(h, hipMemset(h->rad_sq, 0, sizeof(double)));
  if (h->c.C.lsoil && h->cfg.lrad) {  // the soil reads the surface fluxes of the radiation updates
    h->c.C.rad_sw_in = h->cf.C.rad_sw_in = h->rad_mem + 4 * ncol;
    h->c.C.rad_lw_out = h->cf.C.rad_lw_out = h->rad_mem + 5 * ncol;
  }
  return 0;
}

// (ny,nx) Fortran-order host array <-> device column order (ji-1)*nx + (jk-1)
void to_device_cols(const double *host, std::vector<double> &out, int ny, int nx) {
  out.resize((size_t)ny * nx);
  for (int jk = 0; jk < nx; ++jk)
    for (int ji = 0; ji < ny; ++ji) out[(size_t)ji * nx + jk] = host[ji + (size_t)ny * jk];
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------------------------
extern "C" {

int lgame_profile(lgame_handle *h, int enable) {
  if (!h) return fail(h, LGAME_EINVAL, "null handle");
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->prof = enable == 2 ? 2 : (enable != 0 ? 1 : 0);
  h->ev_marks.clear();
  h->ev_next = 0;
  return LGAME_OK;
}

int lgame_profile_read(lgame_handle *h, int group, double *total_ms, long long *count) {
  if (!h || !total_ms || !count) return fail(h, LGAME_EINVAL, "null argument");
  HIPCHK(h, hipStreamSynchronize(h->stream));
  double tot = 0.0;
  long long n = 0;
  for (auto &m : h->ev_marks) {
    if (m.first != group) continue;
    float ms = 0.0f;
    HIPCHK(h, hipEventElapsedTime(&ms, h->ev_pool[m.second.first], h->ev_pool[m.second.second]));
    tot += ms;
    ++n;
  }
  *total_ms = tot;
  *count = n;
  return LGAME_OK;
}

int lgame_profile_reset(lgame_handle *h) {
  if (!h) return fail(h, LGAME_EINVAL, "null handle");
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->ev_marks.clear();
  h->ev_next = 0;
  return LGAME_OK;
}

int lgame_abi_version(void) { return LGAME_ABI_VERSION; }

int lgame_grid_setup(const lgame_dims *dims, const lgame_grid_params *params, lgame_grid_out *out) {
  g_init_error.clear();
  if (!dims || !params || !out) return fail(nullptr, LGAME_EINVAL, "null argument");
  std::string err;
  const int rc = grid_generate(*dims, *params, *out, err);
  if (rc) return fail(nullptr, rc, "%s", err.c_str());
  return LGAME_OK;
}

const char *lgame_last_error(const lgame_handle *h) { return h ? h->err.c_str() : g_init_error.c_str(); }

int lgame_init(const lgame_dims *dims, const lgame_cfg *cfg, const lgame_grid_view *grid, lgame_handle **out) {
  g_init_error.clear();
  if (!dims || !cfg || !grid || !out) return fail(nullptr, LGAME_EINVAL, "null argument");
  *out = nullptr;
  const lgame_dims &d = *dims;
  const lgame_cfg &f = *cfg;
  if (d.ny < 3 || d.nx < 3 || d.n_layers < 3) return fail(nullptr, LGAME_EINVAL, "ny, nx and n_layers must be >= 3");
  if (d.n_oro_layers < 0 || d.n_oro_layers >= d.n_layers)
    return fail(nullptr, LGAME_EINVAL, "it must be 0 <= n_oro_layers < n_layers (mo_run_nml.F90:133-136)");
  if (f.theta_adv_order != 2 && f.theta_adv_order != 3) return fail(nullptr, LGAME_EINVAL, "theta_adv_order must be 2 or 3");
  const bool moist = f.lmoist != 0;
  if (moist ? (d.n_condensed_constituents != 5 || d.n_constituents != 7) : (d.n_condensed_constituents != 0 || d.n_constituents != 1))
    return fail(nullptr, LGAME_EINVAL, "constituent counts must be 1/0 (dry) or 7/5 (moist)");
  if (!f.lperiodic) return fail(nullptr, LGAME_EINVAL, "unsupported: lperiodic=.false. (lateral boundary relaxation)");
  if (!f.luse_bg_state) return fail(nullptr, LGAME_EINVAL, "unsupported: luse_bg_state=.false.");
  if ((f.ltemp_diff_h || f.lmass_diff_h) && !f.lmom_diff_h)
    return fail(nullptr, LGAME_EINVAL,
                "unsupported: horizontal temperature or mass diffusion without horizontal momentum diffusion "
                "(scalar_diff_coeffs, mo_eff_diff_coeffs.F90:540-590, then reads stale vorticity and divergence)");
  if (f.diff_coeff_scheme_h != 0 && f.diff_coeff_scheme_h != 1)
    return fail(nullptr, LGAME_EINVAL, "diff_coeff_scheme_h must be 0 (smag) or 1 (tke)");
  if (f.lpbl && !f.lmom_diff_h)
    return fail(nullptr, LGAME_EINVAL, "lpbl needs lmom_diff_h (the reference stops otherwise, mo_surface_nml.F90)");
  if (f.lsfc_sensible_heat_flux && !f.lprog_soil_temp)
    return fail(nullptr, LGAME_EINVAL,
                "unsupported: the surface sensible heat flux without the prognostic soil (the scalar flux resistance it "
                "divides by is computed only with the soil, mo_pbl.F90:96-104)");
  if (f.lprog_soil_temp && (d.nsoillays < 2 || d.nsoillays > MAX_SOIL || d.nsoillays > d.n_layers + 1))
    return fail(nullptr, LGAME_EINVAL, "with the prognostic soil, nsoillays must be in 2..%d and at most n_layers+1", MAX_SOIL);
  if (f.lprog_soil_temp && d.n_layers < 4) return fail(nullptr, LGAME_EINVAL, "the prognostic soil needs n_layers >= 4");
  // the surface turbulence quantities (update_sfc_turb_quantities, mo_manage_pchevi.F90:52-54)
  const bool sfc_turb = f.lpbl || f.lsfc_sensible_heat_flux || f.lsfc_phase_trans;
  // the diffusion work arrays (temperature and mass diffusion need lmom_diff_h; the PBL friction
  // adds to the momentum diffusion tendencies; the surface fields live there too)
  const bool diffusion = f.lmom_diff_h || f.lmom_diff_v || sfc_turb;
  if (f.lmom_diff_h && !(grid->mean_velocity_area > 0.0)) return fail(nullptr, LGAME_EINVAL, "mean_velocity_area must be positive");
  if (sfc_turb && !(f.karman > 0.0 && f.h_prandtl > 0.0)) return fail(nullptr, LGAME_EINVAL, "karman and h_prandtl must be positive");
  if (f.rad_mode != LGAME_RAD_OFF && f.rad_mode != LGAME_RAD_TEACHER && f.rad_mode != LGAME_RAD_SURROGATE)
    return fail(nullptr, LGAME_EINVAL, "unknown radiation mode %d", f.rad_mode);
  if (f.lrad && f.rad_mode == LGAME_RAD_OFF) return fail(nullptr, LGAME_EINVAL, "lrad needs a radiation mode (LGAME_RAD_*)");
  if (f.lrad && !(f.dtime_rad > 0.0)) return fail(nullptr, LGAME_EINVAL, "dtime_rad must be positive");
  if (f.precision != LGAME_PREC_FP64 && f.precision != LGAME_PREC_MIXED)
    return fail(nullptr, LGAME_EINVAL, "unknown precision mode %d", f.precision);
  if (!(f.dtime > 0.0) || !(f.toa > 0.0)) return fail(nullptr, LGAME_EINVAL, "dtime and toa must be positive");
#define NEED(p) \
  if (!grid->p) return fail(nullptr, LGAME_EINVAL, "grid field " #p " is null");
  NEED(z_scalar) NEED(z_u) NEED(z_v) NEED(z_w) NEED(dx) NEED(dy) NEED(dz) NEED(area_x) NEED(area_y) NEED(area_z) NEED(volume)
  NEED(slope_x) NEED(slope_y) NEED(inner_product_weights) NEED(area_dual_x) NEED(area_dual_y) NEED(area_dual_z) NEED(z_area_dual_z)
  NEED(fvec_x) NEED(fvec_y) NEED(fvec_z) NEED(trsk_weights_u) NEED(trsk_weights_v) NEED(exner_bg_grad_u) NEED(exner_bg_grad_v)
  NEED(exner_bg_grad_w) NEED(theta_v_bg) NEED(exner_bg) NEED(gravity_m_v) NEED(gravity_potential)
  if (f.lmom_diff_h) {
    NEED(dx_dual) NEED(dy_dual) NEED(layer_thickness)
  }
  if (sfc_turb || f.lprog_soil_temp) {
    NEED(land_fraction) NEED(roughness_length)
  }
  if (f.lprog_soil_temp) {
    NEED(lake_fraction) NEED(sfc_rho_c) NEED(t_conduc_soil) NEED(t_const_soil) NEED(sst) NEED(z_soil_interface) NEED(z_soil_center)
  }
#undef NEED

  lgame_handle *h = new lgame_handle();
  h->dims = d;
  h->cfg = f;
  Geom &G = h->c.G;
  G.ny = d.ny;
  G.nx = d.nx;
  G.nl = d.n_layers;
  G.nlev = d.n_layers + 1;
  G.nflat = d.n_layers - d.n_oro_layers;
  G.nc = d.n_constituents;
  G.ncond = d.n_condensed_constituents;
  G.pitch = ((d.nx + 1) + 15) / 16 * 16;
  G.lvl = (long long)(d.ny + 1) * G.pitch;
  G.fsize = (long long)G.nlev * G.lvl;
  if ((unsigned long long)G.fsize * sizeof(double) >= (1ull << 32)) {
    delete h;
    return fail(nullptr, LGAME_EINVAL, "one field of %lld points exceeds 4 GiB", G.fsize);
  }
  h->tiles.seg = default_segments(G);
  h->tiles.ver_passes = col_waves(G) < VER_SPLIT_WAVES && G.nl >= 3 ? 3 : 1;
  Cfg &C = h->c.C;
  C.dtime = f.dtime;
  C.toa = f.toa;
  C.dxn = f.dx;
  C.dyn = f.dy;
  C.klemp_damp_max = f.klemp_damp_max;
  C.klemp_begin_rel = f.klemp_begin_rel;
  C.order = f.theta_adv_order;
  C.llinear = f.llinear != 0;
  C.lcorio = f.lcorio != 0;
  C.lmoist = moist;
  C.lklemp = f.lklemp != 0;
  C.lfreeslip = f.lfreeslip != 0;
  C.lmom_h = f.lmom_diff_h != 0;
  C.lmom_v = f.lmom_diff_v != 0;
  C.ltemp_h = f.ltemp_diff_h != 0;
  C.ltemp_v = f.ltemp_diff_v != 0;
  C.lmass_h = f.lmass_diff_h != 0;
  C.lmass_v = f.lmass_diff_v != 0;
  C.scheme_tke = f.diff_coeff_scheme_h == 1;
  C.c_s = f.c_s;
  C.bg_shear = f.bg_shear;
  C.eff_hor_res = std::sqrt(f.dx * f.dy);  // mo_run_nml.F90:93
  C.mva = grid->mean_velocity_area;
  C.lpbl = f.lpbl != 0;
  C.h_prandtl = f.h_prandtl;
  C.karman = f.karman;
  C.lsoil = f.lprog_soil_temp != 0;
  C.lsens = f.lsfc_sensible_heat_flux != 0;
  C.lsfc_turb = sfc_turb;
  C.nsoil = C.lsoil ? d.nsoillays : 0;

  // n_damping_levels, as mo_grid_generator.F90:1050-1058 derives it from z_w
  {
    const int ny = d.ny, nx = d.nx, nlev = G.nlev;
    C.n_damping_levels = 0;
    for (int jl = 1; jl <= nlev; ++jl) {
      double mx = -1e300;
      for (size_t q = 0; q < (size_t)ny * nx; ++q) mx = std::fmax(mx, grid->z_w[q + (size_t)(jl - 1) * ny * nx]);
      if (mx > f.klemp_begin_rel * f.toa) C.n_damping_levels += 1;
    }
  }

  hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    fail(nullptr, LGAME_EHIP, "hipStreamCreate: %s", hipGetErrorString(e));
    delete h;
    return LGAME_EHIP;
  }
  // field budget: 2 states (FP64), the FP64 grid, column scratch (FP64), the work arrays of the
  // stencil kernels (FP64 plain, FP32 mixed) and, for the mixed form, the FP32 grid copy
  const int nc = d.n_constituents;
  const bool mixed = f.precision == LGAME_PREC_MIXED;
  h->mixed = mixed;
  constexpr int NGRID = 40;
  const size_t fs = (size_t)G.fsize;
  const size_t nstate = (size_t)2 * (nc + 6 + (C.lsoil ? 1 : 0));
  const size_t ndiff = diffusion ? (size_t)14 + nc + 3 * (nc + 1) : 0;  // the diffusion and surface fields of WorkT
  const size_t nwork = (size_t)nc + 4 + 8 + 6 + 4 + 7 + 10 + 2 + ndiff;
  const size_t ncol = 13;  // 12 column-scratch fields and the cloud-water sink velocity
  const size_t wbytes = mixed ? sizeof(float) : sizeof(double);
  h->slab_bytes = (nstate + NGRID + ncol) * fs * sizeof(double) + nwork * fs * wbytes + (mixed ? NGRID * fs * sizeof(float) : 0);
  e = hipMalloc(&h->slab, h->slab_bytes);
  if (e != hipSuccess) {
    fail(nullptr, LGAME_ENOMEM, "hipMalloc of %zu bytes: %s", h->slab_bytes, hipGetErrorString(e));
    (void)hipStreamDestroy(h->stream);
    delete h;
    return LGAME_ENOMEM;
  }
  size_t used = 0;  // bytes handed out; every field starts on a 128-byte boundary (pitch is 16 doubles)
  auto take8 = [&]() {
    double *p = reinterpret_cast<double *>(h->slab + used);
    used += fs * sizeof(double);
    return p;
  };
  auto take4 = [&]() {
    float *p = reinterpret_cast<float *>(h->slab + used);
    used += fs * sizeof(float);
    return p;
  };
  for (int s = 0; s < 2; ++s) {
    StateDev &S = h->slot[s];
    for (int jc = 0; jc < MAX_CONST; ++jc) S.rho[jc] = jc < nc ? take8() : nullptr;
    S.rhotheta_v = take8();
    S.theta_v_pert = take8();
    S.exner_pert = take8();
    S.wind_u = take8();
    S.wind_v = take8();
    S.wind_w = take8();
    S.temperature_soil = C.lsoil ? take8() : nullptr;
  }
  h->state_bytes = used;
  // grid: NGRID consecutive FP64 fields in a fixed order (the FP32 copy uses the same order)
  double *grid64 = take8();
  for (int q = 1; q < NGRID; ++q) take8();
  Fld<double> col[14];
  h->work_off = used;
  for (int q = 0; q < (int)ncol; ++q) col[q] = take8();
  auto assign_work = [&](auto &W, auto take) {
    for (int jc = 0; jc < MAX_CONST; ++jc) W.tend_rho[jc] = jc < nc ? take() : nullptr;
    W.tend_rhotheta_v = take();
    W.tend_u = take();
    W.tend_v = take();
    W.tend_w = take();
    W.pnl_u = take();
    W.pnl_v = take();
    W.pnl_w = take();
    W.pl_u = take();
    W.pl_v = take();
    W.pl_w = take();
    W.pold_u = take();
    W.pold_v = take();
    W.vsg_x = take();
    W.vsg_y = take();
    W.vsg_z = take();
    W.pvt_x = take();
    W.pvt_y = take();
    W.pvt_z = take();
    W.heating_diss = take();
    W.csh = take();
    W.temperature = take();
    W.radiation_tendency = take();
    W.eta_x = take();
    W.eta_y = take();
    W.eta_z = take();
    W.u_ph = take();
    W.v_ph = take();
    W.w_ph = take();
    W.v_squared = take();
    for (int q = 0; q < 2; ++q) {
      W.tgu[q] = take();
      W.tgv[q] = take();
      W.tfd[q] = take();
      W.tu[q] = take();
      W.tv[q] = take();
    }
    W.decel = take();
    W.pgc = take();
    if (diffusion) {
      W.tke = take();
      W.tke_next = take();
      W.visc = take();
      W.kv = take();
      W.vdiv = take();
      W.zeta = take();
      W.vz = take();
      W.mom_x = take();
      W.mom_y = take();
      W.mom_z = take();
      W.tdh = take();
      W.rsum = take();
      W.sfc = take();
      W.sfc2 = take();
      for (int jc = 0; jc < MAX_CONST; ++jc) W.mdt[jc] = jc < nc ? take() : nullptr;
      for (int q = 0; q <= MAX_CONST; ++q) {
        W.dfu[q] = q <= nc ? take() : nullptr;
        W.dfv[q] = q <= nc ? take() : nullptr;
        W.dfw[q] = q <= nc ? take() : nullptr;
      }
    }
    for (int q = 0; q < 12; ++q) W.col[q] = col[q];
    W.vfall_cw = col[12];
  };
  if (mixed)
    assign_work(h->cf.w, take4);
  else
    assign_work(h->c.w, take8);
  h->work_bytes = used - h->work_off;
  float *grid32 = nullptr;
  if (mixed) {
    grid32 = take4();
    for (int q = 1; q < NGRID; ++q) take4();
  }
  if (used > h->slab_bytes) {
    fail(nullptr, LGAME_EINVAL, "internal: field budget exceeded");
    lgame_finalize(h);
    return LGAME_EINVAL;
  }
  e = hipMemsetAsync(h->slab, 0, h->slab_bytes, h->stream);
  h->staging_elems = (size_t)d.ny * d.nx * d.n_layers * (nc > 6 ? nc : 6) + (size_t)(d.ny + 1) * (d.nx + 1) * (d.n_layers + 1);
  if (e == hipSuccess) e = hipMalloc(&h->staging, h->staging_elems * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&h->red, (size_t)2 * NRED * RED_GRID * sizeof(double) + NRED * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&h->counter, 2 * sizeof(unsigned long long));
  if (e == hipSuccess) e = hipMemsetAsync(h->counter, 0, 2 * sizeof(unsigned long long), h->stream);
  h->c.C.nonfinite = h->counter + 1;
  if (e == hipSuccess) e = hipMalloc(&h->trsk, (size_t)(3 * d.ny + 2 * (d.ny + 1)) * (sizeof(double) + sizeof(float)));
  if (e != hipSuccess) {
    fail(nullptr, LGAME_ENOMEM, "device allocation: %s", hipGetErrorString(e));
    lgame_finalize(h);
    return LGAME_ENOMEM;
  }
  const int ny = d.ny, nx = d.nx, nl = d.n_layers, nlev = nl + 1;
  // grid field q of the block
  enum {
    G_Z_SCALAR, G_Z_U, G_Z_V, G_Z_W, G_DX, G_DY, G_DZ, G_AREA_X, G_AREA_Y, G_AREA_Z, G_VOLUME, G_SLOPE_X, G_SLOPE_Y,
    G_IPW, G_AREA_DUAL_X = G_IPW + 6, G_AREA_DUAL_Y, G_AREA_DUAL_Z, G_Z_AREA_DUAL_Z, G_FVEC_X, G_FVEC_Y, G_FVEC_Z,
    G_EBG_U, G_EBG_V, G_EBG_W, G_THETA_V_BG, G_EXNER_BG, G_GRAVITY_M_V, G_GRAVITY_POTENTIAL, G_KLEMP_PREF, G_DX_DUAL,
    G_DY_DUAL, G_LAYER_THICKNESS, G_LAND_FRACTION, G_ROUGHNESS_LENGTH, G_SOIL, G_COUNT
  };
  static_assert(G_COUNT <= NGRID, "grid field budget");
  auto gf = [&](int q) { return grid64 + (size_t)q * fs; };
  int rc = 0;
  rc = rc ? rc : upload(h, gf(G_Z_SCALAR), grid->z_scalar, ny, nx, nl);
  rc = rc ? rc : upload(h, gf(G_Z_U), grid->z_u, ny, nx + 1, nl);
  rc = rc ? rc : upload(h, gf(G_Z_V), grid->z_v, ny + 1, nx, nl);
  rc = rc ? rc : upload(h, gf(G_Z_W), grid->z_w, ny, nx, nlev);
  rc = rc ? rc : upload(h, gf(G_DX), grid->dx, ny, nx + 1, nl);
  rc = rc ? rc : upload(h, gf(G_DY), grid->dy, ny + 1, nx, nl);
  rc = rc ? rc : upload(h, gf(G_DZ), grid->dz, ny, nx, nlev);
  rc = rc ? rc : upload(h, gf(G_AREA_X), grid->area_x, ny, nx + 1, nl);
  rc = rc ? rc : upload(h, gf(G_AREA_Y), grid->area_y, ny + 1, nx, nl);
  rc = rc ? rc : upload(h, gf(G_AREA_Z), grid->area_z, ny, nx, nlev);
  rc = rc ? rc : upload(h, gf(G_VOLUME), grid->volume, ny, nx, nl);
  rc = rc ? rc : upload(h, gf(G_SLOPE_X), grid->slope_x, ny, nx + 1, nl);
  rc = rc ? rc : upload(h, gf(G_SLOPE_Y), grid->slope_y, ny + 1, nx, nl);
  rc = rc ? rc : upload(h, gf(G_IPW), grid->inner_product_weights, ny, nx, nl, 6, 6);
  rc = rc ? rc : upload(h, gf(G_AREA_DUAL_X), grid->area_dual_x, ny + 1, nx, nlev);
  rc = rc ? rc : upload(h, gf(G_AREA_DUAL_Y), grid->area_dual_y, ny, nx + 1, nlev);
  rc = rc ? rc : upload(h, gf(G_AREA_DUAL_Z), grid->area_dual_z, ny + 1, nx + 1, nl);
  rc = rc ? rc : upload(h, gf(G_Z_AREA_DUAL_Z), grid->z_area_dual_z, ny + 1, nx + 1, nl);
  rc = rc ? rc : upload(h, gf(G_FVEC_X), grid->fvec_x, ny + 1, nx, 1);
  rc = rc ? rc : upload(h, gf(G_FVEC_Y), grid->fvec_y, ny, nx + 1, 1);
  rc = rc ? rc : upload(h, gf(G_FVEC_Z), grid->fvec_z, ny + 1, nx + 1, 1);
  rc = rc ? rc : upload(h, gf(G_EBG_U), grid->exner_bg_grad_u, ny, nx + 1, nl);
  rc = rc ? rc : upload(h, gf(G_EBG_V), grid->exner_bg_grad_v, ny + 1, nx, nl);
  rc = rc ? rc : upload(h, gf(G_EBG_W), grid->exner_bg_grad_w, ny, nx, nlev);
  rc = rc ? rc : upload(h, gf(G_THETA_V_BG), grid->theta_v_bg, ny, nx, nl);
  rc = rc ? rc : upload(h, gf(G_EXNER_BG), grid->exner_bg, ny, nx, nl);
  rc = rc ? rc : upload(h, gf(G_GRAVITY_M_V), grid->gravity_m_v, ny, nx, nlev);
  rc = rc ? rc : upload(h, gf(G_GRAVITY_POTENTIAL), grid->gravity_potential, ny, nx, nl);
  if (f.lmom_diff_h) {
    rc = rc ? rc : upload(h, gf(G_DX_DUAL), grid->dx_dual, ny + 1, nx, nl);
    rc = rc ? rc : upload(h, gf(G_DY_DUAL), grid->dy_dual, ny, nx + 1, nl);
    rc = rc ? rc : upload(h, gf(G_LAYER_THICKNESS), grid->layer_thickness, ny, nx, nl);
  }
  if (sfc_turb || C.lsoil) {
    rc = rc ? rc : upload(h, gf(G_LAND_FRACTION), grid->land_fraction, ny, nx, 1);
    rc = rc ? rc : upload(h, gf(G_ROUGHNESS_LENGTH), grid->roughness_length, ny, nx, 1);
  }
  if (C.lsoil) {
    const double *soil2d[5] = {grid->lake_fraction, grid->sfc_rho_c, grid->t_conduc_soil, grid->t_const_soil, grid->sst};
    for (int q = 0; q < 5; ++q) rc = rc ? rc : upload(h, gf(G_SOIL) + (size_t)q * G.lvl, soil2d[q], ny, nx, 1);
    const int ns = d.nsoillays;
    std::vector<double> z((size_t)2 * ns + 2);
    std::memcpy(z.data(), grid->z_soil_interface, sizeof(double) * (ns + 1));
    std::memcpy(z.data() + ns + 1, grid->z_soil_center, sizeof(double) * ns);
    z[2 * ns + 1] = grid->z_t_const;
    if (!rc) {
      e = hipMalloc(&h->soil_z, z.size() * sizeof(double));
      if (e == hipSuccess) e = hipMemcpy(h->soil_z, z.data(), z.size() * sizeof(double), hipMemcpyHostToDevice);
      if (e != hipSuccess) rc = fail(h, LGAME_EHIP, "soil grid: %s", hipGetErrorString(e));
    }
  }
  if (!rc) {
    std::vector<double> kp((size_t)ny * nx * nlev);
    for (size_t q = 0; q < kp.size(); ++q) kp[q] = klemp_prefactor(f, grid->z_w[q]);
    rc = upload(h, gf(G_KLEMP_PREF), kp.data(), ny, nx, nlev);
  }
  const size_t ntw = (size_t)3 * ny + 2 * (ny + 1);
  float *trsk32 = reinterpret_cast<float *>(h->trsk + ntw);
  if (!rc) {
    std::vector<double> tw(ntw);
    std::memcpy(tw.data(), grid->trsk_weights_u, sizeof(double) * 3 * ny);
    std::memcpy(tw.data() + 3 * ny, grid->trsk_weights_v, sizeof(double) * 2 * (ny + 1));
    std::vector<float> tw32(tw.begin(), tw.end());
    e = hipMemcpy(h->trsk, tw.data(), tw.size() * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(trsk32, tw32.data(), tw32.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e != hipSuccess) rc = fail(h, LGAME_EHIP, "hipMemcpy: %s", hipGetErrorString(e));
  }
  if (!rc && mixed) {
    hipLaunchKernelGGL(k_to_float, dim3(4096), dim3(256), 0, h->stream, grid32, (const double *)grid64, (size_t)NGRID * fs);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) rc = fail(h, LGAME_EHIP, "grid conversion: %s", hipGetErrorString(e));
  }
  if (rc) {
    g_init_error = h->err;
    lgame_finalize(h);
    return rc;
  }
  auto assign_grid = [&](auto &g, auto *base, auto *trsk) {
    auto at = [&](int q) { return base + (size_t)q * fs; };
    g.z_scalar = at(G_Z_SCALAR);
    g.z_u = at(G_Z_U);
    g.z_v = at(G_Z_V);
    g.z_w = at(G_Z_W);
    g.dx = at(G_DX);
    g.dy = at(G_DY);
    g.dz = at(G_DZ);
    g.area_x = at(G_AREA_X);
    g.area_y = at(G_AREA_Y);
    g.area_z = at(G_AREA_Z);
    g.volume = at(G_VOLUME);
    g.slope_x = at(G_SLOPE_X);
    g.slope_y = at(G_SLOPE_Y);
    for (int q = 0; q < 6; ++q) g.ipw[q] = at(G_IPW + q);
    g.area_dual_x = at(G_AREA_DUAL_X);
    g.area_dual_y = at(G_AREA_DUAL_Y);
    g.area_dual_z = at(G_AREA_DUAL_Z);
    g.z_area_dual_z = at(G_Z_AREA_DUAL_Z);
    g.fvec_x = at(G_FVEC_X);
    g.fvec_y = at(G_FVEC_Y);
    g.fvec_z = at(G_FVEC_Z);
    g.ebg_u = at(G_EBG_U);
    g.ebg_v = at(G_EBG_V);
    g.ebg_w = at(G_EBG_W);
    g.theta_v_bg = at(G_THETA_V_BG);
    g.exner_bg = at(G_EXNER_BG);
    g.gravity_m_v = at(G_GRAVITY_M_V);
    g.gravity_potential = at(G_GRAVITY_POTENTIAL);
    g.klemp_pref = at(G_KLEMP_PREF);
    g.dx_dual = at(G_DX_DUAL);
    g.dy_dual = at(G_DY_DUAL);
    g.layer_thickness = at(G_LAYER_THICKNESS);
    g.land_fraction = at(G_LAND_FRACTION);
    g.roughness_length = at(G_ROUGHNESS_LENGTH);
    g.soil = at(G_SOIL);
    g.soil_z = h->soil_z;
    g.trsk_u = trsk;
    g.trsk_v = trsk + 3 * ny;
  };
  assign_grid(h->c.g, grid64, h->trsk);
  h->c.g64 = h->c.g;
  if (mixed) {
    h->cf.G = h->c.G;
    h->cf.C = h->c.C;
    assign_grid(h->cf.g, grid32, trsk32);
    h->cf.g64 = h->c.g;
  }
  e = init_surface(h);
  if (e != hipSuccess) {
    fail(nullptr, LGAME_EHIP, "surface initialisation: %s", hipGetErrorString(e));
    lgame_finalize(h);
    return LGAME_EHIP;
  }
  *out = h;
  return LGAME_OK;
}

static int state_io(lgame_handle *h, int slot, lgame_state_view *st, bool up) {
  if (!h || !st) return fail(h, LGAME_EINVAL, "null argument");
  if (slot != 0 && slot != 1) return fail(h, LGAME_EINVAL, "slot must be 0 or 1");
  const int ny = h->dims.ny, nx = h->dims.nx, nl = h->dims.n_layers, nc = h->dims.n_constituents;
  StateDev &S = h->slot[slot];
  if (!st->rho || !st->rhotheta_v || !st->theta_v_pert || !st->exner_pert || !st->wind_u || !st->wind_v || !st->wind_w)
    return fail(h, LGAME_EINVAL, "null state field");
  const bool soil = h->c.C.lsoil != 0;
  if (soil && !st->temperature_soil) return fail(h, LGAME_EINVAL, "temperature_soil is null (the prognostic soil is on)");
  int rc = 0;
  const size_t ncell = (size_t)ny * nx * nl;
  for (int jc = 0; jc < nc && !rc; ++jc)
    rc = up ? upload(h, S.rho[jc].p, st->rho + jc * ncell, ny, nx, nl) : download(h, st->rho + jc * ncell, S.rho[jc].p, ny, nx, nl);
  struct F {
    double *host, *dev;
    int n1, n2, n3;
  } fs[] = {{st->rhotheta_v, S.rhotheta_v.p, ny, nx, nl},     {st->theta_v_pert, S.theta_v_pert.p, ny, nx, nl},
            {st->exner_pert, S.exner_pert.p, ny, nx, nl},     {st->wind_u, S.wind_u.p, ny, nx + 1, nl},
            {st->wind_v, S.wind_v.p, ny + 1, nx, nl},         {st->wind_w, S.wind_w.p, ny, nx, nl + 1}};
  for (auto &x : fs) {
    if (rc) break;
    rc = up ? upload(h, x.dev, x.host, x.n1, x.n2, x.n3) : download(h, x.host, x.dev, x.n1, x.n2, x.n3);
  }
  if (!rc && soil)
    rc = up ? upload(h, S.temperature_soil.p, st->temperature_soil, ny, nx, h->dims.nsoillays)
            : download(h, st->temperature_soil, S.temperature_soil.p, ny, nx, h->dims.nsoillays);
  if (!rc && up) h->have_state[slot] = 1;
  return rc;
}

int lgame_linear_combine(lgame_handle *h, int slot_1, int slot_2, double coeff_1, double coeff_2, lgame_state_view *out) {
  if (!h || !out) return fail(h, LGAME_EINVAL, "null argument");
  for (int s : {slot_1, slot_2})
    if ((s != 0 && s != 1) || !h->have_state[s]) return fail(h, LGAME_ESTATE, "slot %d holds no state", s);
  if (!out->rho || !out->rhotheta_v || !out->theta_v_pert || !out->exner_pert || !out->wind_u || !out->wind_v || !out->wind_w)
    return fail(h, LGAME_EINVAL, "null state field");
  const bool soil = h->c.C.lsoil != 0;
  if (soil && !out->temperature_soil) return fail(h, LGAME_EINVAL, "temperature_soil is null (the prognostic soil is on)");
  const int ny = h->dims.ny, nx = h->dims.nx, nl = h->dims.n_layers, nc = h->dims.n_constituents;
  const StateDev &A = h->slot[slot_1], &B = h->slot[slot_2];
  const double *tb = h->c.g64.theta_v_bg.p;
  const int g = h->c.G.ncond;
  struct F {
    double *host;
    const double *a, *b;
    int n1, n2, n3;
    bool theta;
  };
  std::vector<F> fs;
  const size_t ncell = (size_t)ny * nx * nl;
  for (int jc = 0; jc < nc; ++jc) fs.push_back({out->rho + jc * ncell, A.rho[jc].p, B.rho[jc].p, ny, nx, nl, false});
  fs.push_back({out->rhotheta_v, A.rhotheta_v.p, B.rhotheta_v.p, ny, nx, nl, false});
  fs.push_back({out->theta_v_pert, A.rhotheta_v.p, B.rhotheta_v.p, ny, nx, nl, true});
  fs.push_back({out->exner_pert, A.exner_pert.p, B.exner_pert.p, ny, nx, nl, false});
  fs.push_back({out->wind_u, A.wind_u.p, B.wind_u.p, ny, nx + 1, nl, false});
  fs.push_back({out->wind_v, A.wind_v.p, B.wind_v.p, ny + 1, nx, nl, false});
  fs.push_back({out->wind_w, A.wind_w.p, B.wind_w.p, ny, nx, nl + 1, false});
  if (soil) fs.push_back({out->temperature_soil, A.temperature_soil.p, B.temperature_soil.p, ny, nx, h->dims.nsoillays, false});
  for (const F &f : fs) {
    const size_t n = (size_t)f.n1 * f.n2 * f.n3;
    if (n > h->staging_elems) return fail(h, LGAME_EINVAL, "internal: staging buffer too small");
    const dim3 blk(BX, BY, 1), grd((f.n2 + BX - 1) / BX, (f.n1 + BY - 1) / BY, f.n3);
    hipLaunchKernelGGL(k_combine_to_ref, grd, blk, 0, h->stream, h->staging, f.a, f.b, f.theta ? (const double *)A.rho[g].p : nullptr,
                       f.theta ? (const double *)B.rho[g].p : nullptr, f.theta ? tb : nullptr, coeff_1, coeff_2, f.n1, f.n2, f.n3,
                       h->c.G);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipMemcpyAsync(f.host, h->staging, n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
  }
  return LGAME_OK;
}

int lgame_set_state(lgame_handle *h, int slot, const lgame_state_view *state) {
  const int rc = state_io(h, slot, const_cast<lgame_state_view *>(state), true);
  if (!rc) HIPCHK(h, hipMemsetAsync(h->counter + 1, 0, sizeof(unsigned long long), h->stream));  // a new state
  return rc;
}

int lgame_get_state(lgame_handle *h, int slot, lgame_state_view *state) {
  if (h && (slot == 0 || slot == 1) && !h->have_state[slot]) return fail(h, LGAME_ESTATE, "slot %d holds no state", slot);
  return state_io(h, slot, state, false);
}

int lgame_pchevi(lgame_handle *h, int slot_old, int slot_new, int total_step_counter, int lrad_update, double t_0) {
  if (!h) return fail(h, LGAME_EINVAL, "null handle");
  if ((slot_old != 0 && slot_old != 1) || (slot_new != 0 && slot_new != 1) || slot_old == slot_new)
    return fail(h, LGAME_EINVAL, "slots must be 0 and 1, distinct");
  if (!h->have_state[slot_old]) return fail(h, LGAME_ESTATE, "slot %d holds no state", slot_old);
  if (!h->have_state[slot_new]) return fail(h, LGAME_ESTATE, "slot %d holds no state (copy state_old there first)", slot_new);
  if (total_step_counter < 0) return fail(h, LGAME_EINVAL, "negative step counter");
  RadCall rc{};
  const RadCall *rad = nullptr;
  if (lrad_update && h->cfg.lrad) {
    if (!h->rad_surface) return fail(h, LGAME_ESTATE, "radiation update before lgame_rad_set_surface");
    if (h->cfg.rad_mode == LGAME_RAD_SURROGATE && !h->mlp)
      return fail(h, LGAME_ESTATE, "radiation surrogate update before lgame_rad_attach_mlp");
    rc.mode = h->cfg.rad_mode;
    rc.t_rad = t_0 + 0.5 * h->cfg.dtime_rad;  // mo_manage_radiation_calls.F90:53
    rc.dev = rad_dev(h);
    rc.dev.call = h->rad_calls;
    rc.mlp = h->mlp;
    rc.work = h->mlp_work;
    rad = &rc;
    ++h->rad_calls;
  }
  EventProf ep(h, h->prof == 2);
  ProfHook *ph = h->prof ? &ep : nullptr;
  if (ph) ep.begin(LGAME_PROF_STEP);
  int erc;
  if (h->mixed)
    erc = enqueue_pchevi(h->cf, h->slot[slot_old], h->slot[slot_new], total_step_counter, h->stream, ph, rad, h->tiles);
  else
    erc = enqueue_pchevi(h->c, h->slot[slot_old], h->slot[slot_new], total_step_counter, h->stream, ph, rad, h->tiles);
  if (erc) return fail(h, erc, "the radiation update could not be enqueued (%s)", hipGetErrorString(hipGetLastError()));
  if (ph) {
    ep.end(LGAME_PROF_STEP);
    if (ep.failed) return fail(h, LGAME_EHIP, "event recording failed");
  }
  HIPCHK(h, hipGetLastError());
  h->stepped = 1;
  return LGAME_OK;
}

static int nonfinite_check(lgame_handle *h);

int lgame_pchevi_host(lgame_handle *h, const lgame_state_view *state_old, lgame_state_view *state_new, int total_step_counter,
                      int lrad_update, double t_0) {
  const int a = total_step_counter % 2 == 0 ? 0 : 1, b = 1 - a;
  int rc = lgame_set_state(h, a, state_old);
  if (!rc) rc = lgame_set_state(h, b, state_new);
  if (!rc) rc = lgame_pchevi(h, a, b, total_step_counter, lrad_update, t_0);
  if (!rc) rc = lgame_get_state(h, b, state_new);
  if (!rc) rc = nonfinite_check(h);
  return rc;
}

// The step's own check (mo_write_out.F90:63-67 checks the state for NaN at every output): the
// corrector's column solvers flag every new state value that is not finite.
static int nonfinite_check(lgame_handle *h) {
  unsigned long long flag = 0;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  HIPCHK(h, hipMemcpy(&flag, h->counter + 1, sizeof flag, hipMemcpyDeviceToHost));
  if (flag) return fail(h, LGAME_ENONFINITE, "a step produced a non-finite value in the state (the model crashed)");
  return LGAME_OK;
}

int lgame_sync(lgame_handle *h) {
  if (!h) return fail(h, LGAME_EINVAL, "null handle");
  return nonfinite_check(h);
}

int lgame_integrals(lgame_handle *h, int slot, double *out) {
  if (!h || !out) return fail(h, LGAME_EINVAL, "null argument");
  if ((slot != 0 && slot != 1) || !h->have_state[slot]) return fail(h, LGAME_ESTATE, "slot %d holds no state", slot);
  const int nq = 4 + h->dims.n_constituents;
  double *partial = h->red, *res = h->red + 2 * NRED * RED_GRID;
  if (!h->stepped) {  // before the first step the temperature is diagnosed from this state (control.F90:484)
    const Geom &G = h->c.G;
    if (h->mixed)
      hipLaunchKernelGGL(k_temperature<float>, grid3(G.nx, G.ny, G.nl), dim3(BX, BY, 1), 0, h->stream, h->cf, h->slot[slot]);
    else
      hipLaunchKernelGGL(k_temperature<double>, grid3(G.nx, G.ny, G.nl), dim3(BX, BY, 1), 0, h->stream, h->c, h->slot[slot]);
  }
  if (h->mixed)
    hipLaunchKernelGGL(k_integrals<float>, dim3(RED_GRID), dim3(RED_BLOCK), 0, h->stream, h->c, h->slot[slot],
                       Fld<const float>(h->cf.w.temperature), partial);
  else
    hipLaunchKernelGGL(k_integrals<double>, dim3(RED_GRID), dim3(RED_BLOCK), 0, h->stream, h->c, h->slot[slot],
                       Fld<const double>(h->c.w.temperature), partial);
  hipLaunchKernelGGL(k_integrals_final, dim3(1), dim3(RED_BLOCK), 0, h->stream, (const double *)partial, nq, res);
  HIPCHK(h, hipGetLastError());
  double tmp[NRED];
  HIPCHK(h, hipMemcpyAsync(tmp, res, nq * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  tmp[LGAME_INT_KINETIC] = 0.5 * tmp[LGAME_INT_KINETIC];
  std::memcpy(out, tmp, nq * sizeof(double));
  return LGAME_OK;
}

// ---- checkpoint of the carried fields ------------------------------------------------------
}  // extern "C"

namespace {

// The device fields a step carries to the next besides the two state slots, each with the bytes
// it occupies: the pressure-gradient terms whose negated sum becomes the next step's old
// gradient (mo_p_grad.F90:38-43), the lagged momentum advection with the mass fluxes and the
// kinetic energy it was formed from (mo_vector_tend_expl.F90:58; the kinetic energy of the lowest
// layer also drives the next step's surface turbulence, mo_pbl.F90:19-109), the sediment
// heating, the dissipative heating (the
// next step's TKE production, mo_tke.F90:19-67), the radiation tendency, and with diffusion or
// the surface physics the TKE, viscosity, momentum-diffusion tendencies, heating and
// mass-diffusion tendencies and the surface turbulence quantities; then the radiation's surface
// fluxes.
struct Carried {
  void *p;
  size_t bytes;
};

std::vector<Carried> carried_fields(lgame_handle *h) {
  std::vector<Carried> v;
  const size_t es = h->mixed ? sizeof(float) : sizeof(double);
  const size_t box = (size_t)h->c.G.fsize * es, lvl = (size_t)h->c.G.lvl * es;
  auto add_work = [&](const auto &w) {
    for (auto f : {w.pnl_u, w.pnl_v, w.pnl_w, w.pl_u, w.pl_v, w.pl_w, w.pold_u, w.pold_v, w.vsg_x, w.vsg_y, w.vsg_z, w.pvt_x,
                   w.pvt_y, w.pvt_z, w.u_ph, w.v_ph, w.w_ph, w.v_squared, w.csh, w.heating_diss, w.radiation_tendency, w.tke,
                   w.visc, w.mom_x, w.mom_y, w.mom_z, w.tdh})
      if (f.p) v.push_back({(void *)f.p, box});
    for (int jc = 0; jc < h->dims.n_constituents; ++jc)
      if (w.mdt[jc].p) v.push_back({(void *)w.mdt[jc].p, box});
    for (auto f : {w.sfc, w.sfc2})
      if (f.p) v.push_back({(void *)f.p, 3 * lvl});
  };
  if (h->mixed)
    add_work(h->cf.w);
  else
    add_work(h->c.w);
  if (h->rad_mem) {
    const size_t ncol = (size_t)h->dims.ny * h->dims.nx;
    v.push_back({(void *)(h->rad_mem + 4 * ncol), 2 * ncol * sizeof(double)});
  }
  return v;
}

// header of a checkpoint: what it was made from, checked on load
struct CkptHeader {
  char magic[8];
  int abi, precision, ny, nx, nl, nc, nfields, has_rad;
  long long payload;
  unsigned long long rad_calls;
};

long long checkpoint_bytes(lgame_handle *h) {
  long long n = (long long)sizeof(CkptHeader);
  for (auto &f : carried_fields(h)) n += (long long)f.bytes;
  return n;
}

CkptHeader checkpoint_header(lgame_handle *h) {
  CkptHeader hd{};
  std::memcpy(hd.magic, "LGCKPT1", 8);
  hd.abi = LGAME_ABI_VERSION;
  hd.precision = h->mixed ? 1 : 0;
  hd.ny = h->dims.ny;
  hd.nx = h->dims.nx;
  hd.nl = h->dims.n_layers;
  hd.nc = h->dims.n_constituents;
  auto f = carried_fields(h);
  hd.nfields = (int)f.size();
  hd.has_rad = h->rad_mem ? 1 : 0;
  hd.payload = checkpoint_bytes(h) - (long long)sizeof(CkptHeader);
  hd.rad_calls = h->rad_calls;
  return hd;
}

}  // namespace

extern "C" {

int lgame_checkpoint_size(lgame_handle *h, long long *bytes) {
  if (!h || !bytes) return fail(h, LGAME_EINVAL, "null argument");
  *bytes = checkpoint_bytes(h);
  return LGAME_OK;
}

int lgame_checkpoint_save(lgame_handle *h, void *buf, long long bytes) {
  if (!h || !buf) return fail(h, LGAME_EINVAL, "null argument");
  if (bytes < checkpoint_bytes(h)) return fail(h, LGAME_EINVAL, "the buffer holds %lld bytes, the checkpoint needs %lld", bytes,
                                               checkpoint_bytes(h));
  const CkptHeader hd = checkpoint_header(h);
  char *out = (char *)buf;
  std::memcpy(out, &hd, sizeof hd);
  out += sizeof hd;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  for (auto &f : carried_fields(h)) {
    HIPCHK(h, hipMemcpy(out, f.p, f.bytes, hipMemcpyDeviceToHost));
    out += f.bytes;
  }
  return LGAME_OK;
}

int lgame_checkpoint_load(lgame_handle *h, const void *buf, long long bytes) {
  if (!h || !buf) return fail(h, LGAME_EINVAL, "null argument");
  if (bytes < (long long)sizeof(CkptHeader)) return fail(h, LGAME_EINVAL, "not a checkpoint (too short)");
  CkptHeader hd;
  std::memcpy(&hd, buf, sizeof hd);
  const CkptHeader want = checkpoint_header(h);
  if (std::memcmp(hd.magic, want.magic, 8) != 0 || hd.abi != want.abi) return fail(h, LGAME_EINVAL, "not a checkpoint of this library");
  if (hd.precision != want.precision || hd.ny != want.ny || hd.nx != want.nx || hd.nl != want.nl || hd.nc != want.nc ||
      hd.nfields != want.nfields || hd.has_rad != want.has_rad || hd.payload != want.payload)
    return fail(h, LGAME_EINVAL, "the checkpoint was made by a handle of another configuration");
  if (bytes < checkpoint_bytes(h)) return fail(h, LGAME_EINVAL, "the checkpoint is truncated");
  const char *in = (const char *)buf + sizeof hd;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  for (auto &f : carried_fields(h)) {
    HIPCHK(h, hipMemcpy(f.p, in, f.bytes, hipMemcpyHostToDevice));
    in += f.bytes;
  }
  h->rad_calls = hd.rad_calls;
  return LGAME_OK;
}

int lgame_check_finite(lgame_handle *h, int slot, long long *n_bad) {
  if (!h || !n_bad) return fail(h, LGAME_EINVAL, "null argument");
  if ((slot != 0 && slot != 1) || !h->have_state[slot]) return fail(h, LGAME_ESTATE, "slot %d holds no state", slot);
  HIPCHK(h, hipMemsetAsync(h->counter, 0, sizeof(unsigned long long), h->stream));
  const Geom &G = h->c.G;
  hipLaunchKernelGGL(k_count_nonfinite, grid3(G.nx + 1, G.ny + 1, G.nlev), dim3(BX, BY, 1), 0, h->stream, h->c, h->slot[slot], h->counter);
  HIPCHK(h, hipGetLastError());
  unsigned long long v = 0;
  HIPCHK(h, hipMemcpyAsync(&v, h->counter, sizeof v, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  *n_bad = (long long)v;
  return v ? LGAME_ENONFINITE : LGAME_OK;
}

int lgame_get_diag(lgame_handle *h, int which, double *out) {
  if (!h || !out) return fail(h, LGAME_EINVAL, "null argument");
  const int ny = h->dims.ny, nx = h->dims.nx, nl = h->dims.n_layers;
  const bool diffusion = h->cfg.lmom_diff_h || h->cfg.lmom_diff_v || h->c.C.lsfc_turb;
  auto get = [&](auto &w) -> int {
    switch (which) {
      case LGAME_DIAG_TEMPERATURE: return download(h, out, w.temperature.p, ny, nx, nl);
      case LGAME_DIAG_HEATING_DISS: return download(h, out, w.heating_diss.p, ny, nx, nl);
      default: break;
    }
    if (!diffusion) return fail(h, LGAME_EINVAL, "diagnostic %d exists only with diffusion", which);
    switch (which) {
      case LGAME_DIAG_TKE: return download(h, out, w.tke.p, ny, nx, nl);
      case LGAME_DIAG_VISCOSITY: return download(h, out, w.visc.p, ny, nx, nl);
      case LGAME_DIAG_MOM_DIFF_TEND_X: return download(h, out, w.mom_x.p, ny, nx + 1, nl);
      case LGAME_DIAG_MOM_DIFF_TEND_Z: return download(h, out, w.mom_z.p, ny, nx, nl + 1);
      case LGAME_DIAG_TEMP_DIFF_HEATING: return download(h, out, w.tdh.p, ny, nx, nl);
      case LGAME_DIAG_ROUGHNESS_LENGTH:
      case LGAME_DIAG_MONIN_OBUKHOV:
        if (!h->c.C.lsfc_turb)
          return fail(h, LGAME_EINVAL, "diagnostic %d exists only with the surface turbulence quantities", which);
        return download(h, out, w.sfc.p + (which == LGAME_DIAG_MONIN_OBUKHOV ? 2 * (size_t)h->c.G.lvl : 0), ny, nx, 1);
      case LGAME_DIAG_SCALAR_FLUX_RESISTANCE:
      case LGAME_DIAG_PFD_SENS_SOIL:
      case LGAME_DIAG_PFD_SENS_SEA:
        if (!h->c.C.lsens) return fail(h, LGAME_EINVAL, "diagnostic %d exists only with the sensible heat flux", which);
        return download(h, out, w.sfc2.p + (size_t)(which - LGAME_DIAG_SCALAR_FLUX_RESISTANCE) * (size_t)h->c.G.lvl, ny, nx, 1);
      default: return fail(h, LGAME_EINVAL, "unknown diagnostic %d", which);
    }
  };
  return h->mixed ? get(h->cf.w) : get(h->c.w);
}

}  // extern "C"

extern "C" {

static_assert(KI_COUNT == LGAME_N_TILES, "tile table and LGAME_N_TILES");
static_assert(PROF_KERNEL0 == LGAME_PROF_KERNEL0 && KI_SCALAR_DIV == 14, "per-kernel profile groups");

int lgame_autotune(lgame_handle *h, int steps, int *rows) {
  if (!h) return fail(h, LGAME_EINVAL, "null handle");
  if (steps < 1 || steps > 64) return fail(h, LGAME_EINVAL, "steps must be in 1..64");
  if (!h->have_state[0] || !h->have_state[1]) return fail(h, LGAME_ESTATE, "set both state slots first");
  if (h->stepped) return fail(h, LGAME_ESTATE, "autotune before the first step");
  // the tuning steps advance the state: keep a copy of both slots
  char *backup = nullptr;
  HIPCHK(h, hipMalloc(&backup, h->state_bytes));
  HIPCHK(h, hipMemcpyAsync(backup, h->slab, h->state_bytes, hipMemcpyDeviceToDevice, h->stream));
  const int cand[3] = {1, 2, 4};
  double best_ms[KI_COUNT];
  int best_by[KI_COUNT];
  for (int k = 0; k < KI_COUNT; ++k) {
    best_ms[k] = 1e300;
    best_by[k] = BY;
  }
  // run `steps` steps with the current tile table and sum each kernel's time (and the vertical
  // solver's, when solver_ms is given)
  auto measure = [&](double *ms, double *solver_ms = nullptr) {
    int rc = 0;
    h->ev_marks.clear();
    h->ev_next = 0;
    EventProf ep(h, true);
    for (int s = 1; s <= steps + 1; ++s) {  // counters >= 1: the steady-state kernel sequence
      if (s == 2) {                          // the first step warms up and is not counted
        rc = hipStreamSynchronize(h->stream) == hipSuccess ? 0 : LGAME_EHIP;
        h->ev_marks.clear();
        h->ev_next = 0;
      }
      const int a = s % 2, b = 1 - a;
      if (h->mixed)
        (void)enqueue_pchevi(h->cf, h->slot[a], h->slot[b], s, h->stream, &ep, (const RadCall *)nullptr, h->tiles);
      else
        (void)enqueue_pchevi(h->c, h->slot[a], h->slot[b], s, h->stream, &ep, (const RadCall *)nullptr, h->tiles);
    }
    if (rc || ep.failed || hipStreamSynchronize(h->stream) != hipSuccess) return rc ? rc : LGAME_EHIP;
    for (int k = 0; k < KI_COUNT; ++k) ms[k] = 0.0;
    if (solver_ms) *solver_ms = 0.0;
    for (auto &m : h->ev_marks) {
      const int g = m.first - PROF_KERNEL0;
      float t = 0.0f;
      if (m.first == 0 && solver_ms) {
        if (hipEventElapsedTime(&t, h->ev_pool[m.second.first], h->ev_pool[m.second.second]) == hipSuccess) *solver_ms += t;
        continue;
      }
      if (g < 0 || g >= KI_COUNT) continue;
      if (hipEventElapsedTime(&t, h->ev_pool[m.second.first], h->ev_pool[m.second.second]) == hipSuccess) ms[g] += t;
    }
    h->ev_marks.clear();
    h->ev_next = 0;
    return 0;
  };
  int rc = 0;
  double ms[KI_COUNT];
  for (int ci = 0; ci < 3 && !rc; ++ci) {
    for (int k = 0; k < KI_COUNT; ++k) h->tiles.by[k] = cand[ci];
    rc = measure(ms);
    for (int k = 0; k < KI_COUNT && !rc; ++k)
      if (ms[k] > 0.0 && ms[k] < best_ms[k]) {
        best_ms[k] = ms[k];
        best_by[k] = cand[ci];
      }
  }
  for (int k = 0; k < KI_COUNT; ++k) h->tiles.by[k] = best_by[k];
  // the vertical segments of the column-marching scalar kernel, with its best rows per block
  int best_seg = h->tiles.seg;
  double best_seg_ms = 1e300;
  for (int sg = 1; sg <= 8 && !rc; sg *= 2) {
    if (sg > 1 && h->c.G.nl / sg < 4) break;
    h->tiles.seg = sg;
    rc = measure(ms);
    if (!rc && ms[KI_SCALAR_DIV] > 0.0 && ms[KI_SCALAR_DIV] < best_seg_ms) {
      best_seg_ms = ms[KI_SCALAR_DIV];
      best_seg = sg;
    }
  }
  h->tiles.seg = best_seg;
  // the vertical solver in one or three passes, on domains near the size where they trade places
  if (!rc && h->c.G.nl >= 3 && col_waves(h->c.G) < 2 * VER_SPLIT_WAVES) {
    int best_p = h->tiles.ver_passes;
    double best_p_ms = 1e300;
    for (int p : {1, 3}) {
      if (rc) break;
      h->tiles.ver_passes = p;
      double sv = 0.0;
      rc = measure(ms, &sv);
      if (!rc && sv > 0.0 && sv < best_p_ms) {
        best_p_ms = sv;
        best_p = p;
      }
    }
    h->tiles.ver_passes = best_p;
  }
  // back to the state before tuning: both slots, and zero carried work arrays, as at lgame_init
  hipError_t e = hipMemcpyAsync(h->slab, backup, h->state_bytes, hipMemcpyDeviceToDevice, h->stream);
  if (e == hipSuccess) e = hipMemsetAsync(h->slab + h->work_off, 0, h->work_bytes, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e == hipSuccess) e = init_surface(h);
  (void)hipFree(backup);
  if (e != hipSuccess) return fail(h, LGAME_EHIP, "autotune restore: %s", hipGetErrorString(e));
  if (rc) return fail(h, rc, "autotune step failed");
  if (rows)
    for (int k = 0; k < KI_COUNT; ++k) rows[k] = best_by[k];
  return LGAME_OK;
}

int lgame_set_tiles(lgame_handle *h, const int *rows) {
  if (!h || !rows) return fail(h, LGAME_EINVAL, "null argument");
  for (int k = 0; k < KI_COUNT; ++k)
    if (rows[k] != 1 && rows[k] != 2 && rows[k] != 4) return fail(h, LGAME_EINVAL, "rows per block must be 1, 2 or 4");
  for (int k = 0; k < KI_COUNT; ++k) h->tiles.by[k] = rows[k];
  return LGAME_OK;
}

int lgame_set_segments(lgame_handle *h, int nseg) {
  if (!h) return fail(h, LGAME_EINVAL, "null handle");
  if (nseg != 1 && nseg != 2 && nseg != 4 && nseg != 8) return fail(h, LGAME_EINVAL, "segments must be 1, 2, 4 or 8");
  if (nseg > 1 && h->c.G.nl / nseg < 4) return fail(h, LGAME_EINVAL, "fewer than 4 levels per segment");
  h->tiles.seg = nseg;
  return LGAME_OK;
}

int lgame_get_segments(lgame_handle *h, int *nseg) {
  if (!h || !nseg) return fail(h, LGAME_EINVAL, "null argument");
  *nseg = h->tiles.seg;
  return LGAME_OK;
}

int lgame_set_column_passes(lgame_handle *h, int passes) {
  if (!h) return fail(h, LGAME_EINVAL, "null handle");
  if (passes != 1 && passes != 3) return fail(h, LGAME_EINVAL, "column passes must be 1 or 3");
  if (passes == 3 && h->c.G.nl < 3) return fail(h, LGAME_EINVAL, "three passes need at least 3 layers");
  h->tiles.ver_passes = passes;
  return LGAME_OK;
}

int lgame_get_column_passes(lgame_handle *h, int *passes) {
  if (!h || !passes) return fail(h, LGAME_EINVAL, "null argument");
  *passes = h->tiles.ver_passes;
  return LGAME_OK;
}

// ---- ensembles -----------------------------------------------------------------------------
}  // extern "C"

// Members of an ensemble whose steps run the same kernel sequence (same form and the same
// switches of enqueue_pchevi's branches, the same surrogate), advanced with one launch per phase.
struct BatchGroup {
  std::vector<lgame_handle *> m;
  bool mixed = false, rad = false, surrogate = false, o3 = false, moist = false;
  Cfg C{};                       // the switches (of the first member)
  void *mem = nullptr;           // MemberDev<R>[n] on the device
  void *host = nullptr;          // its host copy (pinned), rewritten at radiation updates
  size_t mem_bytes = 0;
  hipEvent_t uploaded = nullptr; // the last upload of `host` has been read
  int *prefix = nullptr;         // [PH_COUNT][n+1] first block of each member, per phase
  int total[PH_COUNT] = {0};     // blocks of each phase
  Tiles tiles;                   // rows per block (of the largest member)
  float *rx = nullptr, *ry = nullptr;  // the surrogate's inputs and outputs of every member's columns
  MlpWork *work = nullptr;
  long long cols = 0;
};

struct lgame_ensemble {
  std::vector<lgame_handle *> m;
  hipStream_t master = nullptr;
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> join;
  hipGraphExec_t exec[2] = {nullptr, nullptr};  // one graph per ping-pong parity
  hipGraph_t graph[2] = {nullptr, nullptr};
  std::vector<BatchGroup> groups;               // the batched launch (mode 2), made at its first step
  bool grouped = false;
  std::string err;
};

namespace {

int efail(lgame_ensemble *e, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (e) e->err = buf;
  return code;
}

#define ECHK(e, call)                                                                            \
  do {                                                                                           \
    hipError_t e_ = (call);                                                                      \
    if (e_ != hipSuccess) return efail(e, LGAME_EHIP, "%s: %s", #call, hipGetErrorString(e_));  \
  } while (0)

// A member's step as a branch of the ensemble's graph.  The branches run concurrently and fill the
// GPU together, so the vertical solver takes its one-pass form (the three-pass form of a small
// domain alone reads its inputs twice: the 48-member ensemble step took 74.2 instead of 71.8 ms).
void enqueue_member(lgame_handle *h, int a, int b, int step, const RadCall *rad) {
  Tiles t = h->tiles;
  t.ver_passes = 1;
  if (h->mixed)
    enqueue_pchevi(h->cf, h->slot[a], h->slot[b], step, h->stream, nullptr, rad, t);
  else
    enqueue_pchevi(h->c, h->slot[a], h->slot[b], step, h->stream, nullptr, rad, t);
}

// all members' steps of one parity as parallel branches of one graph (fork/join over the
// members' own streams)
int capture(lgame_ensemble *e, int parity) {
  const int a = parity, b = 1 - parity;
  ECHK(e, hipStreamBeginCapture(e->master, hipStreamCaptureModeThreadLocal));
  ECHK(e, hipEventRecord(e->fork, e->master));
  for (size_t q = 0; q < e->m.size(); ++q) {
    lgame_handle *h = e->m[q];
    ECHK(e, hipStreamWaitEvent(h->stream, e->fork, 0));
    enqueue_member(h, a, b, 1 + parity, nullptr);  // any counter >= 1 of this parity: the steady sequence
    ECHK(e, hipEventRecord(e->join[q], h->stream));
    ECHK(e, hipStreamWaitEvent(e->master, e->join[q], 0));
  }
  ECHK(e, hipStreamEndCapture(e->master, &e->graph[parity]));
  ECHK(e, hipGraphInstantiate(&e->exec[parity], e->graph[parity], nullptr, nullptr, 0));
  return 0;
}

// the branches of enqueue_pchevi a handle's step takes, and its form: members with the same key
// run the same kernels in the same order
unsigned long long seq_key(const lgame_handle *h) {
  const Cfg &C = h->c.C;
  const int bits[] = {h->mixed, C.order == 3, C.llinear, C.lcorio, C.lmom_h, C.lmom_v, C.lklemp, C.lpbl, C.ltemp_h,
                      C.lmass_h, C.lsoil, C.lsens, C.lsfc_turb, h->c.G.nc > 1, h->cfg.lrad,
                      h->cfg.lrad && h->cfg.rad_mode == LGAME_RAD_SURROGATE};
  unsigned long long k = 0;
  for (int b : bits) k = 2 * k + (b ? 1 : 0);
  return k;
}

void free_group(BatchGroup &g) {
  if (g.mem) (void)hipFree(g.mem);
  if (g.host) (void)hipHostFree(g.host);
  if (g.prefix) (void)hipFree(g.prefix);
  if (g.rx) (void)hipFree(g.rx);
  if (g.ry) (void)hipFree(g.ry);
  if (g.work) mlp_work_destroy(g.work);
  if (g.uploaded) (void)hipEventDestroy(g.uploaded);
  g = BatchGroup();
}

template <class R>
const CtxT<R> &ctx_of(const lgame_handle *h);
template <>
const CtxT<double> &ctx_of<double>(const lgame_handle *h) {
  return h->c;
}
template <>
const CtxT<float> &ctx_of<float>(const lgame_handle *h) {
  return h->cf;
}

// the device table of a group: members, prefix sums of every phase's blocks, and for the
// surrogate the concatenated input and output rows of all members' columns
template <class R>
int build_group(lgame_ensemble *e, BatchGroup &g) {
  const int n = (int)g.m.size();
  const lgame_handle *ref = g.m[0];
  for (auto *h : g.m)
    if ((long long)h->c.G.nx * h->c.G.ny * h->c.G.nl > (long long)ref->c.G.nx * ref->c.G.ny * ref->c.G.nl) ref = h;
  g.tiles = ref->tiles;
  g.mem_bytes = sizeof(MemberDev<R>) * n;
  ECHK(e, hipHostMalloc(&g.host, g.mem_bytes, hipHostMallocDefault));
  ECHK(e, hipMalloc(&g.mem, g.mem_bytes));
  ECHK(e, hipMalloc(&g.prefix, sizeof(int) * PH_COUNT * (n + 1)));
  ECHK(e, hipEventCreateWithFlags(&g.uploaded, hipEventDisableTiming));
  for (auto *h : g.m) g.cols += (long long)h->c.G.nx * h->c.G.ny;
  MemberDev<R> *M = static_cast<MemberDev<R> *>(g.host);
  // The column-marching scalar kernel cuts a small domain's columns into segments to fill the
  // GPU; in one launch over the whole group the members' rows fill it together, so the segments
  // (each with one extra layer) are dropped when the group has four wavefronts per SIMD without
  // them.
  long long waves = 0;
  for (auto *h : g.m) waves += (long long)((h->c.G.nx + SCOL_CELLS - 1) / SCOL_CELLS) * h->c.G.ny;
  const bool one_seg = waves >= 4096;
  for (int q = 0; q < n; ++q) {
    lgame_handle *h = g.m[q];
    MemberDev<R> &d = M[q];
    std::memset(static_cast<void *>(&d), 0, sizeof d);
    d.c = ctx_of<R>(h);
    d.s[0] = h->slot[0];
    d.s[1] = h->slot[1];
    d.seg = one_seg ? 1 : h->tiles.seg;  // the radiation part is filled at each radiation update
  }
  std::vector<int> pf((size_t)PH_COUNT * (n + 1), 0);
  for (int ph = 0; ph < PH_COUNT; ++ph) {
    const int tile = phase_tile(ph);
    const int by = tile >= 0 ? g.tiles.by[tile] : 1;
    int *p = pf.data() + (size_t)ph * (n + 1);
    for (int q = 0; q < n; ++q) {
      int gx, gy, gz;
      phase_grid(ph, g.m[q]->c.G, by, M[q].seg, gx, gy, gz);
      p[q + 1] = p[q] + gx * gy * gz;
    }
    g.total[ph] = p[n];
  }
  ECHK(e, hipMemcpyAsync(g.prefix, pf.data(), pf.size() * sizeof(int), hipMemcpyHostToDevice, e->master));
  ECHK(e, hipMemcpyAsync(g.mem, g.host, g.mem_bytes, hipMemcpyHostToDevice, e->master));
  ECHK(e, hipEventRecord(g.uploaded, e->master));
  ECHK(e, hipStreamSynchronize(e->master));  // pf leaves scope
  return 0;
}

int build_groups(lgame_ensemble *e) {
  for (auto *h : e->m) {
    bool placed = false;
    for (auto &g : e->groups)
      if (!placed && seq_key(g.m[0]) == seq_key(h) && g.m[0]->mlp == h->mlp) {
        g.m.push_back(h);
        placed = true;
      }
    if (!placed) {
      e->groups.emplace_back();
      BatchGroup &g = e->groups.back();
      g.m.push_back(h);
      g.mixed = h->mixed;
      g.C = h->c.C;
      g.o3 = h->c.C.order == 3;
      g.moist = h->c.G.nc > 1;
      g.rad = h->cfg.lrad != 0;
      g.surrogate = g.rad && h->cfg.rad_mode == LGAME_RAD_SURROGATE;
    }
  }
  for (auto &g : e->groups) {
    const int rc = g.mixed ? build_group<float>(e, g) : build_group<double>(e, g);
    if (rc) {  // leave no half-made groups behind: the next step starts over
      for (auto &q : e->groups) free_group(q);
      e->groups.clear();
      return rc;
    }
  }
  e->grouped = true;
  return 0;
}

// The radiation part of a group's device table at an update: every member's own radiation data,
// with the surrogate's inputs and outputs in the group's matrices (made at the first update, with
// the members' network, which must be one and the same), the call index and the time.
template <class R>
int prepare_group_radiation(lgame_ensemble *e, BatchGroup &g, double t_0) {
  if (g.surrogate) {
    const lgame_mlp *mlp = g.m[0]->mlp;
    for (auto *h : g.m)
      if (h->mlp != mlp) return efail(e, LGAME_ESTATE, "the members of a batched group need the same surrogate network");
    if (!g.work) {
      ECHK(e, hipMalloc(&g.rx, (size_t)g.cols * LGAME_RAD_N_IN * sizeof(float)));
      ECHK(e, hipMalloc(&g.ry, (size_t)g.cols * LGAME_RAD_N_OUT * sizeof(float)));
      const int rc = mlp_work_create(mlp, g.cols < 262144 ? g.cols : 262144, &g.work);
      if (rc) return efail(e, rc, "surrogate inference buffers of the group");
    }
  }
  ECHK(e, hipEventSynchronize(g.uploaded));  // the launches that read the previous table are done with it
  MemberDev<R> *H = static_cast<MemberDev<R> *>(g.host);
  long long row = 0;
  for (size_t q = 0; q < g.m.size(); ++q) {
    lgame_handle *h = g.m[q];
    H[q].c = ctx_of<R>(h);  // with the surface-flux pointers of a radiation set up after the grouping
    H[q].rad = rad_dev(h);
    if (g.surrogate) {
      H[q].rad.x = g.rx + (size_t)row * LGAME_RAD_N_IN;
      H[q].rad.ypred = g.ry + (size_t)row * LGAME_RAD_N_OUT;
    }
    H[q].rad.call = h->rad_calls++;
    H[q].t_rad = t_0 + 0.5 * h->cfg.dtime_rad;  // mo_manage_radiation_calls.F90:53
    row += (long long)h->c.G.nx * h->c.G.ny;
  }
  ECHK(e, hipMemcpyAsync(g.mem, g.host, g.mem_bytes, hipMemcpyHostToDevice, e->master));
  ECHK(e, hipEventRecord(g.uploaded, e->master));
  return 0;
}

// One step of a group: enqueue_pchevi's sequence with every phase one launch over all members.
template <class R>
int enqueue_group(lgame_ensemble *e, BatchGroup &g, int a_slot, int b_slot, int step, bool rad_update, double t_0) {
  const int n = (int)g.m.size();
  const hipStream_t st = e->master;
  const MemberDev<R> *M = static_cast<const MemberDev<R> *>(g.mem);
  const Cfg &C = g.C;
  auto launch = [&](int ph, const BatchArgs &a) {
    const dim3 grd((unsigned)g.total[ph]);
    const int *pf = g.prefix + (size_t)ph * (n + 1);
    dim3 blk(BX, 1, 1);
    if (ph == PH_SOLVE_VER)
      blk = dim3(BCOL);
    else if (ph == PH_SOLVE_GEN)
      blk = dim3(BCOL, NTR * GEN_GROUPS);
    else
      blk = dim3(BX, g.tiles.by[phase_tile(ph)]);
    if (ph == PH_SCALAR) {
      batch_launch_scalar<R>(grd, blk, st, M, pf, n, a, g.moist, g.o3, C.ltemp_h || C.lmass_h);
      return;
    }
    if (!batch_launch_step<R>(ph, grd, blk, st, M, pf, n, a) && !batch_launch_diff<R>(ph, grd, blk, st, M, pf, n, a))
      (void)batch_launch_solve<R>(ph, grd, blk, st, M, pf, n, a);
  };
  BatchArgs base{a_slot, b_slot, a_slot, 1, step == 0 ? 1 : 0, 0, 0, C.llinear ? 0 : 1};
  launch(PH_TEMPERATURE, base);
  // update_sfc_turb_quantities and update_rad_fluxes with state_old (mo_manage_pchevi.F90:52-58)
  if (C.lsfc_turb) launch(PH_SFC_TURB, base);
  if (rad_update && g.rad) {
    const int rrc = prepare_group_radiation<R>(e, g, t_0);
    if (rrc) return rrc;
    const dim3 grd((unsigned)g.total[PH_RAD]), blk(BCOL);
    const int *pf = g.prefix + (size_t)PH_RAD * (n + 1);
    BatchArgs a = base;
    if (!g.surrogate) {
      a.flags = RAD_TEND;
      batch_launch_radiation<R>(RPH_SOLVER, grd, blk, st, M, pf, n, a);
    } else {
      batch_launch_radiation<R>(RPH_GATHER, grd, blk, st, M, pf, n, a);
      const int rc = mlp_infer_device(g.m[0]->mlp, g.work, g.rx, g.cols, g.ry, st);
      if (rc) return efail(e, rc, "surrogate inference of the group");
      batch_launch_radiation<R>(RPH_APPLY, grd, blk, st, M, pf, n, a);
      batch_launch_radiation<R>(RPH_SELECT, grd, blk, st, M, pf, n, a);
      a.flags = RAD_AUDIT;
      batch_launch_radiation<R>(RPH_SOLVER, grd, blk, st, M, pf, n, a);
    }
  }
  for (int rk = 1; rk <= 2; ++rk) {
    BatchArgs a = base;
    a.rk = rk;
    a.state_slot = rk == 1 ? a_slot : b_slot;
    if (rk == 1) {
      if (g.o3) {
        BatchArgs t = a;
        t.state_slot = a_slot;
        t.nfield = 2;
        for (int ph : {PH_THETA_GRAD, PH_THETA_DIV, PH_THETA_EDGE}) launch(ph, t);
      }
    }
    bool winds_done = false;
    if ((rk == 2 || step == 0) && (!C.llinear || C.lcorio)) {
      launch(PH_POT_VORT, a);  // with the mass fluxes and the kinetic energy
      // with the gradient of the kinetic energy, and at rk 2 the wind update
      launch(rk == 2 ? PH_VORT_FLUX_VT : PH_VORT_FLUX, a);
      winds_done = rk == 2;
    }
    if (rk == 1 && C.lmom_h)
      for (int ph : {PH_DIFF_ZETA, PH_DIFF_COEF, PH_DIFF_DUAL, PH_DIFF_MOM}) launch(ph, a);
    if (rk == 1 && C.lpbl) launch(PH_PBL_WIND, a);
    if (rk == 1 && (C.lklemp || C.lmom_h || C.lmom_v || C.lpbl)) launch(PH_DISSIPATION, a);
    if (rk == 1)
      launch(PH_PGRAD, a);  // the pressure gradient with the wind update
    else if (!winds_done)
      launch(PH_VEC_TEND, a);
    if (rk == 1 && (C.ltemp_h || C.lmass_h)) {
      launch(PH_DIFF_FLUX, a);
      launch(PH_DIFF_DIV, a);
    }
    if (g.o3 && rk == 2) {
      BatchArgs t = a;
      t.state_slot = b_slot;
      t.nfield = 1;
      for (int ph : {PH_THETA_GRAD, PH_THETA_DIV, PH_THETA_EDGE}) launch(ph, t);
    }
    launch(PH_SCALAR, a);
    if (C.lsoil || C.lsens) launch(PH_SURFACE, a);
    launch(PH_SOLVE_VER, a);
    if (g.moist) launch(PH_SOLVE_GEN, a);
  }
  ECHK(e, hipGetLastError());
  return 0;
}

}  // namespace

extern "C" {

int lgame_ensemble_create(lgame_handle *const *members, int n, lgame_ensemble **out) {
  if (!members || n < 1 || !out) return LGAME_EINVAL;
  *out = nullptr;
  lgame_ensemble *e = new lgame_ensemble();
  for (int q = 0; q < n; ++q) {
    // a member listed twice would be stepped twice per ensemble step on one stream
    bool dup = false;
    for (int p = 0; p < q; ++p) dup = dup || members[p] == members[q];
    if (!members[q] || dup) {
      delete e;
      return LGAME_EINVAL;
    }
    e->m.push_back(members[q]);
  }
  bool ok = hipStreamCreateWithFlags(&e->master, hipStreamNonBlocking) == hipSuccess;
  ok = ok && hipEventCreateWithFlags(&e->fork, hipEventDisableTiming) == hipSuccess;
  e->join.resize(n, nullptr);
  for (int q = 0; q < n && ok; ++q) ok = hipEventCreateWithFlags(&e->join[q], hipEventDisableTiming) == hipSuccess;
  if (!ok) {
    lgame_ensemble_destroy(e);
    return LGAME_EHIP;
  }
  *out = e;
  return LGAME_OK;
}

// One step of every member (the members' own ping-pong, counter total_step_counter).  Reading a
// member's state or budget sums needs lgame_ensemble_sync first (graph work runs on the
// ensemble's own stream).  Steps at
// counter 0 and steps with a radiation update run eagerly on the members' streams (the radiation
// surrogate's buffers are shared, so those members run one after another); all other steps
// replay one captured graph for the whole ensemble.  mode 1 runs every member eagerly and
// serially instead (the comparison without the batched launch).
int lgame_ensemble_step(lgame_ensemble *e, int total_step_counter, int lrad_update, double t_0, int mode) {
  if (!e) return LGAME_EINVAL;
  const int a = total_step_counter % 2, b = 1 - a;
  for (auto *h : e->m)
    if (!h->have_state[0] || !h->have_state[1]) return efail(e, LGAME_ESTATE, "every member needs both state slots");
  if (mode == 2) {  // one launch per phase over every member of a group
    for (auto *h : e->m)
      if (h->cfg.lrad && lrad_update && (!h->rad_surface || (h->cfg.rad_mode == LGAME_RAD_SURROGATE && !h->mlp)))
        return efail(e, LGAME_ESTATE, "a member's radiation is not set up (surface data, surrogate)");
    if (!e->grouped) {
      const int rc = build_groups(e);
      if (rc) return rc;
    }
    for (size_t q = 0; q < e->m.size(); ++q) {
      ECHK(e, hipEventRecord(e->join[q], e->m[q]->stream));
      ECHK(e, hipStreamWaitEvent(e->master, e->join[q], 0));
    }
    for (auto &g : e->groups) {
      const int rc = g.mixed ? enqueue_group<float>(e, g, a, b, total_step_counter, lrad_update != 0, t_0)
                             : enqueue_group<double>(e, g, a, b, total_step_counter, lrad_update != 0, t_0);
      if (rc) return rc;
    }
    for (auto *h : e->m) h->stepped = 1;
    return LGAME_OK;
  }
  if (mode == 0 && total_step_counter > 0 && !lrad_update) {
    if (!e->exec[a]) {
      int rc = capture(e, a);
      if (rc) return rc;
    }
    // the graph runs on the master stream after everything already queued on the members' streams
    for (size_t q = 0; q < e->m.size(); ++q) {
      ECHK(e, hipEventRecord(e->join[q], e->m[q]->stream));
      ECHK(e, hipStreamWaitEvent(e->master, e->join[q], 0));
    }
    ECHK(e, hipGraphLaunch(e->exec[a], e->master));
    for (auto *h : e->m) h->stepped = 1;
    return LGAME_OK;
  }
  // eager steps on the members' streams start after the graph launches queued on the master
  ECHK(e, hipEventRecord(e->fork, e->master));
  for (auto *h : e->m) ECHK(e, hipStreamWaitEvent(h->stream, e->fork, 0));
  const bool serial = mode == 1 || lrad_update;
  for (auto *h : e->m) {
    int rc = lgame_pchevi(h, a, b, total_step_counter, lrad_update, t_0);
    if (rc) return efail(e, rc, "member step: %s", h->err.c_str());
    if (serial) ECHK(e, hipStreamSynchronize(h->stream));
  }
  return LGAME_OK;
}

int lgame_ensemble_sync(lgame_ensemble *e) {
  if (!e) return LGAME_EINVAL;
  ECHK(e, hipStreamSynchronize(e->master));
  for (auto *h : e->m) ECHK(e, hipStreamSynchronize(h->stream));
  return LGAME_OK;
}

const char *lgame_ensemble_last_error(const lgame_ensemble *e) { return e ? e->err.c_str() : ""; }

void lgame_ensemble_destroy(lgame_ensemble *e) {
  if (!e) return;
  if (e->master) (void)hipStreamSynchronize(e->master);
  for (int p = 0; p < 2; ++p) {
    if (e->exec[p]) (void)hipGraphExecDestroy(e->exec[p]);
    if (e->graph[p]) (void)hipGraphDestroy(e->graph[p]);
  }
  for (auto &g : e->groups) free_group(g);
  for (auto ev : e->join)
    if (ev) (void)hipEventDestroy(ev);
  if (e->fork) (void)hipEventDestroy(e->fork);
  if (e->master) (void)hipStreamDestroy(e->master);
  delete e;
}

int lgame_rad_set_surface(lgame_handle *h, const lgame_rad_view *v) {
  if (!h || !v || !v->lat_geo || !v->lon_geo || !v->sfc_albedo || !v->temp_sfc) return fail(h, LGAME_EINVAL, "null argument");
  int rc = rad_alloc(h);
  if (rc) return rc;
  const int ny = h->dims.ny, nx = h->dims.nx;
  const size_t ncol = (size_t)ny * nx;
  const double *src[4] = {v->lat_geo, v->lon_geo, v->sfc_albedo, v->temp_sfc};
  std::vector<double> tmp;
  for (int q = 0; q < 4; ++q) {
    to_device_cols(src[q], tmp, ny, nx);
    HIPCHK(h, hipMemcpy(h->rad_mem + q * ncol, tmp.data(), ncol * sizeof(double), hipMemcpyHostToDevice));
  }
  h->rad_surface = 1;
  return LGAME_OK;
}

int lgame_rad_columns(lgame_handle *h, int slot, double t, float *x, float *y) {
  if (!h) return fail(h, LGAME_EINVAL, "null handle");
  if ((slot != 0 && slot != 1) || !h->have_state[slot]) return fail(h, LGAME_ESTATE, "slot %d holds no state", slot);
  if (!h->rad_surface) return fail(h, LGAME_ESTATE, "lgame_rad_set_surface first");
  const Geom &G = h->c.G;
  const int ny = G.ny, nx = G.nx, nl = G.nl, nin = 5 * nl + 4, nout = nl + 2;
  const size_t ncol = (size_t)ny * nx;
  RadDev r = rad_dev(h);
  const dim3 gc((unsigned)((ncol + BCOL - 1) / BCOL)), bc(BCOL);
  if (h->mixed) {
    hipLaunchKernelGGL(k_temperature<float>, grid3(G.nx, G.ny, G.nl), dim3(BX, BY, 1), 0, h->stream, h->cf, h->slot[slot]);
    hipLaunchKernelGGL(k_rad<float>, gc, bc, 0, h->stream, h->cf, h->slot[slot], r, t, RAD_X | RAD_Y);
  } else {
    hipLaunchKernelGGL(k_temperature<double>, grid3(G.nx, G.ny, G.nl), dim3(BX, BY, 1), 0, h->stream, h->c, h->slot[slot]);
    hipLaunchKernelGGL(k_rad<double>, gc, bc, 0, h->stream, h->c, h->slot[slot], r, t, RAD_X | RAD_Y);
  }
  HIPCHK(h, hipGetLastError());
  std::vector<float> hx(ncol * nin), hy(ncol * nout);
  HIPCHK(h, hipMemcpyAsync(hx.data(), h->rad_x, hx.size() * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipMemcpyAsync(hy.data(), h->rad_y, hy.size() * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  for (int jk = 0; jk < nx; ++jk)
    for (int ji = 0; ji < ny; ++ji) {
      const size_t dc = (size_t)ji * nx + jk, fc = ji + (size_t)ny * jk;
      if (x) std::memcpy(x + fc * nin, hx.data() + dc * nin, nin * sizeof(float));
      if (y) std::memcpy(y + fc * nout, hy.data() + dc * nout, nout * sizeof(float));
    }
  return LGAME_OK;
}

int lgame_rad_get(lgame_handle *h, int which, double *out) {
  if (!h || !out) return fail(h, LGAME_EINVAL, "null argument");
  const int ny = h->dims.ny, nx = h->dims.nx, nl = h->dims.n_layers;
  if (which == LGAME_RAD_OUT_TENDENCY) {
    if (h->mixed) return download(h, out, h->cf.w.radiation_tendency.p, ny, nx, nl);
    return download(h, out, h->c.w.radiation_tendency.p, ny, nx, nl);
  }
  if (which != LGAME_RAD_OUT_SW_IN && which != LGAME_RAD_OUT_LW_OUT) return fail(h, LGAME_EINVAL, "unknown radiation output %d", which);
  if (!h->rad_mem) return fail(h, LGAME_ESTATE, "no radiation data");
  const size_t ncol = (size_t)ny * nx;
  std::vector<double> tmp(ncol);
  HIPCHK(h, hipStreamSynchronize(h->stream));
  HIPCHK(h, hipMemcpy(tmp.data(), h->rad_mem + (which == LGAME_RAD_OUT_SW_IN ? 4 : 5) * ncol, ncol * sizeof(double),
                      hipMemcpyDeviceToHost));
  for (int jk = 0; jk < nx; ++jk)
    for (int ji = 0; ji < ny; ++ji) out[ji + (size_t)ny * jk] = tmp[(size_t)ji * nx + jk];
  return LGAME_OK;
}

int lgame_rad_attach_mlp(lgame_handle *h, lgame_mlp *m, double tol_k_per_day, double audit_fraction, unsigned long long seed) {
  if (!h || !m) return fail(h, LGAME_EINVAL, "null argument");
  if (h->cfg.rad_mode != LGAME_RAD_SURROGATE) return fail(h, LGAME_EINVAL, "the handle is not configured for the surrogate");
  int n_in = 0, n_out = 0;
  mlp_dims(m, &n_in, &n_out);
  if (n_in != LGAME_RAD_N_IN || n_out != LGAME_RAD_N_OUT)
    return fail(h, LGAME_EINVAL, "the network must map %d inputs to %d outputs", LGAME_RAD_N_IN, LGAME_RAD_N_OUT);
  if (h->dims.n_layers < 2) return fail(h, LGAME_EINVAL, "the surrogate needs at least 2 layers");
  if (!mlp_has_norm(m))
    return fail(h, LGAME_EINVAL, "the network has no normalisation (lgame_mlp_set_norm): its output scales define the envelope");
  if (!(tol_k_per_day > 0.0) || !(audit_fraction >= 0.0 && audit_fraction <= 1.0))
    return fail(h, LGAME_EINVAL, "tolerance must be positive and the audit fraction in [0, 1]");
  int rc = rad_alloc(h);
  if (rc) return rc;
  // the handle's own inference buffers: the step allocates nothing, and handles sharing the
  // network do not share activations
  HIPCHK(h, hipStreamSynchronize(h->stream));
  mlp_work_destroy(h->mlp_work);
  h->mlp_work = nullptr;
  const long long ncol = (long long)h->dims.ny * h->dims.nx;
  if (mlp_work_create(m, ncol < 65536 ? ncol : 65536, &h->mlp_work))
    return fail(h, LGAME_ENOMEM, "inference buffers of the surrogate");
  h->mlp = m;
  h->rad_tol = tol_k_per_day;
  h->rad_thresh = audit_fraction >= 1.0 ? 0xffffffffu : (unsigned)(audit_fraction * 4294967296.0);
  h->rad_seed = seed;
  h->rad_calls = seed * 1000003ull;
  return LGAME_OK;
}

int lgame_rad_stats(lgame_handle *h, long long *calls, long long *audited, long long *fallbacks, double *audit_rmse) {
  if (!h || !calls || !audited || !fallbacks || !audit_rmse) return fail(h, LGAME_EINVAL, "null argument");
  *calls = *audited = *fallbacks = 0;
  *audit_rmse = 0.0;
  if (!h->rad_counts) return LGAME_OK;
  unsigned long long c[2];
  double sq = 0.0;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  HIPCHK(h, hipMemcpy(c, h->rad_counts, sizeof c, hipMemcpyDeviceToHost));
  HIPCHK(h, hipMemcpy(&sq, h->rad_sq, sizeof sq, hipMemcpyDeviceToHost));
  *calls = (long long)(h->rad_calls - h->rad_seed * 1000003ull);
  *audited = (long long)c[0];
  *fallbacks = (long long)c[1];
  *audit_rmse = c[0] ? std::sqrt(sq / ((double)c[0] * h->dims.n_layers)) : 0.0;
  return LGAME_OK;
}

}  // extern "C"

extern "C" {

void lgame_finalize(lgame_handle *h) {
  if (!h) return;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  if (h->slab) (void)hipFree(h->slab);
  if (h->staging) (void)hipFree(h->staging);
  if (h->red) (void)hipFree(h->red);
  if (h->counter) (void)hipFree(h->counter);
  if (h->trsk) (void)hipFree(h->trsk);
  if (h->soil_z) (void)hipFree(h->soil_z);
  for (void *p : {(void *)h->rad_mem, (void *)h->rad_x, (void *)h->rad_y, (void *)h->rad_counts, (void *)h->rad_sq,
                  (void *)h->rad_list, (void *)h->rad_list_count})
    if (p) (void)hipFree(p);
  mlp_work_destroy(h->mlp_work);
  for (auto e : h->ev_pool) (void)hipEventDestroy(e);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

}  // extern "C"
