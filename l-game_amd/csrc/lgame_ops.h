// lgame_ops.h -- per-point device functions of the HEVI step operators.
//
// Each function returns the value the reference operator stores at one grid point, with the
// same floating-point operation order as the reference Fortran (the library is built with
// -ffp-contract=off, like the reference build).  Periodic boundaries: neighbour indices wrap;
// the duplicate edge points (x index nx+1, y index ny+1) follow the reference's semantics for
// each operator (a copy of index 1, or their own metrics where the reference computes them).
//
// The functions are templates on the arithmetic type R of the stencil kernels (double in the
// plain form, float in the mixed form).  Operands that are FP64 state fields are converted to R
// on load, and literals are written R(...), so that with R = double every expression is the
// FP64 expression of the reference and with R = float all arithmetic is FP32.
#pragma once

#include "lgame_dev.h"

namespace lgame {

#define AT(p, ji, jk, jl) (p)[LG_IX(c.G, ji, jk, jl)]
// a field value converted to the stencil type R
#define ATR(p, ji, jk, jl) static_cast<R>(AT(p, ji, jk, jl))

__device__ __forceinline__ int wrapm(int i, int n) { return i == 1 ? n : i - 1; }

// ---- vertical fields seen through remap_ver2hor -----------------------------------------
// grad_vert of a scalar field (mo_gradient_operators.F90:65-83); the reference leaves levels 1
// and n_levels of its result untouched and they hold zero whenever the remap reads them.
template <class R, class P>
struct VGrad {
  P s;
  __device__ R operator()(const CtxT<R> &c, int ji, int jk, int jl) const {
    if (jl >= c.G.nlev || jl <= 1) return R(0.0);
    return (ATR(s, ji, jk, jl - 1) - ATR(s, ji, jk, jl)) / AT(c.g.dz, ji, jk, jl);
  }
};
template <class R, class P>
struct VField {
  P w;
  __device__ R operator()(const CtxT<R> &c, int ji, int jk, int jl) const { return ATR(w, ji, jk, jl); }
};

// remap_ver2hor_x, mo_averaging.F90:56-95 (periodic)
template <class R, class VZ>
__device__ R remap_x(const CtxT<R> &c, const VZ &vz, int ji, int jk, int jl) {
  const int nx = c.G.nx;
  int ka, kb;
  if (jk == 1 || jk == nx + 1) {
    ka = nx;
    kb = 1;
  } else {
    ka = jk - 1;
    kb = jk;
  }
  R r = R(0.0);
  r = r + AT(c.g.ipw[4], ji, ka, jl) * vz(c, ji, ka, jl);
  r = r + AT(c.g.ipw[4], ji, kb, jl) * vz(c, ji, kb, jl);
  if (jl < c.G.nl || c.C.lfreeslip) {
    r = r + AT(c.g.ipw[5], ji, ka, jl) * vz(c, ji, ka, jl + 1);
    r = r + AT(c.g.ipw[5], ji, kb, jl) * vz(c, ji, kb, jl + 1);
  }
  return R(0.5) * r;
}

// remap_ver2hor_y, mo_averaging.F90:97-136 (periodic)
template <class R, class VZ>
__device__ R remap_y(const CtxT<R> &c, const VZ &vz, int ji, int jk, int jl) {
  const int ny = c.G.ny;
  int ia, ib;
  if (ji == 1 || ji == ny + 1) {
    ia = ny;
    ib = 1;
  } else {
    ia = ji - 1;
    ib = ji;
  }
  R r = R(0.0);
  r = r + AT(c.g.ipw[4], ia, jk, jl) * vz(c, ia, jk, jl);
  r = r + AT(c.g.ipw[4], ib, jk, jl) * vz(c, ib, jk, jl);
  if (jl < c.G.nl || c.C.lfreeslip) {
    r = r + AT(c.g.ipw[5], ia, jk, jl) * vz(c, ia, jk, jl + 1);
    r = r + AT(c.g.ipw[5], ib, jk, jl) * vz(c, ib, jk, jl + 1);
  }
  return R(0.5) * r;
}

// grad_hor x-component (mo_gradient_operators.F90:17-63,85-130); at jk = nx+1 the covariant part
// is the copy of jk = 1 and the terrain correction uses the metrics of jk = nx+1.
template <class R, class P>
__device__ __forceinline__ R grad_hor_x(const CtxT<R> &c, P s, int ji, int jk, int jl) {
  const int nx = c.G.nx;
  R r;
  if (jk == 1 || jk == nx + 1)
    r = (ATR(s, ji, 1, jl) - ATR(s, ji, nx, jl)) / AT(c.g.dx, ji, 1, jl);
  else
    r = (ATR(s, ji, jk, jl) - ATR(s, ji, jk - 1, jl)) / AT(c.g.dx, ji, jk, jl);
  if (jl > c.G.nflat) r = r - AT(c.g.slope_x, ji, jk, jl) * remap_x(c, VGrad<R, P>{s}, ji, jk, jl);
  return r;
}

template <class R, class P>
__device__ __forceinline__ R grad_hor_y(const CtxT<R> &c, P s, int ji, int jk, int jl) {
  const int ny = c.G.ny;
  R r;
  if (ji == 1 || ji == ny + 1)
    r = (ATR(s, ny, jk, jl) - ATR(s, 1, jk, jl)) / AT(c.g.dy, 1, jk, jl);
  else
    r = (ATR(s, ji - 1, jk, jl) - ATR(s, ji, jk, jl)) / AT(c.g.dy, ji, jk, jl);
  if (jl > c.G.nflat) r = r - AT(c.g.slope_y, ji, jk, jl) * remap_y(c, VGrad<R, P>{s}, ji, jk, jl);
  return r;
}

// vertical_contravariant_corr, mo_averaging.F90:16-54
template <class R, class P>
__device__ R vcc(const CtxT<R> &c, P vx, P vy, int ji, int jk, int jl) {
  R r = R(0.0);
  if (jl >= c.G.nflat + 1) {
    if (jl != c.G.nflat + 1) {
      const int l = jl - 1;
      r = r - R(0.5) * ATR(vx, ji, jk + 1, l) * AT(c.g.slope_x, ji, jk + 1, l) * AT(c.g.ipw[0], ji, jk, l) -
          R(0.5) * ATR(vy, ji, jk, l) * AT(c.g.slope_y, ji, jk, l) * AT(c.g.ipw[1], ji, jk, l) -
          R(0.5) * ATR(vx, ji, jk, l) * AT(c.g.slope_x, ji, jk, l) * AT(c.g.ipw[2], ji, jk, l) -
          R(0.5) * ATR(vy, ji + 1, jk, l) * AT(c.g.slope_y, ji + 1, jk, l) * AT(c.g.ipw[3], ji, jk, l);
    }
    r = r - R(0.5) * ATR(vx, ji, jk + 1, jl) * AT(c.g.slope_x, ji, jk + 1, jl) * AT(c.g.ipw[0], ji, jk, jl) -
        R(0.5) * ATR(vy, ji, jk, jl) * AT(c.g.slope_y, ji, jk, jl) * AT(c.g.ipw[1], ji, jk, jl) -
        R(0.5) * ATR(vx, ji, jk, jl) * AT(c.g.slope_x, ji, jk, jl) * AT(c.g.ipw[2], ji, jk, jl) -
        R(0.5) * ATR(vy, ji + 1, jk, jl) * AT(c.g.slope_y, ji + 1, jk, jl) * AT(c.g.ipw[3], ji, jk, jl);
  }
  return r;
}

// div_h, mo_divergence_operators.F90:16-68
template <class R, class P>
__device__ R div_h(const CtxT<R> &c, P vx, P vy, int ji, int jk, int jl) {
  R comp_h = ATR(vx, ji, jk + 1, jl) * AT(c.g.area_x, ji, jk + 1, jl) + ATR(vy, ji, jk, jl) * AT(c.g.area_y, ji, jk, jl) -
             ATR(vx, ji, jk, jl) * AT(c.g.area_x, ji, jk, jl) - ATR(vy, ji + 1, jk, jl) * AT(c.g.area_y, ji + 1, jk, jl);
  R comp_v = R(0.0);
  const int nflat = c.G.nflat, nl = c.G.nl;
  if (jl == nflat) {
    R cl = vcc(c, vx, vy, ji, jk, jl + 1);
    comp_v = -cl * AT(c.g.area_z, ji, jk, jl + 1);
  } else if (jl == nl) {
    R cu = vcc(c, vx, vy, ji, jk, jl);
    comp_v = cu * AT(c.g.area_z, ji, jk, jl);
  } else if (jl > nflat) {
    R cu = vcc(c, vx, vy, ji, jk, jl);
    R cl = vcc(c, vx, vy, ji, jk, jl + 1);
    comp_v = cu * AT(c.g.area_z, ji, jk, jl) - cl * AT(c.g.area_z, ji, jk, jl + 1);
  }
  return (comp_h + comp_v) / AT(c.g.volume, ji, jk, jl);
}

// horizontal_covariant_x/_y, mo_averaging.F90:138-176 (vertical component = wind_w)
template <class R>
__device__ __forceinline__ R hcov_x(const CtxT<R> &c, Fld<const double> u, Fld<const double> w, int ji, int jk, int jl) {
  R r = ATR(u, ji, jk, jl);
  if (jl > c.G.nflat) r = r + AT(c.g.slope_x, ji, jk, jl) * remap_x(c, VField<R, Fld<const double>>{w}, ji, jk, jl);
  return r;
}
template <class R>
__device__ __forceinline__ R hcov_y(const CtxT<R> &c, Fld<const double> v, Fld<const double> w, int ji, int jk, int jl) {
  R r = ATR(v, ji, jk, jl);
  if (jl > c.G.nflat) r = r + AT(c.g.slope_y, ji, jk, jl) * remap_y(c, VField<R, Fld<const double>>{w}, ji, jk, jl);
  return r;
}

// sum of the densities of the condensates and the gas, sum(rho(:,:,:,1:ncond+1),4)
template <class R>
__device__ __forceinline__ R rho_sum(const CtxT<R> &c, const StateDev &s, ix_t i) {
  R r = R(0.0);
  for (int jc = 0; jc <= c.G.ncond; ++jc) r = r + R(s.rho[jc][i]);
  return r;
}

// c_p_water / c_p_ice / c_p_cond, mo_dictionary.F90:112-189
template <class R>
__device__ __forceinline__ R c_p_water(R temperature) {
  R temp_c = temperature - R(T_0), r;
  if (temp_c >= R(0.0)) {
    if (temp_c > R(35.0)) temp_c = R(35.0);
    R d = temp_c - R(35.0), d2 = d * d, d4 = d2 * d2;
    r = R(0.9979) + R(3.1e-6) * d2 + R(3.8e-9) * d4;
  } else {
    if (temp_c < R(-37.0)) temp_c = R(-37.0);
    R t2 = temp_c * temp_c, t3 = t2 * temp_c, t4 = t2 * t2;
    r = R(1.000938) - R(2.7052e-3) * temp_c - R(2.3235e-5) * t2 + R(4.3778e-6) * t3 + R(2.7136e-7) * t4;
  }
  return R(4186.8) * r;
}
template <class R>
__device__ __forceinline__ R c_p_ice(R temperature) {
  R t = temperature;
  if (t > R(T_0)) t = R(T_0);
  if (t < R(20.0)) t = R(20.0);
  R q = t / R(125.1);
  R r = R(-2.0572) + R(0.14644) * t + R(0.06163) * t * exp(-(q * q));
  return r / R(M_V);
}
template <class R>
__device__ __forceinline__ R c_p_cond(int jc, R temperature) {
  return ((jc - 1) % 2 == 0) ? c_p_ice(temperature) : c_p_water(temperature);
}

// c_v_mass_weighted_air, mo_derived.F90:157-187 (jc is 1-based)
template <class R>
__device__ __forceinline__ R c_v_mass_weighted_air(const CtxT<R> &c, const StateDev &s, R temperature, ix_t i) {
  R r = R(0.0);
  const int ncond = c.G.ncond;
  for (int jc = 1; jc <= ncond; ++jc) r = r + R(s.rho[jc - 1][i]) * c_p_cond(jc, temperature);
  if (c.C.lmoist) {
    r = r + (R(s.rho[ncond][i]) - R(s.rho[ncond + 1][i])) * R(C_D_V);
    r = r + R(s.rho[ncond + 1][i]) * R(C_V_V);
  } else {
    r = R(s.rho[ncond][i]) * R(C_D_V);
  }
  return r;
}

// v_fall_liquid (mo_derived.F90:218-283, with calc_diff_coeff :189-216) of a cloud-water droplet
// of the given radius in layer jl, at the given temperature; FP64 in both forms.  Both branches
// of the Reynolds-number test are evaluated and one is selected, so callers stay branch-free.
template <class R>
__device__ __forceinline__ double v_fall_liquid(const CtxT<R> &c, const StateDev &s, int ji, int jk, int jl, double temperature,
                                                double radius) {
  const int nl = c.G.nl;
  double gl;
  if (jl == 1)
    gl = AT(c.g64.gravity_m_v, ji, jk, 2);
  else if (jl == nl)
    gl = AT(c.g64.gravity_m_v, ji, jk, jl);
  else
    gl = AT(c.g64.ipw[4], ji, jk, jl) * AT(c.g64.gravity_m_v, ji, jk, jl) + AT(c.g64.ipw[5], ji, jk, jl) * AT(c.g64.gravity_m_v, ji, jk, jl + 1);
  const double reynolds_crit = 10.0;
  const double rho = AT(s.rho[c.G.ncond], ji, jk, jl);
  const double particle_radius = 130e-12;
  const double particle_mass = M_D / N_A;
  const double thermal_velocity = sqrt(8.0 * K_B * temperature / (MPI * particle_mass));
  const double particle_density = rho / particle_mass;
  const double cross_section = 4.0 * MPI * (particle_radius * particle_radius);
  const double mean_free_path = 1.0 / (sqrt(2.0) * particle_density * cross_section);
  const double kin = (1.0 / 3.0) * thermal_velocity * mean_free_path;
  const double v = 2.0 * (radius * radius) * (RHO_H2O - rho) * gl / (9.0 * rho * kin);
  const double re = v * radius / kin;
  const double c_w = 0.8;
  const double v_turb = sqrt(8.0 * radius * RHO_H2O * gl / (3.0 * rho * c_w));
  return re > reynolds_crit ? v_turb : v;
}

// radius of the cloud-water droplets whose sink velocity the generalized-density solver uses
constexpr double CLOUD_WATER_RADIUS = 12.e-6;

// inner_product of the wind with itself, mo_inner_product.F90:16-50
template <class R>
__device__ __forceinline__ R ke_point(const CtxT<R> &c, const StateDev &s, int ji, int jk, int jl) {
  const R u1 = ATR(s.wind_u, ji, jk + 1, jl), v0 = ATR(s.wind_v, ji, jk, jl), u0 = ATR(s.wind_u, ji, jk, jl);
  const R v1 = ATR(s.wind_v, ji + 1, jk, jl), w0 = ATR(s.wind_w, ji, jk, jl), w1 = ATR(s.wind_w, ji, jk, jl + 1);
  return AT(c.g.ipw[0], ji, jk, jl) * u1 * u1 + AT(c.g.ipw[1], ji, jk, jl) * v0 * v0 + AT(c.g.ipw[2], ji, jk, jl) * u0 * u0 +
         AT(c.g.ipw[3], ji, jk, jl) * v1 * v1 + AT(c.g.ipw[4], ji, jk, jl) * w0 * w0 + AT(c.g.ipw[5], ji, jk, jl) * w1 * w1;
}

// ---- thread -> grid point maps of the stencil kernels ---------------------------------------
// Every kernel body takes its block's coordinates `bid` as an argument: from blockIdx when the
// kernel runs for one domain, from an ensemble's work list when one launch covers several.
struct Bid {
  int x, y, z;
};
__device__ __forceinline__ Bid bid_here() { return Bid{(int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z}; }

#define KIDX                                                        \
  const int jk = bid.x * blockDim.x + threadIdx.x + 1;              \
  const int ji = bid.y * blockDim.y + threadIdx.y + 1;              \
  const int jl = bid.z + 1;

// Tile order for kernels that read many values of the levels above and below: the kernel is
// launched on a 1-D grid of gx*gy*gz blocks, which the hardware deals out round-robin to the 8
// XCDs.  XCD k takes a contiguous range of tiles, the level running fastest, so the blocks of one
// tile at adjacent levels run one after another on the same XCD and its L2 still holds the
// shared layers.  Kernels that mostly stream level by level keep the x-fastest order, which
// reads DRAM pages in sequence.
__device__ __forceinline__ void tile_zfast(int b, int gx, int gy, int gz, int &bx, int &by, int &bz) {
  const int n = gx * gy * gz;
  const int k = b % N_XCD, s = b / N_XCD, q = n / N_XCD, r = n % N_XCD;
  const int t = k * q + (k < r ? k : r) + s;  // tile index, level fastest
  bz = t % gz;
  const int txy = t / gz;
  bx = txy % gx;
  by = txy / gx;
}
#define KIDX_ZFAST(ex, ey, ez)                                                            \
  int bx_, by_, bz_;                                                                      \
  tile_zfast(bid.x, ((ex) + BX - 1) / BX, ((ey) + (int)blockDim.y - 1) / (int)blockDim.y, (ez), bx_, by_, bz_); \
  const int jk = bx_ * blockDim.x + threadIdx.x + 1;                                      \
  const int ji = by_ * blockDim.y + threadIdx.y + 1;                                      \
  const int jl = bz_ + 1;

// rel_vort_z_local, mo_vorticities.F90:423-538, at any dual point (ji <= ny+1, jk <= nx+1).  The
// dual points of row ny+1 and column nx+1 use their own metrics (rel_vort, :177-184, evaluates
// them; calc_pot_vort then overwrites its eta_z there with row 1 and column 1, :391-417).
template <class R>
__device__ R rel_vort_z(const CtxT<R> &c, const StateDev &s, int ji, int jk, int jl) {
  const int ny = c.G.ny, nx = c.G.nx;
  const int j_i[4] = {ji, wrapm(ji, ny), ji, ji == ny + 1 ? 1 : ji};
  const int j_k[4] = {jk == nx + 1 ? 1 : jk, jk, wrapm(jk, nx), jk};
  const R sgn[4] = {R(1.0), -R(1.0), -R(1.0), R(1.0)};
  R r = R(0.0);
  if (jl <= c.G.nflat) {
    r = r + sgn[0] * AT(c.g.dy, j_i[0], j_k[0], jl) * ATR(s.wind_v, j_i[0], j_k[0], jl) +
        sgn[1] * AT(c.g.dx, j_i[1], j_k[1], jl) * ATR(s.wind_u, j_i[1], j_k[1], jl) +
        sgn[2] * AT(c.g.dy, j_i[2], j_k[2], jl) * ATR(s.wind_v, j_i[2], j_k[2], jl) +
        sgn[3] * AT(c.g.dx, j_i[3], j_k[3], jl) * ATR(s.wind_u, j_i[3], j_k[3], jl);
    return r;
  }
  const R zd = AT(c.g.z_area_dual_z, ji, jk, jl);
#pragma unroll
  for (int jm = 0; jm < 4; ++jm) {
    const bool isv = (jm == 0 || jm == 2);
    const Fld<const R> zf = isv ? c.g.z_v : c.g.z_u;
    const Fld<const double> wf = isv ? Fld<const double>(s.wind_v) : Fld<const double>(s.wind_u);
    const Fld<const R> len = isv ? c.g.dy : c.g.dx;
    const R zp = AT(zf, j_i[jm], j_k[jm], jl);
    const R l_rescale = (R(R_E) + zd) / (R(R_E) + zp);
    const R delta_z = zd - zp;
    int sh = 1;
    if (delta_z > R(0.0) || jl == c.G.nl) sh = -1;
    if (jl == 1) sh = 1;
    const R vg = (ATR(wf, j_i[jm], j_k[jm], jl) - ATR(wf, j_i[jm], j_k[jm], jl + sh)) / (zp - AT(zf, j_i[jm], j_k[jm], jl + sh));
    r = r + sgn[jm] * l_rescale * AT(len, j_i[jm], j_k[jm], jl) * (ATR(wf, j_i[jm], j_k[jm], jl) + delta_z * vg);
  }
  return r;
}

template <class R>
__device__ __forceinline__ R rhog(const CtxT<R> &c, const StateDev &s, int ji, int jk, int jl) {
  return ATR(s.rho[c.G.ncond], ji, jk, jl);
}

// ---- vertical_contravariant_corr of flux fields at the faces of one cell -----------------------
// slopes at the four faces of cell (ji, jk) and its inner-product weights 0..3, layer l
template <class R>
struct VccGeo {
  R sx1, sy0, sx0, sy1, w0, w1, w2, w3;
};
template <class R>
__device__ __forceinline__ VccGeo<R> vcc_geo(const CtxT<R> &c, int ji, int jk, int l) {
  return VccGeo<R>{AT(c.g.slope_x, ji, jk + 1, l), AT(c.g.slope_y, ji, jk, l),    AT(c.g.slope_x, ji, jk, l),
                   AT(c.g.slope_y, ji + 1, jk, l), AT(c.g.ipw[0], ji, jk, l),     AT(c.g.ipw[1], ji, jk, l),
                   AT(c.g.ipw[2], ji, jk, l),      AT(c.g.ipw[3], ji, jk, l)};
}

// a vector field at the four faces of cell (ji, jk): x at jk+1, y at ji, x at jk, y at ji+1
template <class R>
struct Faces {
  R e, n, w, s;
};
template <class R, class P>
__device__ __forceinline__ Faces<R> faces(P vx, P vy, const Geom &G, int ji, int jk, int l) {
  return Faces<R>{R(vx[LG_IX(G, ji, jk + 1, l)]), R(vy[LG_IX(G, ji, jk, l)]), R(vx[LG_IX(G, ji, jk, l)]),
                  R(vy[LG_IX(G, ji + 1, jk, l)])};
}

// r minus the four terms of one layer of vertical_contravariant_corr (mo_averaging.F90:29-50)
template <class R>
__device__ __forceinline__ R vcc_sub(R r, const VccGeo<R> &g, const Faces<R> &f) {
  return r - R(0.5) * f.e * g.sx1 * g.w0 - R(0.5) * f.n * g.sy0 * g.w1 - R(0.5) * f.w * g.sx0 * g.w2 - R(0.5) * f.s * g.sy1 * g.w3;
}

}  // namespace lgame
