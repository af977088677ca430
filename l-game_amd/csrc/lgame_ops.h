// lgame_ops.h -- per-point device functions of the HEVI step operators (FP64).
//
// Each function returns the value the reference operator stores at one grid point, with the
// same floating-point operation order as the reference Fortran (the library is built with
// -ffp-contract=off, like the reference build).  Periodic boundaries: neighbour indices wrap;
// the duplicate edge points (x index nx+1, y index ny+1) follow the reference's semantics for
// each operator (a copy of index 1, or their own metrics where the reference computes them).
#pragma once

#include "lgame_dev.h"

namespace lgame {

#define AT(p, ji, jk, jl) (p)[LG_IX(c.G, ji, jk, jl)]

__device__ __forceinline__ int wrapm(int i, int n) { return i == 1 ? n : i - 1; }

// ---- vertical fields seen through remap_ver2hor -----------------------------------------
// grad_vert of a scalar field (mo_gradient_operators.F90:65-83); the reference leaves levels 1
// and n_levels of its result untouched and they hold zero whenever the remap reads them.
struct VGrad {
  const double *s;
  __device__ double operator()(const Ctx &c, int ji, int jk, int jl) const {
    if (jl >= c.G.nlev || jl <= 1) return 0.0;
    return (AT(s, ji, jk, jl - 1) - AT(s, ji, jk, jl)) / AT(c.g.dz, ji, jk, jl);
  }
};
struct VField {
  const double *w;
  __device__ double operator()(const Ctx &c, int ji, int jk, int jl) const { return AT(w, ji, jk, jl); }
};

// remap_ver2hor_x, mo_averaging.F90:56-95 (periodic)
template <class VZ>
__device__ double remap_x(const Ctx &c, const VZ &vz, int ji, int jk, int jl) {
  const int nx = c.G.nx;
  int ka, kb;
  if (jk == 1 || jk == nx + 1) {
    ka = nx;
    kb = 1;
  } else {
    ka = jk - 1;
    kb = jk;
  }
  double r = 0.0;
  r = r + AT(c.g.ipw[4], ji, ka, jl) * vz(c, ji, ka, jl);
  r = r + AT(c.g.ipw[4], ji, kb, jl) * vz(c, ji, kb, jl);
  if (jl < c.G.nl || c.C.lfreeslip) {
    r = r + AT(c.g.ipw[5], ji, ka, jl) * vz(c, ji, ka, jl + 1);
    r = r + AT(c.g.ipw[5], ji, kb, jl) * vz(c, ji, kb, jl + 1);
  }
  return 0.5 * r;
}

// remap_ver2hor_y, mo_averaging.F90:97-136 (periodic)
template <class VZ>
__device__ double remap_y(const Ctx &c, const VZ &vz, int ji, int jk, int jl) {
  const int ny = c.G.ny;
  int ia, ib;
  if (ji == 1 || ji == ny + 1) {
    ia = ny;
    ib = 1;
  } else {
    ia = ji - 1;
    ib = ji;
  }
  double r = 0.0;
  r = r + AT(c.g.ipw[4], ia, jk, jl) * vz(c, ia, jk, jl);
  r = r + AT(c.g.ipw[4], ib, jk, jl) * vz(c, ib, jk, jl);
  if (jl < c.G.nl || c.C.lfreeslip) {
    r = r + AT(c.g.ipw[5], ia, jk, jl) * vz(c, ia, jk, jl + 1);
    r = r + AT(c.g.ipw[5], ib, jk, jl) * vz(c, ib, jk, jl + 1);
  }
  return 0.5 * r;
}

// grad_hor x-component (mo_gradient_operators.F90:17-63,85-130); at jk = nx+1 the covariant part
// is the copy of jk = 1 and the terrain correction uses the metrics of jk = nx+1.
__device__ __forceinline__ double grad_hor_x(const Ctx &c, const double *s, int ji, int jk, int jl) {
  const int nx = c.G.nx;
  double r;
  if (jk == 1 || jk == nx + 1)
    r = (AT(s, ji, 1, jl) - AT(s, ji, nx, jl)) / AT(c.g.dx, ji, 1, jl);
  else
    r = (AT(s, ji, jk, jl) - AT(s, ji, jk - 1, jl)) / AT(c.g.dx, ji, jk, jl);
  if (jl > c.G.nflat) r = r - AT(c.g.slope_x, ji, jk, jl) * remap_x(c, VGrad{s}, ji, jk, jl);
  return r;
}

__device__ __forceinline__ double grad_hor_y(const Ctx &c, const double *s, int ji, int jk, int jl) {
  const int ny = c.G.ny;
  double r;
  if (ji == 1 || ji == ny + 1)
    r = (AT(s, ny, jk, jl) - AT(s, 1, jk, jl)) / AT(c.g.dy, 1, jk, jl);
  else
    r = (AT(s, ji - 1, jk, jl) - AT(s, ji, jk, jl)) / AT(c.g.dy, ji, jk, jl);
  if (jl > c.G.nflat) r = r - AT(c.g.slope_y, ji, jk, jl) * remap_y(c, VGrad{s}, ji, jk, jl);
  return r;
}

// vertical_contravariant_corr, mo_averaging.F90:16-54
__device__ double vcc(const Ctx &c, const double *vx, const double *vy, int ji, int jk, int jl) {
  double r = 0.0;
  if (jl >= c.G.nflat + 1) {
    if (jl != c.G.nflat + 1) {
      const int l = jl - 1;
      r = r - 0.5 * AT(vx, ji, jk + 1, l) * AT(c.g.slope_x, ji, jk + 1, l) * AT(c.g.ipw[0], ji, jk, l) -
          0.5 * AT(vy, ji, jk, l) * AT(c.g.slope_y, ji, jk, l) * AT(c.g.ipw[1], ji, jk, l) -
          0.5 * AT(vx, ji, jk, l) * AT(c.g.slope_x, ji, jk, l) * AT(c.g.ipw[2], ji, jk, l) -
          0.5 * AT(vy, ji + 1, jk, l) * AT(c.g.slope_y, ji + 1, jk, l) * AT(c.g.ipw[3], ji, jk, l);
    }
    r = r - 0.5 * AT(vx, ji, jk + 1, jl) * AT(c.g.slope_x, ji, jk + 1, jl) * AT(c.g.ipw[0], ji, jk, jl) -
        0.5 * AT(vy, ji, jk, jl) * AT(c.g.slope_y, ji, jk, jl) * AT(c.g.ipw[1], ji, jk, jl) -
        0.5 * AT(vx, ji, jk, jl) * AT(c.g.slope_x, ji, jk, jl) * AT(c.g.ipw[2], ji, jk, jl) -
        0.5 * AT(vy, ji + 1, jk, jl) * AT(c.g.slope_y, ji + 1, jk, jl) * AT(c.g.ipw[3], ji, jk, jl);
  }
  return r;
}

// div_h, mo_divergence_operators.F90:16-68
__device__ double div_h(const Ctx &c, const double *vx, const double *vy, int ji, int jk, int jl) {
  double comp_h = AT(vx, ji, jk + 1, jl) * AT(c.g.area_x, ji, jk + 1, jl) + AT(vy, ji, jk, jl) * AT(c.g.area_y, ji, jk, jl) -
                  AT(vx, ji, jk, jl) * AT(c.g.area_x, ji, jk, jl) - AT(vy, ji + 1, jk, jl) * AT(c.g.area_y, ji + 1, jk, jl);
  double comp_v = 0.0;
  const int nflat = c.G.nflat, nl = c.G.nl;
  if (jl == nflat) {
    double cl = vcc(c, vx, vy, ji, jk, jl + 1);
    comp_v = -cl * AT(c.g.area_z, ji, jk, jl + 1);
  } else if (jl == nl) {
    double cu = vcc(c, vx, vy, ji, jk, jl);
    comp_v = cu * AT(c.g.area_z, ji, jk, jl);
  } else if (jl > nflat) {
    double cu = vcc(c, vx, vy, ji, jk, jl);
    double cl = vcc(c, vx, vy, ji, jk, jl + 1);
    comp_v = cu * AT(c.g.area_z, ji, jk, jl) - cl * AT(c.g.area_z, ji, jk, jl + 1);
  }
  return (comp_h + comp_v) / AT(c.g.volume, ji, jk, jl);
}

// horizontal_covariant_x/_y, mo_averaging.F90:138-176 (vertical component = wind_w)
__device__ __forceinline__ double hcov_x(const Ctx &c, const double *u, const double *w, int ji, int jk, int jl) {
  double r = AT(u, ji, jk, jl);
  if (jl > c.G.nflat) r = r + AT(c.g.slope_x, ji, jk, jl) * remap_x(c, VField{w}, ji, jk, jl);
  return r;
}
__device__ __forceinline__ double hcov_y(const Ctx &c, const double *v, const double *w, int ji, int jk, int jl) {
  double r = AT(v, ji, jk, jl);
  if (jl > c.G.nflat) r = r + AT(c.g.slope_y, ji, jk, jl) * remap_y(c, VField{w}, ji, jk, jl);
  return r;
}

// sum of the densities of the condensates and the gas, sum(rho(:,:,:,1:ncond+1),4)
__device__ __forceinline__ double rho_sum(const Ctx &c, const StateDev &s, size_t i) {
  double r = 0.0;
  for (int jc = 0; jc <= c.G.ncond; ++jc) r = r + s.rho[jc][i];
  return r;
}

// c_p_water / c_p_ice / c_p_cond, mo_dictionary.F90:112-189
__device__ __forceinline__ double c_p_water(double temperature) {
  double temp_c = temperature - T_0, r;
  if (temp_c >= 0.0) {
    if (temp_c > 35.0) temp_c = 35.0;
    double d = temp_c - 35.0, d2 = d * d, d4 = d2 * d2;
    r = 0.9979 + 3.1e-6 * d2 + 3.8e-9 * d4;
  } else {
    if (temp_c < -37.0) temp_c = -37.0;
    double t2 = temp_c * temp_c, t3 = t2 * temp_c, t4 = t2 * t2;
    r = 1.000938 - 2.7052e-3 * temp_c - 2.3235e-5 * t2 + 4.3778e-6 * t3 + 2.7136e-7 * t4;
  }
  return 4186.8 * r;
}
__device__ __forceinline__ double c_p_ice(double temperature) {
  double t = temperature;
  if (t > T_0) t = T_0;
  if (t < 20.0) t = 20.0;
  double q = t / 125.1;
  double r = -2.0572 + 0.14644 * t + 0.06163 * t * exp(-(q * q));
  return r / M_V;
}
__device__ __forceinline__ double c_p_cond(int jc, double temperature) {
  return ((jc - 1) % 2 == 0) ? c_p_ice(temperature) : c_p_water(temperature);
}

// c_v_mass_weighted_air, mo_derived.F90:157-187 (jc is 1-based)
__device__ __forceinline__ double c_v_mass_weighted_air(const Ctx &c, const StateDev &s, double temperature, size_t i) {
  double r = 0.0;
  const int ncond = c.G.ncond;
  for (int jc = 1; jc <= ncond; ++jc) r = r + s.rho[jc - 1][i] * c_p_cond(jc, temperature);
  if (c.C.lmoist) {
    r = r + (s.rho[ncond][i] - s.rho[ncond + 1][i]) * C_D_V;
    r = r + s.rho[ncond + 1][i] * C_V_V;
  } else {
    r = s.rho[ncond][i] * C_D_V;
  }
  return r;
}

// inner_product of the wind with itself, mo_inner_product.F90:16-50
__device__ __forceinline__ double ke_point(const Ctx &c, const StateDev &s, int ji, int jk, int jl) {
  const double u1 = AT(s.wind_u, ji, jk + 1, jl), v0 = AT(s.wind_v, ji, jk, jl), u0 = AT(s.wind_u, ji, jk, jl);
  const double v1 = AT(s.wind_v, ji + 1, jk, jl), w0 = AT(s.wind_w, ji, jk, jl), w1 = AT(s.wind_w, ji, jk, jl + 1);
  return AT(c.g.ipw[0], ji, jk, jl) * u1 * u1 + AT(c.g.ipw[1], ji, jk, jl) * v0 * v0 + AT(c.g.ipw[2], ji, jk, jl) * u0 * u0 +
         AT(c.g.ipw[3], ji, jk, jl) * v1 * v1 + AT(c.g.ipw[4], ji, jk, jl) * w0 * w0 + AT(c.g.ipw[5], ji, jk, jl) * w1 * w1;
}

}  // namespace lgame
