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

// ---- two x-adjacent points per thread -------------------------------------------------------
// A stencil can be evaluated at the columns k and k+1 of one row at once (k odd, k+1 <= nx).  Its
// values are then pairs: the loads of the own columns are single 2-element vector loads (aligned,
// since k-1 is even and rows, levels and fields start at multiples of 16 elements), and the west
// neighbours of the pair are the columns (west of k, k), the east ones (k+1, k+2), each loaded per
// element.  Every operation on a pair is the same operation on each element, so each
// element is bitwise the value of the single-point evaluation.  The helpers below take the x index
// as a template parameter X: int (one point), XP (the columns k, k+1) or XG (one column per
// element); VT<R, X> is the matching value type.
template <class R>
struct P2 {
  R a, b;
  P2() = default;
  __device__ __forceinline__ P2(R x) : a(x), b(x) {}
  __device__ __forceinline__ P2(R x, R y) : a(x), b(y) {}
  friend __device__ __forceinline__ P2 operator+(P2 x, P2 y) { return P2(x.a + y.a, x.b + y.b); }
  friend __device__ __forceinline__ P2 operator-(P2 x, P2 y) { return P2(x.a - y.a, x.b - y.b); }
  friend __device__ __forceinline__ P2 operator*(P2 x, P2 y) { return P2(x.a * y.a, x.b * y.b); }
  friend __device__ __forceinline__ P2 operator/(P2 x, P2 y) { return P2(x.a / y.a, x.b / y.b); }
  friend __device__ __forceinline__ P2 operator-(P2 x) { return P2(-x.a, -x.b); }
};
struct XP {
  int k;  // the columns k and k+1
};
struct XG {
  int k0, k1;  // the column of each element
};
struct LP {
  int a, b;  // the level of each element
};
template <class R, class X>
struct ValOf {
  using type = P2<R>;
};
template <class R>
struct ValOf<R, int> {
  using type = R;
};
template <class R, class X>
using VT = typename ValOf<R, X>::type;

// west neighbours: k+1 >= 2 never wraps
__device__ __forceinline__ XG wrapm(XP x, int n) { return XG{wrapm(x.k, n), x.k}; }
__device__ __forceinline__ XG wrapm(XG x, int n) { return XG{wrapm(x.k0, n), wrapm(x.k1, n)}; }
// the value of one cell, or of the pair's first cell
template <class R>
__device__ __forceinline__ R first(R v) {
  return v;
}
template <class R>
__device__ __forceinline__ R first(P2<R> v) {
  return v.a;
}
// the cell (or the pair's first cell) is in column 1
__device__ __forceinline__ bool col1(int k) { return k == 1; }
__device__ __forceinline__ bool col1(XP x) { return x.k == 1; }
// east neighbours: k+2 <= nx+1
__device__ __forceinline__ int east(int k) { return k + 1; }
__device__ __forceinline__ XG east(XP x) { return XG{x.k + 1, x.k + 2}; }
// column nx+1 read as column 1: k+1 <= nx
__device__ __forceinline__ int edge1(int k, int n) { return k == n + 1 ? 1 : k; }
__device__ __forceinline__ XP edge1(XP x, int) { return x; }
__device__ __forceinline__ XG edge1(XG x, int n) { return XG{edge1(x.k0, n), edge1(x.k1, n)}; }

// loads of a field at one point, at a pair of columns, or at one column per element
template <class T>
__device__ __forceinline__ typename std::remove_cv<T>::type ldv(Fld<T> p, const Geom &G, int ji, int k, int jl) {
  return p[LG_IX(G, ji, k, jl)];
}
template <class T>
__device__ __forceinline__ P2<typename std::remove_cv<T>::type> ldv(Fld<T> p, const Geom &G, int ji, XP x, int jl) {
  using E = typename std::remove_cv<T>::type;
  using E2 = E __attribute__((ext_vector_type(2)));
  const E2 v = *reinterpret_cast<const E2 *>(&p[LG_IX(G, ji, x.k, jl)]);
  return P2<E>(v.x, v.y);
}
template <class T>
__device__ __forceinline__ P2<typename std::remove_cv<T>::type> ldv(Fld<T> p, const Geom &G, int ji, XG x, int jl) {
  using E = typename std::remove_cv<T>::type;
  return P2<E>(p[LG_IX(G, ji, x.k0, jl)], p[LG_IX(G, ji, x.k1, jl)]);
}
// theta_v = theta_v_bg + theta_v_pert of a state, formed where it is loaded (the full theta_v of
// theta_v_adv_3rd_order, mo_multiplications.F90:116-204, without a stored copy)
template <class R>
struct ThetaFull {
  Fld<const R> bg;
  Fld<const double> pert;
};
template <class R>
__device__ __forceinline__ R ldv(ThetaFull<R> p, const Geom &G, int ji, int k, int jl) {
  const ix_t i = LG_IX(G, ji, k, jl);
  return p.bg[i] + R(p.pert[i]);
}
// a load at one level per element
template <class T>
__device__ __forceinline__ typename std::remove_cv<T>::type ldv_l(Fld<T> p, const Geom &G, int ji, int k, int jl) {
  return p[LG_IX(G, ji, k, jl)];
}
template <class T>
__device__ __forceinline__ P2<typename std::remove_cv<T>::type> ldv_l(Fld<T> p, const Geom &G, int ji, XP x, LP l) {
  using E = typename std::remove_cv<T>::type;
  return P2<E>(p[LG_IX(G, ji, x.k, l.a)], p[LG_IX(G, ji, x.k + 1, l.b)]);
}
template <class T>
__device__ __forceinline__ P2<typename std::remove_cv<T>::type> ldv_l(Fld<T> p, const Geom &G, int ji, XG x, LP l) {
  using E = typename std::remove_cv<T>::type;
  return P2<E>(p[LG_IX(G, ji, x.k0, l.a)], p[LG_IX(G, ji, x.k1, l.b)]);
}
template <class R, class E>
__device__ __forceinline__ R cvt(E x) {
  return static_cast<R>(x);
}
template <class R, class E>
__device__ __forceinline__ P2<R> cvt(P2<E> x) {
  return P2<R>(static_cast<R>(x.a), static_cast<R>(x.b));
}
// a store at one point or at a pair of columns
template <class T>
__device__ __forceinline__ void stv(Fld<T> p, const Geom &G, int ji, int k, int jl, T v) {
  p[LG_IX(G, ji, k, jl)] = v;
}
template <class T>
__device__ __forceinline__ void stv(Fld<T> p, const Geom &G, int ji, XP x, int jl, P2<T> v) {
  using E2 = T __attribute__((ext_vector_type(2)));
  E2 w;
  w.x = v.a;
  w.y = v.b;
  *reinterpret_cast<E2 *>(&p[LG_IX(G, ji, x.k, jl)]) = w;
}
// generic forms of AT and ATR for any x index type
#define LDG(p, ji, jk, jl) ldv(p, c.G, ji, jk, jl)
#define LDGR(p, ji, jk, jl) cvt<R>(ldv(p, c.G, ji, jk, jl))

// ---- vertical fields seen through remap_ver2hor -----------------------------------------
// grad_vert of a scalar field (mo_gradient_operators.F90:65-83); the reference leaves levels 1
// and n_levels of its result untouched and they hold zero whenever the remap reads them.
template <class R, class P>
struct VGrad {
  P s;
  template <class X>
  __device__ VT<R, X> operator()(const CtxT<R> &c, int ji, X jk, int jl) const {
    if (jl >= c.G.nlev || jl <= 1) return R(0.0);
    return (LDGR(s, ji, jk, jl - 1) - LDGR(s, ji, jk, jl)) / LDG(c.g.dz, ji, jk, jl);
  }
};
template <class R, class P>
struct VField {
  P w;
  template <class X>
  __device__ VT<R, X> operator()(const CtxT<R> &c, int ji, X jk, int jl) const {
    return LDGR(w, ji, jk, jl);
  }
};

// the two columns remap_ver2hor_x averages over: (nx, 1) at columns 1 and nx+1, else (jk-1, jk)
__device__ __forceinline__ int remap_ka(int jk, int nx) { return (jk == 1 || jk == nx + 1) ? nx : jk - 1; }
__device__ __forceinline__ int remap_kb(int jk, int nx) { return (jk == 1 || jk == nx + 1) ? 1 : jk; }
__device__ __forceinline__ XG remap_ka(XP x, int nx) { return XG{remap_ka(x.k, nx), x.k}; }  // k+1 in 2..nx
__device__ __forceinline__ XP remap_kb(XP x, int) { return x; }
__device__ __forceinline__ XG remap_ka(XG x, int nx) { return XG{remap_ka(x.k0, nx), remap_ka(x.k1, nx)}; }
__device__ __forceinline__ XG remap_kb(XG x, int nx) { return XG{remap_kb(x.k0, nx), remap_kb(x.k1, nx)}; }

// remap_ver2hor_x, mo_averaging.F90:56-95 (periodic)
template <class R, class VZ, class X>
__device__ VT<R, X> remap_x(const CtxT<R> &c, const VZ &vz, int ji, X jk, int jl) {
  const int nx = c.G.nx;
  const auto ka = remap_ka(jk, nx);
  const auto kb = remap_kb(jk, nx);
  VT<R, X> r = R(0.0);
  r = r + LDG(c.g.ipw[4], ji, ka, jl) * vz(c, ji, ka, jl);
  r = r + LDG(c.g.ipw[4], ji, kb, jl) * vz(c, ji, kb, jl);
  if (jl < c.G.nl || c.C.lfreeslip) {
    r = r + LDG(c.g.ipw[5], ji, ka, jl) * vz(c, ji, ka, jl + 1);
    r = r + LDG(c.g.ipw[5], ji, kb, jl) * vz(c, ji, kb, jl + 1);
  }
  return R(0.5) * r;
}

// remap_ver2hor_y, mo_averaging.F90:97-136 (periodic)
template <class R, class VZ, class X>
__device__ VT<R, X> remap_y(const CtxT<R> &c, const VZ &vz, int ji, X jk, int jl) {
  const int ny = c.G.ny;
  int ia, ib;
  if (ji == 1 || ji == ny + 1) {
    ia = ny;
    ib = 1;
  } else {
    ia = ji - 1;
    ib = ji;
  }
  VT<R, X> r = R(0.0);
  r = r + LDG(c.g.ipw[4], ia, jk, jl) * vz(c, ia, jk, jl);
  r = r + LDG(c.g.ipw[4], ib, jk, jl) * vz(c, ib, jk, jl);
  if (jl < c.G.nl || c.C.lfreeslip) {
    r = r + LDG(c.g.ipw[5], ia, jk, jl) * vz(c, ia, jk, jl + 1);
    r = r + LDG(c.g.ipw[5], ib, jk, jl) * vz(c, ib, jk, jl + 1);
  }
  return R(0.5) * r;
}

// grad_hor x-component (mo_gradient_operators.F90:17-63,85-130); at jk = nx+1 the covariant part
// is the copy of jk = 1 and the terrain correction uses the metrics of jk = nx+1.
// (columns 1 and nx+1 take (s(1) - s(nx)) / dx(1): edge1 gives column 1 there, wrapm column nx)
template <class R, class P, class X>
__device__ __forceinline__ VT<R, X> grad_hor_x(const CtxT<R> &c, P s, int ji, X jk, int jl) {
  const int nx = c.G.nx;
  const auto ka = edge1(jk, nx);
  const auto kb = wrapm(jk, nx);
  VT<R, X> r = (LDGR(s, ji, ka, jl) - LDGR(s, ji, kb, jl)) / LDG(c.g.dx, ji, ka, jl);
  if (jl > c.G.nflat) r = r - LDG(c.g.slope_x, ji, jk, jl) * remap_x(c, VGrad<R, P>{s}, ji, jk, jl);
  return r;
}

template <class R, class P, class X>
__device__ __forceinline__ VT<R, X> grad_hor_y(const CtxT<R> &c, P s, int ji, X jk, int jl) {
  const int ny = c.G.ny;
  VT<R, X> r;
  if (ji == 1 || ji == ny + 1)
    r = (LDGR(s, ny, jk, jl) - LDGR(s, 1, jk, jl)) / LDG(c.g.dy, 1, jk, jl);
  else
    r = (LDGR(s, ji - 1, jk, jl) - LDGR(s, ji, jk, jl)) / LDG(c.g.dy, ji, jk, jl);
  if (jl > c.G.nflat) r = r - LDG(c.g.slope_y, ji, jk, jl) * remap_y(c, VGrad<R, P>{s}, ji, jk, jl);
  return r;
}

// vertical_contravariant_corr, mo_averaging.F90:16-54
template <class R, class P, class X>
__device__ VT<R, X> vcc(const CtxT<R> &c, P vx, P vy, int ji, X jk, int jl) {
  const auto ke = east(jk);
  VT<R, X> r = R(0.0);
  if (jl >= c.G.nflat + 1) {
    if (jl != c.G.nflat + 1) {
      const int l = jl - 1;
      r = r - R(0.5) * LDGR(vx, ji, ke, l) * LDG(c.g.slope_x, ji, ke, l) * LDG(c.g.ipw[0], ji, jk, l) -
          R(0.5) * LDGR(vy, ji, jk, l) * LDG(c.g.slope_y, ji, jk, l) * LDG(c.g.ipw[1], ji, jk, l) -
          R(0.5) * LDGR(vx, ji, jk, l) * LDG(c.g.slope_x, ji, jk, l) * LDG(c.g.ipw[2], ji, jk, l) -
          R(0.5) * LDGR(vy, ji + 1, jk, l) * LDG(c.g.slope_y, ji + 1, jk, l) * LDG(c.g.ipw[3], ji, jk, l);
    }
    r = r - R(0.5) * LDGR(vx, ji, ke, jl) * LDG(c.g.slope_x, ji, ke, jl) * LDG(c.g.ipw[0], ji, jk, jl) -
        R(0.5) * LDGR(vy, ji, jk, jl) * LDG(c.g.slope_y, ji, jk, jl) * LDG(c.g.ipw[1], ji, jk, jl) -
        R(0.5) * LDGR(vx, ji, jk, jl) * LDG(c.g.slope_x, ji, jk, jl) * LDG(c.g.ipw[2], ji, jk, jl) -
        R(0.5) * LDGR(vy, ji + 1, jk, jl) * LDG(c.g.slope_y, ji + 1, jk, jl) * LDG(c.g.ipw[3], ji, jk, jl);
  }
  return r;
}

// div_h, mo_divergence_operators.F90:16-68
template <class R, class P, class X>
__device__ VT<R, X> div_h(const CtxT<R> &c, P vx, P vy, int ji, X jk, int jl) {
  using V = VT<R, X>;
  const auto ke = east(jk);
  V comp_h = LDGR(vx, ji, ke, jl) * LDG(c.g.area_x, ji, ke, jl) + LDGR(vy, ji, jk, jl) * LDG(c.g.area_y, ji, jk, jl) -
             LDGR(vx, ji, jk, jl) * LDG(c.g.area_x, ji, jk, jl) - LDGR(vy, ji + 1, jk, jl) * LDG(c.g.area_y, ji + 1, jk, jl);
  V comp_v = R(0.0);
  const int nflat = c.G.nflat, nl = c.G.nl;
  if (jl == nflat) {
    V cl = vcc(c, vx, vy, ji, jk, jl + 1);
    comp_v = -cl * LDG(c.g.area_z, ji, jk, jl + 1);
  } else if (jl == nl) {
    V cu = vcc(c, vx, vy, ji, jk, jl);
    comp_v = cu * LDG(c.g.area_z, ji, jk, jl);
  } else if (jl > nflat) {
    V cu = vcc(c, vx, vy, ji, jk, jl);
    V cl = vcc(c, vx, vy, ji, jk, jl + 1);
    comp_v = cu * LDG(c.g.area_z, ji, jk, jl) - cl * LDG(c.g.area_z, ji, jk, jl + 1);
  }
  return (comp_h + comp_v) / LDG(c.g.volume, ji, jk, jl);
}

// horizontal_covariant_x/_y, mo_averaging.F90:138-176 (vertical component = wind_w)
template <class R, class X>
__device__ __forceinline__ VT<R, X> hcov_x(const CtxT<R> &c, Fld<const double> u, Fld<const double> w, int ji, X jk, int jl) {
  VT<R, X> r = LDGR(u, ji, jk, jl);
  if (jl > c.G.nflat) r = r + LDG(c.g.slope_x, ji, jk, jl) * remap_x(c, VField<R, Fld<const double>>{w}, ji, jk, jl);
  return r;
}
template <class R, class X>
__device__ __forceinline__ VT<R, X> hcov_y(const CtxT<R> &c, Fld<const double> v, Fld<const double> w, int ji, X jk, int jl) {
  VT<R, X> r = LDGR(v, ji, jk, jl);
  if (jl > c.G.nflat) r = r + LDG(c.g.slope_y, ji, jk, jl) * remap_y(c, VField<R, Fld<const double>>{w}, ji, jk, jl);
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

// inner_product of the wind with itself, mo_inner_product.F90:16-50, at one cell or a column pair
template <class R, class X>
__device__ __forceinline__ VT<R, X> ke_point(const CtxT<R> &c, const StateDev &s, int ji, X jk, int jl) {
  using V = VT<R, X>;
  const auto ke = east(jk);
  const V u1 = LDGR(s.wind_u, ji, ke, jl), v0 = LDGR(s.wind_v, ji, jk, jl), u0 = LDGR(s.wind_u, ji, jk, jl);
  const V v1 = LDGR(s.wind_v, ji + 1, jk, jl), w0 = LDGR(s.wind_w, ji, jk, jl), w1 = LDGR(s.wind_w, ji, jk, jl + 1);
  return LDG(c.g.ipw[0], ji, jk, jl) * u1 * u1 + LDG(c.g.ipw[1], ji, jk, jl) * v0 * v0 + LDG(c.g.ipw[2], ji, jk, jl) * u0 * u0 +
         LDG(c.g.ipw[3], ji, jk, jl) * v1 * v1 + LDG(c.g.ipw[4], ji, jk, jl) * w0 * w0 + LDG(c.g.ipw[5], ji, jk, jl) * w1 * w1;
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
// rel_vort_z_local's level for the vertical gradient of one term: below (jl-1) when the dual
// point lies above the term's point or at the lowest layer, above (jl+1) at the top layer and
// otherwise
template <class R>
__device__ __forceinline__ int rv_level(R delta_z, int jl, int nl) {
  int sh = 1;
  if (delta_z > R(0.0) || jl == nl) sh = -1;
  if (jl == 1) sh = 1;
  return jl + sh;
}
template <class R>
__device__ __forceinline__ LP rv_level(P2<R> delta_z, int jl, int nl) {
  return LP{rv_level(delta_z.a, jl, nl), rv_level(delta_z.b, jl, nl)};
}
// one term of rel_vort_z_local above the flat levels: sgn * l_rescale * len * (wind + delta_z * vg)
template <class R, class X>
__device__ __forceinline__ VT<R, X> rv_term(const CtxT<R> &c, VT<R, X> zd, R sgn, Fld<const R> zf, Fld<const double> wf,
                                             Fld<const R> len, int jj, X kk, int jl) {
  using V = VT<R, X>;
  const V zp = LDG(zf, jj, kk, jl);
  const V l_rescale = (R(R_E) + zd) / (R(R_E) + zp);
  const V delta_z = zd - zp;
  const auto lv = rv_level(delta_z, jl, c.G.nl);
  const V vg = (LDGR(wf, jj, kk, jl) - cvt<R>(ldv_l(wf, c.G, jj, kk, lv))) / (zp - ldv_l(zf, c.G, jj, kk, lv));
  return sgn * l_rescale * LDG(len, jj, kk, jl) * (LDGR(wf, jj, kk, jl) + delta_z * vg);
}
template <class R, class X>
__device__ VT<R, X> rel_vort_z(const CtxT<R> &c, const StateDev &s, int ji, X jk, int jl) {
  using V = VT<R, X>;
  const int ny = c.G.ny, nx = c.G.nx;
  // the four faces of the dual cell: (row, column, sign), v at faces 0 and 2, u at 1 and 3
  const int i0 = ji, i1 = wrapm(ji, ny), i2 = ji, i3 = ji == ny + 1 ? 1 : ji;
  const auto k0 = edge1(jk, nx);
  const X k1 = jk, k3 = jk;
  const auto k2 = wrapm(jk, nx);
  V r = R(0.0);
  if (jl <= c.G.nflat) {
    r = r + R(1.0) * LDG(c.g.dy, i0, k0, jl) * LDGR(s.wind_v, i0, k0, jl) +
        -R(1.0) * LDG(c.g.dx, i1, k1, jl) * LDGR(s.wind_u, i1, k1, jl) +
        -R(1.0) * LDG(c.g.dy, i2, k2, jl) * LDGR(s.wind_v, i2, k2, jl) +
        R(1.0) * LDG(c.g.dx, i3, k3, jl) * LDGR(s.wind_u, i3, k3, jl);
    return r;
  }
  const V zd = LDG(c.g.z_area_dual_z, ji, jk, jl);
  r = r + rv_term<R>(c, zd, R(1.0), c.g.z_v, s.wind_v, c.g.dy, i0, k0, jl);
  r = r + rv_term<R>(c, zd, -R(1.0), c.g.z_u, s.wind_u, c.g.dx, i1, k1, jl);
  r = r + rv_term<R>(c, zd, -R(1.0), c.g.z_v, s.wind_v, c.g.dy, i2, k2, jl);
  r = r + rv_term<R>(c, zd, R(1.0), c.g.z_u, s.wind_u, c.g.dx, i3, k3, jl);
  return r;
}

template <class R, class X>
__device__ __forceinline__ VT<R, X> rhog(const CtxT<R> &c, const StateDev &s, int ji, X jk, int jl) {
  return LDGR(s.rho[c.G.ncond], ji, jk, jl);
}

// ---- vertical_contravariant_corr of flux fields at the faces of one cell -----------------------
// slopes at the four faces of cell (ji, jk) and its inner-product weights 0..3, layer l
template <class V>
struct VccGeo {
  V sx1, sy0, sx0, sy1, w0, w1, w2, w3;
};
template <class R, class X>
__device__ __forceinline__ VccGeo<VT<R, X>> vcc_geo(const CtxT<R> &c, int ji, X jk, int l) {
  const auto ke = east(jk);
  return VccGeo<VT<R, X>>{LDG(c.g.slope_x, ji, ke, l), LDG(c.g.slope_y, ji, jk, l),     LDG(c.g.slope_x, ji, jk, l),
                          LDG(c.g.slope_y, ji + 1, jk, l), LDG(c.g.ipw[0], ji, jk, l), LDG(c.g.ipw[1], ji, jk, l),
                          LDG(c.g.ipw[2], ji, jk, l),      LDG(c.g.ipw[3], ji, jk, l)};
}

// a vector field at the four faces of cell (ji, jk): x at jk+1, y at ji, x at jk, y at ji+1
template <class V>
struct Faces {
  V e, n, w, s;
};
template <class R, class P, class X>
__device__ __forceinline__ Faces<VT<R, X>> faces(P vx, P vy, const Geom &G, int ji, X jk, int l) {
  return Faces<VT<R, X>>{cvt<R>(ldv(vx, G, ji, east(jk), l)), cvt<R>(ldv(vy, G, ji, jk, l)), cvt<R>(ldv(vx, G, ji, jk, l)),
                         cvt<R>(ldv(vy, G, ji + 1, jk, l))};
}

// the element type of a value or a pair
template <class V>
struct ElemOf {
  using type = V;
};
template <class R>
struct ElemOf<P2<R>> {
  using type = R;
};

// r minus the four terms of one layer of vertical_contravariant_corr (mo_averaging.F90:29-50)
template <class V>
__device__ __forceinline__ V vcc_sub(V r, const VccGeo<V> &g, const Faces<V> &f) {
  using E = typename ElemOf<V>::type;
  return r - E(0.5) * f.e * g.sx1 * g.w0 - E(0.5) * f.n * g.sy0 * g.w1 - E(0.5) * f.w * g.sx0 * g.w2 - E(0.5) * f.s * g.sy1 * g.w3;
}

}  // namespace lgame
