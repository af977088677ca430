// lgame_grid.hip -- grid and background-state generation on the device (gfx950).
//
// Restates grid_setup and bg_setup (src/io/mo_grid_generator.F90:29-1457) for the analytic
// orographies and periodic boundaries.  Every field is written straight in the reference layout
// (Fortran order, y fastest) into one device block and copied to the caller's arrays at the end,
// so the caller's t_grid receives what the reference's generator would have put there.  The
// work is elementwise or one column per thread (the vertical grid and the hydrostatic background
// state integrate down a column), with consecutive lanes on consecutive y, the fastest index of
// the reference layout, so that the stores of a wavefront are contiguous.  All arithmetic is
// FP64; sin, cos, exp, sinh, pow, asin and atan2 are the device's, so fields agree with a host
// evaluation to a few units in the last place.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "lgame_grid.h"

namespace lgame {
namespace {

// src/mo_constants.F90 and the generator's surface constants (mo_grid_generator.F90:748-757)
constexpr double R_E = 6371000.789927;
constexpr double OMEGA = 7.292115e-5;
constexpr double GRAVITY = 9.80616;
constexpr double P_0 = 100000.0;
constexpr double R_D = 287.057811;
constexpr double C_D_P = 1005.0;
constexpr double T_0 = 273.15;
constexpr double RHO_H2O = 1024.0;
constexpr double C_P_WATER_SFC = 4184.0;
constexpr double T_CONDUCTIVITY_WATER = 1.4e-7;
constexpr double SIGMA_SOIL = 0.352;
constexpr double Z_T_CONST = -10.0;
constexpr double PI = 3.14159265358979323846;
constexpr double EPSILON_SECURITY = 1e-10;
// the standard atmosphere of mo_set_initial_state.F90:384-422
constexpr double LAPSE_RATE = 0.0065;
constexpr double SURFACE_TEMP = 288.15;
constexpr double TROPO_HEIGHT = 11000.0;
constexpr double INV_HEIGHT = 20000.0;
constexpr double T_GRAD_INV = 0.001;
constexpr double P_0_STANDARD = 101325.0;

struct Gen {
  int ny, nx, nl, nflat, noro, oro_id, lsleve, lplane, lfreeslip, n_hills;
  double dx, dy, toa, stretch, dlat, dlon, lat_lu, lon_lu;
  double rot[9];
  const double *hills;
  const double *oro, *oro_s;  // orography_id 1: the orography and its smoothed part, (ny,nx)
  int *bad;  // set when a column's levels are not monotonic or its orography reaches the flat levels
};

// the output fields in the device block, each in the reference layout
struct Out {
  double *z_scalar, *z_u, *z_v, *z_w, *dx, *dy, *dz, *area_x, *area_y, *area_z, *volume, *slope_x, *slope_y, *ipw;
  double *area_dual_x, *area_dual_y, *area_dual_z, *z_area_dual_z, *fvec_x, *fvec_y, *fvec_z;
  double *ebg_u, *ebg_v, *ebg_w, *theta_v_bg, *exner_bg, *gravity_m_v, *gravity_potential;
  double *dx_dual, *dy_dual, *layer_thickness, *land_fraction, *roughness_length, *lake_fraction, *sfc_rho_c, *t_conduc_soil;
  double *t_const_soil, *sst, *lat_geo, *lon_geo;
};

// Fortran-order index of (i, j, k), 0-based, in an (n1, n2, *) array
__host__ __device__ inline size_t F3(int i, int j, int k, int n1, int n2) {
  return (size_t)i + (size_t)n1 * ((size_t)j + (size_t)n2 * (size_t)k);
}

__device__ inline double lat_scalar(const Gen &p, int ji) { return p.lat_lu - p.dlat * ji; }
__device__ inline double lon_scalar(const Gen &p, int jk) { return p.lon_lu + p.dlon * jk; }
// latitude of v row ji (0..ny): half a row north of scalar row ji, the last one half a row south
__device__ inline double lat_vrow(const Gen &p, int ji) {
  return ji < p.ny ? lat_scalar(p, ji) + 0.5 * p.dlat : lat_scalar(p, p.ny - 1) - 0.5 * p.dlat;
}
// longitude of u column jk (0..nx): half a column west of scalar column jk, the last one east
__device__ inline double lon_ucol(const Gen &p, int jk) {
  return jk < p.nx ? lon_scalar(p, jk) - 0.5 * p.dlon : lon_scalar(p, p.nx - 1) + 0.5 * p.dlon;
}

// patch_area, mo_grid_generator.F90:1612-1626
__host__ __device__ inline double patch_area(double lat, double dxa, double dya, bool lplane) {
  if (lplane) return R_E * R_E * dxa * dya + 0.0 * lat;
  return R_E * R_E * dxa * (sin(lat + 0.5 * dya) - sin(lat - 0.5 * dya));
}

// vertical_face_area, mo_grid_generator.F90:1628-1642
__device__ inline double vertical_face_area(double lower_z, double upper_z, double lower_length, bool lplane) {
  if (lplane) return lower_length * (upper_z - lower_z);
  return 0.5 * lower_length * (R_E + upper_z + R_E + lower_z) / (R_E + lower_z) * (upper_z - lower_z);
}

// find_global_normal, the rotation, find_geos (mo_grid_generator.F90:1656-1685)
__device__ inline void rotate_point(const Gen &p, double lat, double lon, double &lat_n, double &lon_n, double *rn_out = nullptr) {
  const double r[3] = {cos(lat) * cos(lon), cos(lat) * sin(lon), sin(lat)};
  double rn[3];
  for (int a = 0; a < 3; ++a) rn[a] = p.rot[3 * a] * r[0] + p.rot[3 * a + 1] * r[1] + p.rot[3 * a + 2] * r[2];
  lat_n = asin(rn[2] / sqrt(rn[0] * rn[0] + rn[1] * rn[1] + rn[2] * rn[2]));
  lon_n = atan2(rn[1], rn[0]);
  if (rn_out)
    for (int a = 0; a < 3; ++a) rn_out[a] = rn[a];
}

// the direction of a rotated local basis vector against the new local east (:1687-1712)
__device__ inline double direction(const Gen &p, double lat, double lon, const double basis[3], double &lat_n) {
  double lon_n;
  rotate_point(p, lat, lon, lat_n, lon_n);
  double bn[3];
  for (int a = 0; a < 3; ++a) bn[a] = p.rot[3 * a] * basis[0] + p.rot[3 * a + 1] * basis[1] + p.rot[3 * a + 2] * basis[2];
  const double li[3] = {-sin(lon_n), cos(lon_n), 0.0};
  const double lj[3] = {-sin(lat_n) * cos(lon_n), -sin(lat_n) * sin(lon_n), cos(lat_n)};
  const double x = li[0] * bn[0] + li[1] * bn[1] + li[2] * bn[2];
  const double y = lj[0] * bn[0] + lj[1] * bn[1] + lj[2] * bn[2];
  return atan2(y, x);
}

// Python's float remainder (the sign of the divisor), for the periodic hill distance
__device__ inline double py_mod(double a, double b) {
  double m = fmod(a, b);
  if (m != 0.0 && ((b < 0.0) != (m < 0.0))) m += b;
  return m;
}

// bg_temp and bg_pres of the standard atmosphere (mo_set_initial_state.F90:384-422)
__device__ inline double bg_temp(double h) {
  if (h < TROPO_HEIGHT) return SURFACE_TEMP - LAPSE_RATE * h;
  if (h < INV_HEIGHT) return SURFACE_TEMP - LAPSE_RATE * TROPO_HEIGHT;
  return SURFACE_TEMP - LAPSE_RATE * TROPO_HEIGHT + T_GRAD_INV * (h - INV_HEIGHT);
}
__device__ inline double bg_pres(double h) {
  return P_0_STANDARD * pow(1.0 - LAPSE_RATE * h / SURFACE_TEMP, GRAVITY / (R_D * LAPSE_RATE));
}
// geopot (mo_set_initial_state.F90:424-438)
__device__ inline double geopot(double h, bool lplane) {
  if (lplane) return GRAVITY * h;
  return -GRAVITY * (R_E * R_E) / (R_E + h) + GRAVITY * R_E;
}

// cos and sin of the latitude of every row of scalar points, interleaved
__global__ void k_lat_trig(const Gen p, double *trig) {
  const int jn = blockIdx.x * blockDim.x + threadIdx.x;
  if (jn >= p.ny) return;
  trig[2 * jn] = cos(lat_scalar(p, jn));
  trig[2 * jn + 1] = sin(lat_scalar(p, jn));
}

// smooth_hor_scalar2 (mo_grid_generator.F90:1512-1557): every point becomes the mean of all the
// others weighted by 1/(d + EPSILON_SECURITY)^1.5, d the great-circle angle between the grid
// coordinates (calculate_distance_h, :1582-1596).  One thread per target point sums over the
// sources in the reference's order (x outer, y inner); the source indices are the same for every
// lane, so the source values and the per-row cos/sin of the latitude (trig) are wave-uniform
// loads.  x^1.5 is x sqrt(x) (within an ulp of pow).  raw and the result are (ny,nx), Fortran
// order.
__global__ void k_smooth_oro(const Gen p, const double *__restrict__ trig, const double *__restrict__ raw,
                             double *out) {
  const int ji = blockIdx.x * blockDim.x + threadIdx.x, jk = blockIdx.y;
  if (ji >= p.ny) return;
  const double lon_a = lon_scalar(p, jk);
  const double cla = trig[2 * ji], sla = trig[2 * ji + 1];
  double acc = 0.0, sum_w = 0.0;
  for (int jm = 0; jm < p.nx; ++jm) {
    const double cdl = cos(lon_scalar(p, jm) - lon_a);
    const double *col = raw + (size_t)jm * p.ny;
    for (int jn = 0; jn < p.ny; ++jn) {
      const double cosang = cla * trig[2 * jn] * cdl + sla * trig[2 * jn + 1];
      const double base = 2.0 * asin(sqrt(0.5 - 0.5 * cosang)) + EPSILON_SECURITY;
      const double weight = (jm == jk && jn == ji) ? 0.0 : 1.0 / (base * sqrt(base));
      acc = acc + weight * col[jn];
      sum_w = sum_w + weight;
    }
  }
  out[F3(ji, jk, 0, p.ny, p.nx)] = acc / sum_w;
}

// the orography of orography_id 1 from its smoothed part (mo_grid_generator.F90:605-616): with
// SLEVE a tenth of the unsmoothed detail is kept, else only the smoothed field
__global__ void k_blend_oro(const Gen p, double *oro, const double *oro_s) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)p.ny * p.nx) return;
  oro[i] = p.lsleve ? oro_s[i] + 0.1 * (oro[i] - oro_s[i]) : oro_s[i];
}

// orography and its smoothed part at column (ji, jk) (mo_grid_generator.F90:326-361, 605-616)
__device__ void orography(const Gen &p, int ji, int jk, double &oro, double &oro_s) {
  const double x = p.dx * (jk + 1) - (p.nx / 2 + 1) * p.dx;
  oro = 0.0;
  oro_s = 0.0;
  if (p.oro_id == 1) {
    oro = p.oro[F3(ji, jk, 0, p.ny, p.nx)];
    oro_s = p.oro_s[F3(ji, jk, 0, p.ny, p.nx)];
    return;
  }
  if (p.oro_id == 2) {
    const double cc = cos(PI * x / 4000.0);
    oro = 250.0 * exp(-(x * x) / (5000.0 * 5000.0)) * (cc * cc);
    oro_s = 0.5 * 250.0 * exp(-(x * x) / (5000.0 * 5000.0));
  } else if (p.oro_id == 3) {
    if (fabs(x) <= 25e3) {
      const double c1 = cos(0.5 * PI * x / 25000.0), c2 = cos(PI * x / 8000.0);
      oro = 3000.0 * (c1 * c1) * (c2 * c2);
    }
  } else if (p.oro_id == 4) {
    const double e = exp(-(x * x) / (4.0 * (5000.0 * 5000.0)));
    oro = 500.0 * (e * e);
  } else if (p.oro_id == 5) {
    const double y = p.dy * (ji + 1) - (p.ny / 2 + 1) * p.dy;
    const double lx = p.nx * p.dx, ly = p.ny * p.dy;
    for (int h = 0; h < p.n_hills; ++h) {
      const double x0 = p.hills[4 * h], y0 = p.hills[4 * h + 1], ht = p.hills[4 * h + 2], w = p.hills[4 * h + 3];
      const double ddx = py_mod(x - x0 + 0.5 * lx, lx) - 0.5 * lx;
      const double ddy = py_mod(y - y0 + 0.5 * ly, ly) - 0.5 * ly;
      oro += ht * exp(-(ddx * ddx + ddy * ddy) / (w * w));
    }
  }
  if (p.oro_id != 2) oro_s = oro;
}

// One thread per column: the vertical grid (:894-951), z_scalar, z_w, dz, layer thickness,
// area_z, volume (:1030-1315), the hydrostatic background state with its vertical gradients
// (bg_setup, :1389-1457) and the per-column surface fields.
__global__ void k_gen_columns(const Gen p, const Out o) {
  const int ji = blockIdx.x * blockDim.x + threadIdx.x, jk = blockIdx.y;
  if (ji >= p.ny || jk >= p.nx) return;
  const int ny = p.ny, nx = p.nx, nl = p.nl, nlev = nl + 1, nflat = p.nflat;
  const bool lplane = p.lplane != 0;
  double oro, oro_s;
  orography(p, ji, jk, oro, oro_s);
  // the level heights vv(l), l = 0..nl; rolling: vv of this and the next level
  auto A = [&](int l) { return pow(1.0 - (double)l / nl, p.stretch) * p.toa; };
  const double toa_oro = A(nflat);
  auto vv = [&](int l) {
    const double a = A(l);
    if (l <= nflat) return a;
    if (p.lsleve) {
      if (p.oro_id == 2)
        return a + sinh((toa_oro - a) / 5.e3) / sinh(toa_oro / 5.e3) * oro_s +
               sinh((toa_oro - a) / 2.e3) / sinh(toa_oro / 2.e3) * (oro - oro_s);
      return a + sinh((toa_oro - a) / (0.5 * toa_oro)) / sinh(toa_oro / (0.5 * toa_oro)) * oro_s +
             sinh((toa_oro - a) / (0.25 * toa_oro)) / sinh(toa_oro / (0.25 * toa_oro)) * (oro - oro_s);
    }
    const double b = (l + 1 - (nflat + 1.0)) / p.noro;
    return a + b * oro;
  };
  if (oro >= toa_oro) *p.bad = 1;
  double v0 = vv(0);
  for (int l = 0; l < nl; ++l) {
    const double v1 = vv(l + 1);
    if (v1 - v0 >= 0.0) *p.bad = 1;
    o.z_scalar[F3(ji, jk, l, ny, nx)] = 0.5 * (v0 + v1);
    v0 = v1;
  }
  auto zs = [&](int l) { return o.z_scalar[F3(ji, jk, l, ny, nx)]; };
  auto zw = [&](int l) { return l == 0 ? p.toa : (l == nl ? oro : 0.5 * (zs(l - 1) + zs(l))); };
  for (int l = 0; l < nlev; ++l) {
    o.z_w[F3(ji, jk, l, ny, nx)] = zw(l);
    double d;
    if (l == 0)
      d = 2.0 * (p.toa - zs(0));
    else if (l == nl)
      d = 2.0 * (zs(nl - 1) - oro);
    else
      d = zs(l - 1) - zs(l);
    o.dz[F3(ji, jk, l, ny, nx)] = d;
  }
  const double lat = lat_scalar(p, ji);
  const double zsfc = oro;
  const double a_sfc = lplane ? 0.0 : patch_area(lat, p.dlon, p.dlat, false) * ((R_E + zsfc) * (R_E + zsfc)) / (R_E * R_E);
  auto area_z = [&](int l) {
    if (lplane) return p.dx * p.dy;
    const double r = R_E + zw(l), rs = R_E + zsfc;
    return a_sfc * (r * r) / (rs * rs);
  };
  for (int l = 0; l < nlev; ++l) o.area_z[F3(ji, jk, l, ny, nx)] = area_z(l);
  for (int l = 0; l < nl; ++l) {
    const size_t i = F3(ji, jk, l, ny, nx);
    const double zu = zw(l), zl = zw(l + 1);
    o.layer_thickness[i] = zu - zl;
    double vol;
    if (lplane) {
      vol = area_z(l + 1) * (zu - zl);
    } else {
      const double ru = R_E + zu, rl = R_E + zl;
      vol = 1.0 / 3.0 * (pow(ru, 3.0) - pow(rl, 3.0)) / (rl * rl) * area_z(l + 1);
    }
    o.volume[i] = vol;
  }
  // bg_setup: geopotential, temperature, Exner pressure integrated up from the lowest layer
  for (int l = 0; l < nl; ++l) o.gravity_potential[F3(ji, jk, l, ny, nx)] = geopot(zs(l), lplane);
  double e1 = pow(bg_pres(zs(nl - 1)) / P_0, R_D / C_D_P);
  o.exner_bg[F3(ji, jk, nl - 1, ny, nx)] = e1;
  for (int l = nl - 2; l >= 0; --l) {
    const double t = bg_temp(zs(l)), t1 = bg_temp(zs(l + 1));
    const double phi = o.gravity_potential[F3(ji, jk, l, ny, nx)], phi1 = o.gravity_potential[F3(ji, jk, l + 1, ny, nx)];
    const double b = -0.5 * e1 / t1 * (t - t1 + 2.0 / C_D_P * (phi - phi1));
    const double c = (e1 * e1) * t / t1;
    const double e = b + sqrt(b * b + c);
    o.exner_bg[F3(ji, jk, l, ny, nx)] = e;
    e1 = e;
  }
  for (int l = 0; l < nl; ++l) {
    const size_t i = F3(ji, jk, l, ny, nx);
    o.theta_v_bg[i] = bg_temp(zs(l)) / o.exner_bg[i];
  }
  // grad_vert of the Exner pressure and of the geopotential (mo_gradient_operators.F90:65-83)
  for (int l = 0; l < nlev; ++l) {
    const size_t i = F3(ji, jk, l, ny, nx);
    double ge = 0.0, gp = 0.0;
    if (l >= 1 && l <= nl - 1) {
      const size_t a = F3(ji, jk, l - 1, ny, nx), b = F3(ji, jk, l, ny, nx);
      ge = (o.exner_bg[a] - o.exner_bg[b]) / o.dz[i];
      gp = (o.gravity_potential[a] - o.gravity_potential[b]) / o.dz[i];
    }
    o.ebg_w[i] = ge;
    o.gravity_m_v[i] = gp;
  }
  // the surface without land (:746-800) and the geographic coordinates
  const size_t i2 = F3(ji, jk, 0, ny, nx);
  double lat_g, lon_g;
  rotate_point(p, lat, lon_scalar(p, jk), lat_g, lon_g);
  o.lat_geo[i2] = lat_g;
  o.lon_geo[i2] = lon_g;
  o.land_fraction[i2] = 0.0;
  o.lake_fraction[i2] = 0.0;
  o.roughness_length[i2] = 0.08;
  o.sfc_rho_c[i2] = RHO_H2O * C_P_WATER_SFC;
  o.t_conduc_soil[i2] = T_CONDUCTIVITY_WATER;
  o.t_const_soil[i2] = T_0 + 25.0 * cos(2.0 * lat_g);
  o.sst[i2] = 280.0;
}

// z_w of u column jk (0..nx) at level l: the mean of the neighbouring columns, extrapolated at
// both ends (the face heights of :1030-1315)
__device__ inline double zw_x(const Gen &p, const Out &o, int ji, int jk, int l) {
  const int ny = p.ny, nx = p.nx;
  auto a = [&](int k) { return o.z_w[F3(ji, k, l, ny, nx)]; };
  if (jk == 0) return a(0) + 0.5 * (a(0) - a(1));
  if (jk == nx) return a(nx - 1) + 0.5 * (a(nx - 1) - a(nx - 2));
  return 0.5 * (a(jk - 1) + a(jk));
}
__device__ inline double zw_y(const Gen &p, const Out &o, int ji, int jk, int l) {
  const int ny = p.ny, nx = p.nx;
  auto a = [&](int j) { return o.z_w[F3(j, jk, l, ny, nx)]; };
  if (ji == 0) return a(0) + 0.5 * (a(0) - a(1));
  if (ji == ny) return a(ny - 1) + 0.5 * (a(ny - 1) - a(ny - 2));
  return 0.5 * (a(ji - 1) + a(ji));
}
// z at u and v points (periodic means of the neighbouring scalar points)
__device__ inline double z_u_at(const Gen &p, const Out &o, int ji, int jk, int l) {
  const int ny = p.ny, nx = p.nx;
  const int k = jk == nx ? 0 : jk;
  const int km = k == 0 ? nx - 1 : k - 1;
  if (k == 0) return 0.5 * (o.z_scalar[F3(ji, 0, l, ny, nx)] + o.z_scalar[F3(ji, nx - 1, l, ny, nx)]);
  return 0.5 * (o.z_scalar[F3(ji, km, l, ny, nx)] + o.z_scalar[F3(ji, k, l, ny, nx)]);
}
__device__ inline double z_v_at(const Gen &p, const Out &o, int ji, int jk, int l) {
  const int ny = p.ny, nx = p.nx;
  const int j = ji == ny ? 0 : ji;
  if (j == 0) return 0.5 * (o.z_scalar[F3(0, jk, l, ny, nx)] + o.z_scalar[F3(ny - 1, jk, l, ny, nx)]);
  return 0.5 * (o.z_scalar[F3(j - 1, jk, l, ny, nx)] + o.z_scalar[F3(j, jk, l, ny, nx)]);
}
__device__ inline double dx_at(const Gen &p, const Out &o, int ji, int jk, int l) {
  if (p.lplane) return p.dx;
  return p.dx * cos(lat_scalar(p, ji)) * (R_E + z_u_at(p, o, ji, jk, l)) / R_E;
}
__device__ inline double dy_at(const Gen &p, const Out &o, int ji, int jk, int l) {
  if (p.lplane) return p.dy;
  return p.dy * (R_E + z_v_at(p, o, ji, jk, l)) / R_E;
}

// u points (ny, nx+1, nl): z_u, dx, slope_x, dy_dual, area_x
__global__ void k_gen_u(const Gen p, const Out o) {
  const int ji = blockIdx.x * blockDim.x + threadIdx.x, jk = blockIdx.y, l = blockIdx.z;
  const int ny = p.ny, nx = p.nx;
  if (jk > nx || ji >= ny || l >= p.nl) return;
  const size_t i = F3(ji, jk, l, ny, nx + 1);
  const double zu = z_u_at(p, o, ji, jk, l);
  o.z_u[i] = zu;
  const double dxm = dx_at(p, o, ji, jk, l);
  o.dx[i] = dxm;
  // grad_hor_cov of z_scalar (mo_gradient_operators.F90:17-63): column nx+1 repeats column 1
  const int k = jk == nx ? 0 : jk, km = k == 0 ? nx - 1 : k - 1;
  const double dxk = jk == nx ? dx_at(p, o, ji, 0, l) : dxm;
  o.slope_x[i] = (o.z_scalar[F3(ji, k, l, ny, nx)] - o.z_scalar[F3(ji, km, l, ny, nx)]) / dxk;
  o.dy_dual[i] = p.lplane ? p.dy : p.dy * (R_E + zu) / R_E;
  const double lower = zw_x(p, o, ji, jk, l + 1), upper = zw_x(p, o, ji, jk, l);
  const double ll = p.lplane ? p.dy : p.dy * (R_E + lower) / R_E;
  o.area_x[i] = vertical_face_area(lower, upper, ll, p.lplane != 0);
}

// v points (ny+1, nx, nl): z_v, dy, slope_y, dx_dual, area_y
__global__ void k_gen_v(const Gen p, const Out o) {
  const int ji = blockIdx.x * blockDim.x + threadIdx.x, jk = blockIdx.y, l = blockIdx.z;
  const int ny = p.ny, nx = p.nx;
  if (jk >= nx || ji > ny || l >= p.nl) return;
  const size_t i = F3(ji, jk, l, ny + 1, nx);
  const double zv = z_v_at(p, o, ji, jk, l);
  o.z_v[i] = zv;
  const double dym = dy_at(p, o, ji, jk, l);
  o.dy[i] = dym;
  const int j = ji == ny ? 0 : ji, jm = j == 0 ? ny - 1 : j - 1;
  const double dyj = ji == ny ? dy_at(p, o, 0, jk, l) : dym;
  o.slope_y[i] = (o.z_scalar[F3(jm, jk, l, ny, nx)] - o.z_scalar[F3(j, jk, l, ny, nx)]) / dyj;
  const double lat_v = lat_vrow(p, ji);
  o.dx_dual[i] = p.lplane ? p.dx : p.dx * cos(lat_v) * (R_E + zv) / R_E;
  const double lower = zw_y(p, o, ji, jk, l + 1), upper = zw_y(p, o, ji, jk, l);
  const double ll = p.lplane ? p.dx : p.dx * cos(lat_v) * (R_E + lower) / R_E;
  o.area_y[i] = vertical_face_area(lower, upper, ll, p.lplane != 0);
}

// dual cells: z_area_dual_z and area_dual_z (ny+1, nx+1, nl), area_dual_x (ny+1, nx, nl+1),
// area_dual_y (ny, nx+1, nl+1); one thread per (ny+1, nx+1) point and level l in 0..nl
__global__ void k_gen_dual(const Gen p, const Out o) {
  const int ji = blockIdx.x * blockDim.x + threadIdx.x, jk = blockIdx.y, l = blockIdx.z;
  const int ny = p.ny, nx = p.nx, nl = p.nl;
  if (jk > nx || ji > ny || l > nl) return;
  const bool lplane = p.lplane != 0;
  const double lat_v = lat_vrow(p, ji);
  if (l < nl) {  // z of the dual z faces: z_v extended over the u columns
    auto zv = [&](int k) { return z_v_at(p, o, ji, k, l); };
    double zad;
    if (jk == 0)
      zad = zv(0) + 0.5 * (zv(0) - zv(1));
    else if (jk == nx)
      zad = zv(nx - 1) + 0.5 * (zv(nx - 1) - zv(nx - 2));
    else
      zad = 0.5 * (zv(jk - 1) + zv(jk));
    const size_t i = F3(ji, jk, l, ny + 1, nx + 1);
    o.z_area_dual_z[i] = zad;
    if (lplane) {
      o.area_dual_z[i] = p.dx * p.dy;
    } else {
      const double r = R_E + zad;
      o.area_dual_z[i] = patch_area(lat_v, p.dlon, p.dlat, false) * (r * r) / (R_E * R_E);
    }
  }
  if (jk < nx) {  // area_dual_x at v point (ji, jk), level l
    double lz, ln;
    if (l == nl) {
      lz = zw_y(p, o, ji, jk, nl);
      ln = dy_at(p, o, ji, jk, nl - 1) * (R_E + lz) / (R_E + z_v_at(p, o, ji, jk, nl - 1));
    } else {
      lz = z_v_at(p, o, ji, jk, l);
      ln = dy_at(p, o, ji, jk, l);
    }
    const double uz = l == 0 ? zw_y(p, o, ji, jk, 0) : z_v_at(p, o, ji, jk, l - 1);
    if (lplane) ln = p.dy;
    o.area_dual_x[F3(ji, jk, l, ny + 1, nx)] = vertical_face_area(lz, uz, ln, lplane);
  }
  if (ji < ny) {  // area_dual_y at u point (ji, jk), level l
    double lz, ln;
    if (l == nl) {
      lz = zw_x(p, o, ji, jk, nl);
      ln = dx_at(p, o, ji, jk, nl - 1) * (R_E + lz) / (R_E + z_u_at(p, o, ji, jk, nl - 1));
    } else {
      lz = z_u_at(p, o, ji, jk, l);
      ln = dx_at(p, o, ji, jk, l);
    }
    const double uz = l == 0 ? zw_x(p, o, ji, jk, 0) : z_u_at(p, o, ji, jk, l - 1);
    if (lplane) ln = p.dx;
    o.area_dual_y[F3(ji, jk, l, ny, nx + 1)] = vertical_face_area(lz, uz, ln, lplane);
  }
}

// inner product weights (6, ny, nx, nl) (:1285-1315)
__global__ void k_gen_ipw(const Gen p, const Out o) {
  const int ji = blockIdx.x * blockDim.x + threadIdx.x, jk = blockIdx.y, l = blockIdx.z;
  const int ny = p.ny, nx = p.nx;
  if (jk >= nx || ji >= ny || l >= p.nl) return;
  const size_t i = F3(ji, jk, l, ny, nx);
  const double vol2 = 2.0 * o.volume[i];
  const size_t u0 = F3(ji, jk, l, ny, nx + 1), u1 = F3(ji, jk + 1, l, ny, nx + 1);
  const size_t v0 = F3(ji, jk, l, ny + 1, nx), v1 = F3(ji + 1, jk, l, ny + 1, nx);
  const size_t w0 = F3(ji, jk, l, ny, nx), w1 = F3(ji, jk, l + 1, ny, nx);
  double *w = o.ipw + 6 * i;
  w[0] = o.area_x[u1] * o.dx[u1] / vol2;
  w[1] = o.area_y[v0] * o.dy[v0] / vol2;
  w[2] = o.area_x[u0] * o.dx[u0] / vol2;
  w[3] = o.area_y[v1] * o.dy[v1] / vol2;
  w[4] = o.area_z[w0] * o.dz[w0] / vol2;
  w[5] = o.area_z[w1] * o.dz[w1] / vol2;
}

// grad_hor of the background Exner pressure with the terrain correction (mo_gradient_operators.F90:
// 85-130, remap_ver2hor_x/y of mo_averaging.F90:56-136), u points then v points
__global__ void k_gen_bg_grad(const Gen p, const Out o) {
  const int ji = blockIdx.x * blockDim.x + threadIdx.x, jk = blockIdx.y, l = blockIdx.z;
  const int ny = p.ny, nx = p.nx, nl = p.nl;
  if (jk > nx || ji > ny || l >= nl) return;
  const bool below = p.lfreeslip || l != nl - 1;
  auto remap = [&](int ja, int ka, int jb, int kb) {  // neighbours a, b of the point
    const size_t a = F3(ja, ka, l, ny, nx), b = F3(jb, kb, l, ny, nx);
    const size_t a1 = F3(ja, ka, l + 1, ny, nx), b1 = F3(jb, kb, l + 1, ny, nx);
    double r = (0.0 + o.ipw[6 * a + 4] * o.ebg_w[a]) + o.ipw[6 * b + 4] * o.ebg_w[b];
    if (below) r = (r + o.ipw[6 * a + 5] * o.ebg_w[a1]) + o.ipw[6 * b + 5] * o.ebg_w[b1];
    return r;
  };
  if (ji < ny) {  // u point (ji, jk)
    const int k = jk == nx ? 0 : jk, km = k == 0 ? nx - 1 : k - 1;
    const size_t i = F3(ji, jk, l, ny, nx + 1), ik = F3(ji, k, l, ny, nx + 1);
    double g = (o.exner_bg[F3(ji, k, l, ny, nx)] - o.exner_bg[F3(ji, km, l, ny, nx)]) / o.dx[ik];
    if (l >= p.nflat) g = g - o.slope_x[i] * (0.5 * remap(ji, km, ji, k));
    o.ebg_u[i] = g;
  }
  if (jk < nx) {  // v point (ji, jk)
    const int j = ji == ny ? 0 : ji, jm = j == 0 ? ny - 1 : j - 1;
    const size_t i = F3(ji, jk, l, ny + 1, nx), ij = F3(j, jk, l, ny + 1, nx);
    double g = (o.exner_bg[F3(jm, jk, l, ny, nx)] - o.exner_bg[F3(j, jk, l, ny, nx)]) / o.dy[ij];
    if (l >= p.nflat) g = g - o.slope_y[i] * (0.5 * remap(jm, jk, j, jk));
    o.ebg_v[i] = g;
  }
}

// Coriolis vectors (:213-322): fvec_x at v points (ny+1, nx), fvec_y at u points (ny, nx+1),
// fvec_z at dual points (ny+1, nx+1)
__global__ void k_gen_coriolis(const Gen p, const Out o) {
  const int ji = blockIdx.x * blockDim.x + threadIdx.x, jk = blockIdx.y;
  const int ny = p.ny, nx = p.nx;
  if (jk > nx || ji > ny) return;
  if (jk < nx) {
    const double lat = lat_vrow(p, ji), lon = lon_scalar(p, jk);
    const double basis[3] = {-sin(lat) * cos(lon), -sin(lat) * sin(lon), cos(lat)};
    double lat_n;
    const double d = direction(p, lat, lon, basis, lat_n);
    o.fvec_x[F3(ji, jk, 0, ny + 1, nx)] = 2.0 * OMEGA * cos(lat_n) * sin(d - 0.5 * PI);
  }
  if (ji < ny) {
    const double lat = lat_scalar(p, ji), lon = lon_ucol(p, jk);
    const double basis[3] = {-sin(lon), cos(lon), 0.0};
    double lat_n;
    const double d = direction(p, lat, lon, basis, lat_n);
    o.fvec_y[F3(ji, jk, 0, ny, nx + 1)] = 2.0 * OMEGA * cos(lat_n) * sin(d + 0.5 * PI);
  }
  double lat_n, lon_n;
  rotate_point(p, lat_vrow(p, ji), lon_ucol(p, jk), lat_n, lon_n);
  o.fvec_z[F3(ji, jk, 0, ny + 1, nx + 1)] = 2.0 * OMEGA * sin(lat_n);
}

}  // namespace

int grid_generate(const lgame_dims &d, const lgame_grid_params &gp, lgame_grid_out &out, std::string &err) {
  const int ny = d.ny, nx = d.nx, nl = d.n_layers, nlev = nl + 1;
  if (ny < 3 || nx < 3 || nl < 3 || d.n_oro_layers < 1 || d.n_oro_layers >= nl) {
    err = "dims: ny, nx, n_layers >= 3 and 1 <= n_oro_layers < n_layers";
    return LGAME_EINVAL;
  }
  const int id = gp.orography_id;
  if (id < 0 || id > 5) {
    err = "orography_id must be 0 to 5";
    return LGAME_EINVAL;
  }
  if (id == 1 && !gp.oro) {
    err = "orography_id 1 needs the orography (oro): the reference reads it from ETOPO files";
    return LGAME_EINVAL;
  }
  if (id == 5 && (gp.n_hills < 0 || (gp.n_hills > 0 && !gp.hills))) {
    err = "orography_id 5 needs n_hills >= 0 and the hills";
    return LGAME_EINVAL;
  }
  if (!(gp.dx > 0.0) || !(gp.dy > 0.0) || !(gp.toa > 0.0)) {
    err = "dx, dy and toa must be positive";
    return LGAME_EINVAL;
  }
  double *const *ptrs = &out.z_scalar;
  const int n_ptrs = (int)((&out.lon_geo_scalar - &out.z_scalar) + 1);
  for (int q = 0; q < n_ptrs; ++q)
    if (!ptrs[q]) {
      err = "an output array of lgame_grid_out is null";
      return LGAME_EINVAL;
    }
  Gen p{};
  p.ny = ny;
  p.nx = nx;
  p.nl = nl;
  p.nflat = nl - d.n_oro_layers;
  p.noro = d.n_oro_layers;
  p.oro_id = id;
  p.lsleve = gp.lsleve != 0;
  p.lplane = gp.lplane != 0;
  p.lfreeslip = gp.lfreeslip != 0;
  p.n_hills = id == 5 ? gp.n_hills : 0;
  p.dx = gp.dx;
  p.dy = gp.dy;
  p.toa = gp.toa;
  p.stretch = gp.stretching_parameter;
  p.dlat = gp.dy / R_E;
  p.dlon = gp.dx / R_E;
  p.lat_lu = (ny - 1.0) / 2.0 * p.dlat;
  p.lon_lu = -(nx - 1.0) / 2.0 * p.dlon;
  {  // rotation: about z by lon_center after about y by lat_center
    const double lc = gp.lat_center, oc = gp.lon_center;
    const double ry[9] = {std::cos(lc), 0.0, -std::sin(lc), 0.0, 1.0, 0.0, std::sin(lc), 0.0, std::cos(lc)};
    const double rz[9] = {std::cos(oc), -std::sin(oc), 0.0, std::sin(oc), std::cos(oc), 0.0, 0.0, 0.0, 1.0};
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) p.rot[3 * a + b] = rz[3 * a] * ry[b] + rz[3 * a + 1] * ry[3 + b] + rz[3 * a + 2] * ry[6 + b];
  }
  // the device block: every output field in the reference layout
  const size_t c2 = (size_t)ny * nx, cu = (size_t)ny * (nx + 1), cv = (size_t)(ny + 1) * nx, cd = (size_t)(ny + 1) * (nx + 1);
  struct Field {
    double **dev;
    double *host;
    size_t n;
  };
  Out o{};
  std::vector<Field> fields = {
      {&o.z_scalar, out.z_scalar, c2 * nl},          {&o.z_u, out.z_u, cu * nl},
      {&o.z_v, out.z_v, cv * nl},                    {&o.z_w, out.z_w, c2 * nlev},
      {&o.dx, out.dx, cu * nl},                      {&o.dy, out.dy, cv * nl},
      {&o.dz, out.dz, c2 * nlev},                    {&o.area_x, out.area_x, cu * nl},
      {&o.area_y, out.area_y, cv * nl},              {&o.area_z, out.area_z, c2 * nlev},
      {&o.volume, out.volume, c2 * nl},              {&o.slope_x, out.slope_x, cu * nl},
      {&o.slope_y, out.slope_y, cv * nl},            {&o.ipw, out.inner_product_weights, 6 * c2 * nl},
      {&o.area_dual_x, out.area_dual_x, cv * nlev},  {&o.area_dual_y, out.area_dual_y, cu * nlev},
      {&o.area_dual_z, out.area_dual_z, cd * nl},    {&o.z_area_dual_z, out.z_area_dual_z, cd * nl},
      {&o.fvec_x, out.fvec_x, cv},                   {&o.fvec_y, out.fvec_y, cu},
      {&o.fvec_z, out.fvec_z, cd},                   {&o.ebg_u, out.exner_bg_grad_u, cu * nl},
      {&o.ebg_v, out.exner_bg_grad_v, cv * nl},      {&o.ebg_w, out.exner_bg_grad_w, c2 * nlev},
      {&o.theta_v_bg, out.theta_v_bg, c2 * nl},      {&o.exner_bg, out.exner_bg, c2 * nl},
      {&o.gravity_m_v, out.gravity_m_v, c2 * nlev},  {&o.gravity_potential, out.gravity_potential, c2 * nl},
      {&o.dx_dual, out.dx_dual, cv * nl},            {&o.dy_dual, out.dy_dual, cu * nl},
      {&o.layer_thickness, out.layer_thickness, c2 * nl}, {&o.land_fraction, out.land_fraction, c2},
      {&o.roughness_length, out.roughness_length, c2}, {&o.lake_fraction, out.lake_fraction, c2},
      {&o.sfc_rho_c, out.sfc_rho_c, c2},             {&o.t_conduc_soil, out.t_conduc_soil, c2},
      {&o.t_const_soil, out.t_const_soil, c2},       {&o.sst, out.sst, c2},
      {&o.lat_geo, out.lat_geo_scalar, c2},          {&o.lon_geo, out.lon_geo_scalar, c2},
  };
  size_t total = 0;
  for (auto &f : fields) total += f.n;
  char *block = nullptr;
  double *hills = nullptr, *oro = nullptr;
  int *bad = nullptr;
  hipStream_t st = nullptr;
  auto cleanup = [&]() {
    if (st) (void)hipStreamSynchronize(st);
    if (block) (void)hipFree(block);
    if (hills) (void)hipFree(hills);
    if (oro) (void)hipFree(oro);
    if (bad) (void)hipFree(bad);
    if (st) (void)hipStreamDestroy(st);
  };
  auto hip_fail = [&](hipError_t e, const char *what) {
    err = std::string(what) + ": " + hipGetErrorString(e);
    cleanup();
    return e == hipErrorOutOfMemory ? LGAME_ENOMEM : LGAME_EHIP;
  };
  hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
  if (e != hipSuccess) return hip_fail(e, "hipStreamCreate");
  e = hipMalloc(&block, total * sizeof(double));
  if (e != hipSuccess) return hip_fail(e, "hipMalloc");
  e = hipMalloc(&bad, sizeof(int));
  if (e == hipSuccess) e = hipMemsetAsync(bad, 0, sizeof(int), st);
  if (e == hipSuccess && p.n_hills > 0) {
    e = hipMalloc(&hills, sizeof(double) * 4 * p.n_hills);
    if (e == hipSuccess) e = hipMemcpyAsync(hills, gp.hills, sizeof(double) * 4 * p.n_hills, hipMemcpyHostToDevice, st);
  }
  if (e == hipSuccess && id == 1) {  // the orography and its smoothed part
    e = hipMalloc(&oro, (2 * c2 + 2 * (size_t)ny) * sizeof(double));
    if (e == hipSuccess) e = hipMemcpyAsync(oro, gp.oro, c2 * sizeof(double), hipMemcpyHostToDevice, st);
  }
  if (e != hipSuccess) return hip_fail(e, "device set-up");
  p.hills = hills;
  p.bad = bad;
  if (id == 1) {
    p.oro = oro;
    p.oro_s = oro + c2;
    double *trig = oro + 2 * c2;
    hipLaunchKernelGGL(k_lat_trig, dim3((ny + 63) / 64), dim3(64), 0, st, p, trig);
    hipLaunchKernelGGL(k_smooth_oro, dim3((ny + 63) / 64, nx), dim3(64), 0, st, p, trig, oro, oro + c2);
    hipLaunchKernelGGL(k_blend_oro, dim3((unsigned)((c2 + 255) / 256)), dim3(256), 0, st, p, oro, oro + c2);
  }
  size_t off = 0;
  for (auto &f : fields) {
    *f.dev = reinterpret_cast<double *>(block) + off;
    off += f.n;
  }
  constexpr int B = 64;
  hipLaunchKernelGGL(k_gen_columns, dim3((ny + B - 1) / B, nx), dim3(B), 0, st, p, o);
  hipLaunchKernelGGL(k_gen_u, dim3((ny + B - 1) / B, nx + 1, nl), dim3(B), 0, st, p, o);
  hipLaunchKernelGGL(k_gen_v, dim3((ny + 1 + B - 1) / B, nx, nl), dim3(B), 0, st, p, o);
  hipLaunchKernelGGL(k_gen_dual, dim3((ny + 1 + B - 1) / B, nx + 1, nlev), dim3(B), 0, st, p, o);
  hipLaunchKernelGGL(k_gen_ipw, dim3((ny + B - 1) / B, nx, nl), dim3(B), 0, st, p, o);
  hipLaunchKernelGGL(k_gen_bg_grad, dim3((ny + 1 + B - 1) / B, nx + 1, nl), dim3(B), 0, st, p, o);
  hipLaunchKernelGGL(k_gen_coriolis, dim3((ny + 1 + B - 1) / B, nx + 1), dim3(B), 0, st, p, o);
  e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "grid kernels");
  int bad_h = 0;
  e = hipMemcpyAsync(&bad_h, bad, sizeof(int), hipMemcpyDeviceToHost, st);
  for (auto &f : fields)
    if (e == hipSuccess) e = hipMemcpyAsync(f.host, *f.dev, f.n * sizeof(double), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return hip_fail(e, "grid download");
  cleanup();
  if (bad_h) {
    err = "problem in vertical grid generation: levels not monotonic, or orography reaching the lowest flat level "
          "(mo_grid_generator.F90:936-951)";
    return LGAME_EINVAL;
  }
  // host-side parts: the area of the turbulence parameterisations, the TRSK weights and the soil
  // grid (mo_grid_generator.F90:1111, :1319-1357, :1360-1378)
  double sum = 0.0;
  for (size_t q = 0; q < c2; ++q) sum += out.area_z[(size_t)nl * c2 + q];
  out.mean_velocity_area = 2.0 * sum / (double)(ny * nx);
  const bool lplane = p.lplane != 0;
  const double dlat = p.dlat, dlon = p.dlon, dx = p.dx, dy = p.dy;
  for (int ji = 0; ji < ny; ++ji) {
    const double lat = p.lat_lu - dlat * ji;
    const double base = patch_area(lat, dlon, dlat, lplane);
    const double t1 = 0.5 - patch_area(lat + 0.25 * dlat, 0.5 * dlon, 0.5 * dlat, lplane) / base;
    const double t2 = -(0.5 - patch_area(lat + 0.25 * dlat, dlon, 0.5 * dlat, lplane) / base);
    const double t3 =
        -(0.5 - (patch_area(lat + 0.25 * dlat, dlon, 0.5 * dlat, lplane) + patch_area(lat - 0.25 * dlat, 0.5 * dlon, 0.5 * dlat, lplane)) / base);
    double *tu = out.trsk_weights_u + 3 * ji;
    if (lplane) {
      tu[0] = t1 * dx / dx;
      tu[1] = t2 * dy / dx;
      tu[2] = t3 * dx / dx;
    } else {
      tu[0] = t1 * dx * std::cos(lat + 0.5 * dlat) / (dx * std::cos(lat));
      tu[1] = t2 * dy / (dx * std::cos(lat));
      tu[2] = t3 * dx * std::cos(lat - 0.5 * dlat) / (dx * std::cos(lat));
    }
  }
  for (int ji = 0; ji <= ny; ++ji) {
    const double llc = ji < ny ? p.lat_lu - dlat * ji : (p.lat_lu - dlat * (ny - 1)) - dlat;
    const double luc = ji == 0 ? p.lat_lu + dlat : p.lat_lu - dlat * (ji - 1);
    double *tv = out.trsk_weights_v + 2 * ji;
    tv[0] = -(0.5 - patch_area(llc + 0.25 * dlat, 0.5 * dlon, 0.5 * dlat, lplane) / patch_area(llc, dlon, dlat, lplane));
    tv[1] = -(0.5 - patch_area(luc - 0.25 * dlat, 0.5 * dlon, 0.5 * dlat, lplane) / patch_area(luc, dlon, dlat, lplane));
  }
  const int ns = d.nsoillays > 1 ? d.nsoillays : 1;
  std::vector<double> zi(ns + 1, 0.0);
  for (int jl = 2; jl <= ns + 1; ++jl) zi[jl - 1] = zi[jl - 2] + std::pow(SIGMA_SOIL, (double)(ns - jl));
  const double rescale = Z_T_CONST / zi[ns];
  for (int jl = 2; jl <= ns + 1; ++jl) zi[jl - 1] = rescale * zi[jl - 1];
  for (int jl = 0; jl <= ns; ++jl) out.z_soil_interface[jl] = zi[jl];
  for (int jl = 0; jl < ns; ++jl) out.z_soil_center[jl] = 0.5 * (zi[jl] + zi[jl + 1]);
  out.z_t_const = Z_T_CONST;
  return LGAME_OK;
}

}  // namespace lgame
