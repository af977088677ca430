"""Grid and background-state set-up for idealized runs (host side, once per run).

A vectorized numpy restatement of ``grid_setup`` (``io/mo_grid_generator.F90:29-1387``) for the
analytic orographies, and of ``bg_setup`` (:1389-1457).  The step consumes the resulting read-only
metrics; the Fortran driver of the reference builds the same arrays itself, so this module only
serves runs driven from Python (tests, benchmark).  Returned arrays are float64, Fortran-ordered,
with the reference extents (index 0 here is Fortran index 1).

Orographies: ``orography_id`` 0 (flat), 2 (Schaer), 3, 4 as in ``mo_grid_generator.F90:326-361``,
plus id 5 = a sum of Gaussian hills given in ``RunConfig.hills`` (set-up extension used for the
random-hill workloads; the reference reads real orography from files instead).
"""
from __future__ import annotations

import numpy as np

from .config import RunConfig

# mo_constants.F90
R_E = 6371000.789927
OMEGA = 7.292115e-5
GRAVITY = 9.80616
P_0 = 100000.0
R_D = 287.057811
C_D_P = 1005.0
LAPSE_RATE = 0.0065
SURFACE_TEMP = 288.15
TROPO_HEIGHT = 11000.0
INV_HEIGHT = 20000.0
T_GRAD_INV = 0.001
P_0_STANDARD = 101325.0
T_0 = 273.15
EPSILON_SECURITY = 1e-10
RHO_H2O = 1024.0
# surface and soil, mo_grid_generator.F90:748-757 and :1362-1363
DENSITY_SOIL = 1442.0
C_P_SOIL = 830.0
C_P_WATER_SFC = 4184.0
T_CONDUCTIVITY_WATER = 1.4e-7
T_CONDUCTIVITY_SOIL = 7.5e-7
SIGMA_SOIL = 0.352
Z_T_CONST = -10.0


def _f(shape):
    return np.zeros(shape, dtype=np.float64, order="F")


def bg_temp(h):
    """bg_temp, io/mo_set_initial_state.F90:384-402."""
    h = np.asarray(h, dtype=np.float64)
    return np.where(h < TROPO_HEIGHT, SURFACE_TEMP - LAPSE_RATE * h,
                    np.where(h < INV_HEIGHT, SURFACE_TEMP - LAPSE_RATE * TROPO_HEIGHT,
                             SURFACE_TEMP - LAPSE_RATE * TROPO_HEIGHT + T_GRAD_INV * (h - INV_HEIGHT)))


def bg_pres(h):
    """bg_pres, io/mo_set_initial_state.F90:404-422 (only the first branch is reachable)."""
    h = np.asarray(h, dtype=np.float64)
    if np.any(h >= INV_HEIGHT):
        raise ValueError("bg_pres argument above the inversion height")
    return P_0_STANDARD * (1.0 - LAPSE_RATE * h / SURFACE_TEMP) ** (GRAVITY / (R_D * LAPSE_RATE))


def geopot(h, lplane):
    """geopot, io/mo_set_initial_state.F90:424-438."""
    if lplane:
        return GRAVITY * h
    return -GRAVITY * R_E ** 2 / (R_E + h) + GRAVITY * R_E


def _patch_area(lat, dxa, dya, lplane):
    """patch_area, mo_grid_generator.F90:1612-1626."""
    if lplane:
        return R_E ** 2 * dxa * dya + 0.0 * lat
    return R_E ** 2 * dxa * (np.sin(lat + 0.5 * dya) - np.sin(lat - 0.5 * dya))


def _vertical_face_area(lower_z, upper_z, lower_length, lplane):
    """vertical_face_area, mo_grid_generator.F90:1628-1642."""
    if lplane:
        return lower_length * (upper_z - lower_z)
    return 0.5 * lower_length * (R_E + upper_z + R_E + lower_z) / (R_E + lower_z) * (upper_z - lower_z)


def _rotate_points(lat, lon, rot):
    """find_global_normal -> rotation -> find_geos (mo_grid_generator.F90:1656-1685)."""
    r = np.stack([np.cos(lat) * np.cos(lon), np.cos(lat) * np.sin(lon), np.sin(lat)])
    rn = np.tensordot(rot, r, axes=1)
    lat_n = np.arcsin(rn[2] / np.sqrt(rn[0] ** 2 + rn[1] ** 2 + rn[2] ** 2))
    lon_n = np.arctan2(rn[1], rn[0])
    return lat_n, lon_n


def _local_i(lon):
    return np.stack([-np.sin(lon), np.cos(lon), np.zeros_like(lon)])


def _local_j(lat, lon):
    return np.stack([-np.sin(lat) * np.cos(lon), -np.sin(lat) * np.sin(lon), np.cos(lat)])


def _direction(lat, lon, basis_old, rot):
    lat_n, lon_n = _rotate_points(lat, lon, rot)
    bn = np.tensordot(rot, basis_old, axes=1)
    x = np.sum(_local_i(lon_n) * bn, axis=0)
    y = np.sum(_local_j(lat_n, lon_n) * bn, axis=0)
    return lat_n, lon_n, np.arctan2(y, x)


def _grad_hor_cov(s, dxm, dym):
    """grad_hor_cov, spatial_operators/mo_gradient_operators.F90:17-63 (periodic)."""
    ny, nx = s.shape[0], s.shape[1]
    rx = _f((ny, nx + 1) + s.shape[2:])
    ry = _f((ny + 1, nx) + s.shape[2:])
    rx[:, 1:nx] = (s[:, 1:nx] - s[:, 0:nx - 1]) / dxm[:, 1:nx]
    ry[1:ny] = (s[0:ny - 1] - s[1:ny]) / dym[1:ny]
    rx[:, 0] = (s[:, 0] - s[:, nx - 1]) / dxm[:, 0]
    rx[:, nx] = rx[:, 0]
    ry[0] = (s[ny - 1] - s[0]) / dym[0]
    ry[ny] = ry[0]
    return rx, ry


def _grad_vert(s, dz):
    """grad_vert, mo_gradient_operators.F90:65-83 (levels 1 and n_levels left at zero)."""
    ny, nx, nl = s.shape
    r = _f((ny, nx, nl + 1))
    r[:, :, 1:nl] = (s[:, :, 0:nl - 1] - s[:, :, 1:nl]) / dz[:, :, 1:nl]
    return r


def _grad_hor(s, rz, g, cfg):
    """grad_hor, mo_gradient_operators.F90:85-130 with remap_ver2hor_x/y (mo_averaging.F90:56-136)."""
    ny, nx, nl = s.shape
    rx, ry = _grad_hor_cov(s, g["dx"], g["dy"])
    ipw = g["inner_product_weights"]
    nflat = cfg.n_flat_layers
    below = np.ones(nl, dtype=bool)
    if not cfg.lfreeslip:
        below[nl - 1] = False
    p5, p6 = ipw[4], ipw[5]

    def wrapx(a):  # columns (nx, 1, 2, ..., nx, 1): left and right neighbours of u points
        return np.concatenate([a[:, nx - 1:nx], a, a[:, 0:1]], axis=1)

    def wrapy(a):
        return np.concatenate([a[ny - 1:ny], a, a[0:1]], axis=0)

    # remap_ver2hor_x: neighbours (jk-1, jk), periodic at jk = 1 and nx+1
    a5, z0 = wrapx(p5), wrapx(rz[:, :, 0:nl])
    a6, z1 = wrapx(p6), wrapx(rz[:, :, 1:nl + 1])
    left, right = slice(0, nx + 1), slice(1, nx + 2)
    r = (0.0 + a5[:, left] * z0[:, left]) + a5[:, right] * z0[:, right]
    r2 = (r + a6[:, left] * z1[:, left]) + a6[:, right] * z1[:, right]
    r = np.where(below[None, None, :], r2, r)
    rx[:, :, nflat:] = rx[:, :, nflat:] - g["slope_x"][:, :, nflat:] * (0.5 * r)[:, :, nflat:]
    a5, z0 = wrapy(p5), wrapy(rz[:, :, 0:nl])
    a6, z1 = wrapy(p6), wrapy(rz[:, :, 1:nl + 1])
    up, dn = slice(0, ny + 1), slice(1, ny + 2)
    r = (0.0 + a5[up] * z0[up]) + a5[dn] * z0[dn]
    r2 = (r + a6[up] * z1[up]) + a6[dn] * z1[dn]
    r = np.where(below[None, None, :], r2, r)
    ry[:, :, nflat:] = ry[:, :, nflat:] - g["slope_y"][:, :, nflat:] * (0.5 * r)[:, :, nflat:]
    return rx, ry


def grid_coordinates(cfg: RunConfig):
    """lat_scalar(ny), lon_scalar(nx): the grid's own (unrotated) coordinates (mo_grid_generator.F90:213-240)."""
    dlat, dlon = cfg.dy / R_E, cfg.dx / R_E
    lat = (cfg.ny - 1.0) / 2.0 * dlat - dlat * np.arange(cfg.ny)
    lon = -(cfg.nx - 1.0) / 2.0 * dlon + dlon * np.arange(cfg.nx)
    return lat, lon


def smooth_hor_scalar2(a, cfg: RunConfig, rows=64):
    """smooth_hor_scalar2, mo_grid_generator.F90:1512-1557: every point becomes the mean of all the
    others weighted by 1/(d + EPSILON_SECURITY)**1.5, d the great-circle distance on the unit
    sphere (calculate_distance_h, :1582-1596) between the points' grid coordinates.  O((ny nx)^2):
    the targets go in blocks of `rows` y rows."""
    a = np.asarray(a, dtype=np.float64)
    ny, nx = a.shape
    lat, lon = grid_coordinates(cfg)
    out = _f((ny, nx))
    latb = lat[:, None]  # sources (jn, jm)
    lonb = lon[None, :]
    for i0 in range(0, ny, rows):
        for jk in range(nx):
            ia = np.arange(i0, min(ny, i0 + rows))
            la = lat[ia][:, None, None]
            cosang = np.cos(la) * np.cos(latb)[None] * np.cos(lonb - lon[jk])[None] + np.sin(la) * np.sin(latb)[None]
            with np.errstate(invalid="ignore"):  # the point itself may round to cos > 1; dropped below
                d = 2.0 * 1.0 * np.arcsin(np.sqrt(0.5 - 0.5 * cosang))
            w = 1.0 / (d + EPSILON_SECURITY) ** 1.5
            w[np.arange(len(ia)), ia, jk] = 0.0  # the point itself is left out
            out[ia, jk] = np.sum(w * a[None], axis=(1, 2)) / np.sum(w, axis=(1, 2))
    return out


def orography(cfg: RunConfig, oro_raw=None):
    """Surface height and its smoothed part (mo_grid_generator.F90:326-361; orography_id 1 is the
    reference's real orography, given here in oro_raw instead of the ETOPO files, smoothed and
    blended as :605-616)."""
    ny, nx, dx, dy = cfg.ny, cfg.nx, cfg.dx, cfg.dy
    oro = _f((ny, nx))
    oro_s = _f((ny, nx))
    jk = np.arange(1, nx + 1)
    x = dx * jk - (nx // 2 + 1) * dx
    if cfg.orography_id == 1:
        if oro_raw is None:
            raise ValueError("orography_id 1 needs the orography (oro_raw)")
        oro_s[:] = smooth_hor_scalar2(oro_raw, cfg)
        oro[:] = oro_s + 0.1 * (np.asarray(oro_raw) - oro_s) if cfg.lsleve else oro_s
        return oro, oro_s
    if cfg.orography_id == 0:
        pass
    elif cfg.orography_id == 2:
        oro[:] = (250.0 * np.exp(-x ** 2 / 5000.0 ** 2) * np.cos(np.pi * x / 4000.0) ** 2)[None, :]
        oro_s[:] = (0.5 * 250.0 * np.exp(-x ** 2 / 5000.0 ** 2))[None, :]
    elif cfg.orography_id == 3:
        h = np.where(np.abs(x) <= 25e3, 3000.0 * np.cos(0.5 * np.pi * x / 25000.0) ** 2 * np.cos(np.pi * x / 8000.0) ** 2, 0.0)
        oro[:] = h[None, :]
    elif cfg.orography_id == 4:
        oro[:] = (500.0 * np.exp(-x ** 2 / (4.0 * 5000.0 ** 2)) ** 2)[None, :]
    elif cfg.orography_id == 5:
        ji = np.arange(1, ny + 1)
        y = dy * ji - (ny // 2 + 1) * dy
        X, Y = np.meshgrid(x, y)
        Lx, Ly = nx * dx, ny * dy
        for (x0, y0, h, w) in cfg.hills:
            ddx = (X - x0 + 0.5 * Lx) % Lx - 0.5 * Lx
            ddy = (Y - y0 + 0.5 * Ly) % Ly - 0.5 * Ly
            oro += h * np.exp(-(ddx ** 2 + ddy ** 2) / w ** 2)
    else:
        raise ValueError(f"orography_id {cfg.orography_id} needs external data")
    if cfg.orography_id != 2:
        oro_s[:] = oro
    return oro, oro_s


def soil_properties(g: dict) -> None:
    """Volumetric heat capacity and temperature conductivity of the surface from the land and lake
    fractions, as mo_grid_generator.F90:746-800 sets them: water values where there is no land,
    land/lake-weighted soil and water values where there is."""
    land, lake = g["land_fraction"], g["lake_fraction"]
    has_land = land > EPSILON_SECURITY
    frac = np.where(has_land, land + lake, 1.0)
    g["sfc_rho_c"] = np.asfortranarray(np.where(
        has_land, (land * DENSITY_SOIL * C_P_SOIL + lake * RHO_H2O * C_P_WATER_SFC) / frac, RHO_H2O * C_P_WATER_SFC))
    g["t_conduc_soil"] = np.asfortranarray(np.where(
        has_land, (land * T_CONDUCTIVITY_SOIL + lake * T_CONDUCTIVITY_WATER) / frac, T_CONDUCTIVITY_WATER))


def grid_setup(cfg: RunConfig, oro_raw=None) -> dict:
    """Grid metrics of grid_setup (mo_grid_generator.F90:29-1387); periodic boundaries.  With
    orography_id 1, oro_raw (ny, nx) is the orography before the smoothing."""
    if not cfg.lperiodic:
        raise ValueError("only periodic grids are set up here")
    ny, nx, nl = cfg.ny, cfg.nx, cfg.n_layers
    nlev = nl + 1
    nflat = cfg.n_flat_layers
    lplane = cfg.lplane
    dx, dy, toa = cfg.dx, cfg.dy, cfg.toa
    g = {}
    # horizontal grid and rotation, :213-322
    dlat = dy / R_E
    dlon = dx / R_E
    lat_lu = (ny - 1.0) / 2.0 * dlat
    lon_lu = -(nx - 1.0) / 2.0 * dlon
    lat_scalar = lat_lu - dlat * np.arange(ny)
    lon_scalar = lon_lu + dlon * np.arange(nx)
    lc, oc = cfg.lat_center, cfg.lon_center
    rot_y = np.array([[np.cos(lc), 0.0, -np.sin(lc)], [0.0, 1.0, 0.0], [np.sin(lc), 0.0, np.cos(lc)]])
    rot_z = np.array([[np.cos(oc), -np.sin(oc), 0.0], [np.sin(oc), np.cos(oc), 0.0], [0.0, 0.0, 1.0]])
    rot = rot_z @ rot_y
    lat_u = np.repeat(lat_scalar[:, None], nx + 1, axis=1)
    lon_u = np.concatenate([lon_scalar - 0.5 * dlon, [lon_scalar[-1] + 0.5 * dlon]])[None, :].repeat(ny, axis=0)
    lat_u_g, _, dir_u = _direction(lat_u, lon_u, _local_i(lon_u), rot)
    lat_v = np.concatenate([lat_scalar + 0.5 * dlat, [lat_scalar[-1] - 0.5 * dlat]])[:, None].repeat(nx, axis=1)
    lon_v = np.repeat(lon_scalar[None, :], ny + 1, axis=0)
    lat_v_g, _, dir_v = _direction(lat_v, lon_v, _local_j(lat_v, lon_v), rot)
    g["fvec_x"] = np.asfortranarray(2.0 * OMEGA * np.cos(lat_v_g) * np.sin(dir_v - 0.5 * np.pi))
    g["fvec_y"] = np.asfortranarray(2.0 * OMEGA * np.cos(lat_u_g) * np.sin(dir_u + 0.5 * np.pi))
    lat_d = np.concatenate([lat_scalar + 0.5 * dlat, [lat_scalar[-1] - 0.5 * dlat]])[:, None].repeat(nx + 1, axis=1)
    lon_d = np.concatenate([lon_scalar - 0.5 * dlon, [lon_scalar[-1] + 0.5 * dlon]])[None, :].repeat(ny + 1, axis=0)
    lat_d_g, _ = _rotate_points(lat_d, lon_d, rot)
    g["fvec_z"] = np.asfortranarray(2.0 * OMEGA * np.sin(lat_d_g))
    # geographic coordinates of the scalar points (lat_geo_scalar, lon_geo_scalar), used by radiation
    lat_geo, lon_geo = _rotate_points(np.repeat(lat_scalar[:, None], nx, axis=1), np.repeat(lon_scalar[None, :], ny, axis=0), rot)
    g["lat_geo_scalar"] = np.asfortranarray(lat_geo)
    g["lon_geo_scalar"] = np.asfortranarray(lon_geo)

    # vertical grid, :894-951
    oro, oro_s = orography(cfg, oro_raw)
    jl = np.arange(1, nlev + 1, dtype=np.float64)
    z_rel = 1.0 - (jl - 1.0) / nl
    A = z_rel ** cfg.stretching_parameter * toa
    vv = np.empty((ny, nx, nlev))
    vv[:] = A[None, None, :]
    toa_oro = A[nflat]
    for l in range(nflat + 1, nlev):
        if cfg.lsleve:
            if cfg.orography_id == 2:
                vv[:, :, l] = A[l] + np.sinh((toa_oro - A[l]) / 5.e3) / np.sinh(toa_oro / 5.e3) * oro_s + \
                    np.sinh((toa_oro - A[l]) / 2.e3) / np.sinh(toa_oro / 2.e3) * (oro - oro_s)
            else:
                vv[:, :, l] = A[l] + np.sinh((toa_oro - A[l]) / (0.5 * toa_oro)) / np.sinh(toa_oro / (0.5 * toa_oro)) * oro_s + \
                    np.sinh((toa_oro - A[l]) / (0.25 * toa_oro)) / np.sinh(toa_oro / (0.25 * toa_oro)) * (oro - oro_s)
        else:
            B = (l + 1 - (nflat + 1.0)) / cfg.n_oro_layers
            vv[:, :, l] = A[l] + B * oro
    if np.any(np.diff(vv, axis=2) >= 0.0):
        raise ValueError("Problem in vertical grid generation (levels not monotonic)")
    if oro.max() >= vv[0, 0, nflat]:
        raise ValueError("Maximum of orography larger or equal to the height of the lowest flat level")
    z_scalar = np.asfortranarray(0.5 * (vv[:, :, :nl] + vv[:, :, 1:]))
    z_w = _f((ny, nx, nlev))
    z_w[:, :, 0] = toa
    z_w[:, :, 1:nl] = 0.5 * (z_scalar[:, :, :nl - 1] + z_scalar[:, :, 1:])
    z_w[:, :, nl] = oro
    z_u = _f((ny, nx + 1, nl))
    z_u[:, 1:nx] = 0.5 * (z_scalar[:, :nx - 1] + z_scalar[:, 1:])
    z_u[:, 0] = 0.5 * (z_scalar[:, 0] + z_scalar[:, nx - 1])
    z_u[:, nx] = z_u[:, 0]
    z_v = _f((ny + 1, nx, nl))
    z_v[1:ny] = 0.5 * (z_scalar[:ny - 1] + z_scalar[1:])
    z_v[0] = 0.5 * (z_scalar[0] + z_scalar[ny - 1])
    z_v[ny] = z_v[0]
    g.update(z_scalar=z_scalar, z_w=z_w, z_u=z_u, z_v=z_v)

    # horizontal distances, :956-980
    if lplane:
        dxm = np.full((ny, nx + 1, nl), dx, order="F")
        dym = np.full((ny + 1, nx, nl), dy, order="F")
    else:
        dxm = np.asfortranarray(dx * np.cos(lat_scalar)[:, None, None] * (R_E + z_u) / R_E)
        dym = np.asfortranarray(dy * (R_E + z_v) / R_E)
    g["dx"], g["dy"] = dxm, dym
    g["slope_x"], g["slope_y"] = _grad_hor_cov(z_scalar, dxm, dym)

    dz = _f((ny, nx, nlev))
    dz[:, :, 0] = 2.0 * (toa - z_scalar[:, :, 0])
    dz[:, :, nl] = 2.0 * (z_scalar[:, :, nl - 1] - z_w[:, :, nl])
    dz[:, :, 1:nl] = z_scalar[:, :, :nl - 1] - z_scalar[:, :, 1:]
    g["dz"] = dz
    # dual-grid distances and layer thicknesses, :1012-1031 and :1063
    if lplane:
        g["dx_dual"] = np.full((ny + 1, nx, nl), dx, order="F")
        g["dy_dual"] = np.full((ny, nx + 1, nl), dy, order="F")
    else:
        lat_dual = np.concatenate([lat_scalar + 0.5 * dlat, [lat_scalar[-1] - 0.5 * dlat]])
        g["dx_dual"] = np.asfortranarray(dx * np.cos(lat_dual)[:, None, None] * (R_E + z_v) / R_E)
        g["dy_dual"] = np.asfortranarray(dy * (R_E + z_u) / R_E)
    g["layer_thickness"] = np.asfortranarray(z_w[:, :, :nl] - z_w[:, :, 1:])

    # areas and volumes, :1030-1315
    if lplane:
        area_z = np.full((ny, nx, nlev), dx * dy, order="F")
    else:
        a_sfc = _patch_area(lat_scalar, dlon, dlat, False)[:, None] * (R_E + z_w[:, :, nl]) ** 2 / R_E ** 2
        area_z = np.asfortranarray(a_sfc[:, :, None] * (R_E + z_w) ** 2 / (R_E + z_w[:, :, nl:nl + 1]) ** 2)
    g["area_z"] = area_z
    # surface: no land unless a case sets a mask (:326, :764-767); roughness length 0.08 m (:774)
    g["land_fraction"] = np.zeros((ny, nx), order="F")
    g["roughness_length"] = np.full((ny, nx), 0.08, order="F")
    g["lake_fraction"] = np.zeros((ny, nx), order="F")
    soil_properties(g)
    # constant soil temperature at depth, :688; the sea surface temperature of the initial soil
    # temperature, 280 K (mo_set_initial_state.F90:236-247)
    g["t_const_soil"] = np.asfortranarray(T_0 + 25.0 * np.cos(2.0 * g["lat_geo_scalar"]))
    g["sst"] = np.full((ny, nx), 280.0, order="F")
    # the soil grid, :1360-1378
    ns = max(cfg.nsoillays, 1)
    zi = np.zeros(ns + 1)
    for jl in range(2, ns + 2):
        zi[jl - 1] = zi[jl - 2] + SIGMA_SOIL ** (ns - jl)
    zi[1:] = Z_T_CONST / zi[ns] * zi[1:]
    g["z_soil_interface"] = zi
    g["z_soil_center"] = 0.5 * (zi[:-1] + zi[1:])
    g["z_t_const"] = Z_T_CONST
    # area used by the turbulence parameterisations, :1111
    g["mean_velocity_area"] = 2.0 * float(np.sum(area_z[:, :, nl])) / (ny * nx)

    def ext_x(a):  # values at the nx+1 u-columns: averages, extrapolated at both ends
        out = np.empty((a.shape[0], nx + 1) + a.shape[2:])
        out[:, 1:nx] = 0.5 * (a[:, :nx - 1] + a[:, 1:])
        out[:, 0] = a[:, 0] + 0.5 * (a[:, 0] - a[:, 1])
        out[:, nx] = a[:, nx - 1] + 0.5 * (a[:, nx - 1] - a[:, nx - 2])
        return out

    def ext_y(a):
        out = np.empty((ny + 1,) + a.shape[1:])
        out[1:ny] = 0.5 * (a[:ny - 1] + a[1:])
        out[0] = a[0] + 0.5 * (a[0] - a[1])
        out[ny] = a[ny - 1] + 0.5 * (a[ny - 1] - a[ny - 2])
        return out

    zwx = ext_x(z_w)
    lower, upper = zwx[:, :, 1:], zwx[:, :, :nl]
    ll = dy if lplane else dy * (R_E + lower) / R_E
    g["area_x"] = np.asfortranarray(_vertical_face_area(lower, upper, ll, lplane))
    zwy = ext_y(z_w)
    lower, upper = zwy[:, :, 1:], zwy[:, :, :nl]
    lat_vv = np.concatenate([lat_scalar + 0.5 * dlat, [lat_scalar[-1] - 0.5 * dlat]])
    ll = dx if lplane else dx * np.cos(lat_vv)[:, None, None] * (R_E + lower) / R_E
    g["area_y"] = np.asfortranarray(_vertical_face_area(lower, upper, ll, lplane))
    zad = ext_x(z_v)
    g["z_area_dual_z"] = np.asfortranarray(zad)
    if lplane:
        g["area_dual_z"] = np.full((ny + 1, nx + 1, nl), dx * dy, order="F")
    else:
        pa = _patch_area(lat_vv, dlon, dlat, False)
        g["area_dual_z"] = np.asfortranarray(pa[:, None, None] * (R_E + zad) ** 2 / R_E ** 2)
    # area_dual_x (ny+1, nx, nlev)
    adx = _f((ny + 1, nx, nlev))
    zwy_s = zwy[:, :, nl]
    for l in range(nlev):
        if l == nl:
            lz = zwy_s
            ln = dym[:, :, nl - 1] * (R_E + lz) / (R_E + z_v[:, :, nl - 1])
        else:
            lz = z_v[:, :, l]
            ln = dym[:, :, l]
        uz = zwy[:, :, 0] if l == 0 else z_v[:, :, l - 1]
        if lplane:
            ln = dy
        adx[:, :, l] = _vertical_face_area(lz, uz, ln, lplane)
    g["area_dual_x"] = adx
    ady = _f((ny, nx + 1, nlev))
    zwx_s = zwx[:, :, nl]
    for l in range(nlev):
        if l == nl:
            lz = zwx_s
            ln = dxm[:, :, nl - 1] * (R_E + lz) / (R_E + z_u[:, :, nl - 1])
        else:
            lz = z_u[:, :, l]
            ln = dxm[:, :, l]
        uz = zwx[:, :, 0] if l == 0 else z_u[:, :, l - 1]
        if lplane:
            ln = dx
        ady[:, :, l] = _vertical_face_area(lz, uz, ln, lplane)
    g["area_dual_y"] = ady
    if lplane:
        vol = area_z[:, :, 1:] * (z_w[:, :, :nl] - z_w[:, :, 1:])
    else:
        vol = 1.0 / 3.0 * ((R_E + z_w[:, :, :nl]) ** 3 - (R_E + z_w[:, :, 1:]) ** 3) / (R_E + z_w[:, :, 1:]) ** 2 * area_z[:, :, 1:]
    g["volume"] = np.asfortranarray(vol)
    ipw = _f((6, ny, nx, nl))
    ax, ay = g["area_x"], g["area_y"]
    ipw[0] = ax[:, 1:] * dxm[:, 1:] / (2.0 * vol)
    ipw[1] = ay[:ny] * dym[:ny] / (2.0 * vol)
    ipw[2] = ax[:, :nx] * dxm[:, :nx] / (2.0 * vol)
    ipw[3] = ay[1:] * dym[1:] / (2.0 * vol)
    ipw[4] = area_z[:, :, :nl] * dz[:, :, :nl] / (2.0 * vol)
    ipw[5] = area_z[:, :, 1:] * dz[:, :, 1:] / (2.0 * vol)
    g["inner_product_weights"] = ipw

    # TRSK weights, :1319-1357
    tu = _f((3, ny))
    base = _patch_area(lat_scalar, dlon, dlat, lplane)
    t1 = 0.5 - _patch_area(lat_scalar + 0.25 * dlat, 0.5 * dlon, 0.5 * dlat, lplane) / base
    t2 = -(0.5 - _patch_area(lat_scalar + 0.25 * dlat, dlon, 0.5 * dlat, lplane) / base)
    t3 = -(0.5 - (_patch_area(lat_scalar + 0.25 * dlat, dlon, 0.5 * dlat, lplane)
                  + _patch_area(lat_scalar - 0.25 * dlat, 0.5 * dlon, 0.5 * dlat, lplane)) / base)
    if lplane:
        tu[0] = t1 * dx / dx
        tu[1] = t2 * dy / dx
        tu[2] = t3 * dx / dx
    else:
        tu[0] = t1 * dx * np.cos(lat_scalar + 0.5 * dlat) / (dx * np.cos(lat_scalar))
        tu[1] = t2 * dy / (dx * np.cos(lat_scalar))
        tu[2] = t3 * dx * np.cos(lat_scalar - 0.5 * dlat) / (dx * np.cos(lat_scalar))
    g["trsk_weights_u"] = tu
    tv = _f((2, ny + 1))
    llc = np.concatenate([lat_scalar, [lat_scalar[-1] - dlat]])
    tv[0] = -(0.5 - _patch_area(llc + 0.25 * dlat, 0.5 * dlon, 0.5 * dlat, lplane) / _patch_area(llc, dlon, dlat, lplane))
    luc = np.concatenate([[lat_scalar[0] + dlat], lat_scalar])
    tv[1] = -(0.5 - _patch_area(luc - 0.25 * dlat, 0.5 * dlon, 0.5 * dlat, lplane) / _patch_area(luc, dlon, dlat, lplane))
    g["trsk_weights_v"] = tv

    bg_setup(g, cfg)
    return g


def bg_setup(g: dict, cfg: RunConfig) -> None:
    """Hydrostatic background state, mo_grid_generator.F90:1389-1457."""
    z = g["z_scalar"]
    ny, nx, nl = z.shape
    phi = np.asfortranarray(geopot(z, cfg.lplane))
    exner = _f((ny, nx, nl))
    T = bg_temp(z)
    exner[:, :, nl - 1] = (bg_pres(z[:, :, nl - 1]) / P_0) ** (R_D / C_D_P)
    for l in range(nl - 2, -1, -1):
        b = -0.5 * exner[:, :, l + 1] / T[:, :, l + 1] * (T[:, :, l] - T[:, :, l + 1] + 2.0 / C_D_P * (phi[:, :, l] - phi[:, :, l + 1]))
        c = exner[:, :, l + 1] ** 2 * T[:, :, l] / T[:, :, l + 1]
        exner[:, :, l] = b + np.sqrt(b ** 2 + c)
    g["gravity_potential"] = phi
    g["exner_bg"] = exner
    g["theta_v_bg"] = np.asfortranarray(T / exner)
    g["exner_bg_grad_w"] = _grad_vert(exner, g["dz"])
    g["exner_bg_grad_u"], g["exner_bg_grad_v"] = _grad_hor(exner, g["exner_bg_grad_w"], g, cfg)
    g["gravity_m_v"] = _grad_vert(phi, g["dz"])
