// lgame_rad.h -- radiation boundary of the step (device data and the enqueue entry point).
#pragma once

#include "lgame_dev.h"
#include "lgame_mlp.h"

namespace lgame {

// what one launch of k_rad does (bit flags)
enum RadFlags {
  RAD_TEND = 1,   // write radiation_tendency and the surface fluxes of every column
  RAD_X = 2,      // write the surrogate inputs of every column
  RAD_Y = 4,      // write the solver's outputs in the surrogate's units (K/day, W m^-2)
  RAD_AUDIT = 8,  // solver on the audit sample only; compare with ypred, fall back above tol
};

// Per-column radiation data of a handle; column index (ji-1)*nx + (jk-1).
struct RadDev {
  const double *lat, *lon, *albedo, *tsfc;
  double *sw_in, *lw_out;
  float *x;           // [col][5*nl+4] (RAD_X), or [col][LGAME_RAD_N_IN] (k_rad_gather)
  float *y;           // [col][nl+2]
  const float *ypred; // [col][LGAME_RAD_N_OUT], the surrogate's outputs on its own levels
  int *list;              // the columns of the audit pass (a column c outside the envelope as ~c)
  unsigned *list_count;   // number of entries in list
  const float *y_mean, *y_scale;  // the surrogate's output normalisation (training envelope)
  unsigned long long *counts;  // [0] audited columns, [1] columns given back to the solver
  double *sq;                  // sum over audited columns and layers of the squared heating-rate error
  double tol;                  // per-column heating-rate RMSE tolerance [K/day]
  unsigned audit_thresh;       // a column is audited if its hash is below this (of 2^32)
  unsigned long long call;     // index of the radiation call (varies the audit sample)
};

struct RadCall {
  int mode;      // LGAME_RAD_TEACHER or LGAME_RAD_SURROGATE
  double t_rad;  // time coordinate of the call, t_0 + dtime_rad/2
  RadDev dev;
  const lgame_mlp *mlp;
  MlpWork *work;  // the handle's own inference buffers
};

template <class R> __global__ void k_rad(const CtxT<R> c, const StateDev s, const RadDev r, double t_rad, int flags);

// update_rad_fluxes (mo_manage_radiation_calls.F90:17-59) for the state `old`, on `stream`;
// returns 0 or an LGAME_E* code (the launches or the inference failed)
template <class R>
int enqueue_radiation(const CtxT<R> &c, const StateDev &old, const RadCall &rc, hipStream_t stream);

}  // namespace lgame
