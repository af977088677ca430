// lgame_mlp.hip -- per-column MLP of the radiation surrogate: BF16 MFMA GEMMs with FP32
// accumulation, GELU, mean-squared-error backward pass and FP32 Adam master weights.
//
// Network: n_in -> hidden x n_hidden_layers (GELU) -> n_out.  Dimensions are padded to multiples
// of 64 in device storage (zero weights, zero inputs), so every GEMM is a whole number of 64x64
// output tiles with a K loop in steps of 32.  Weights are stored [out][in] (in contiguous), the
// activations [column][feature].  Inputs and targets are normalised with per-feature means and
// scales that the caller sets (lgame_mlp_set_norm); the network works on normalised values.
//
// GEMM: one 256-thread block per 64x64 tile of C; each wavefront owns a 32x32 quarter and
// computes it with 2x2 v_mfma_f32_16x16x16_bf16 tiles.  A and B tiles of 32 k-values are staged
// in LDS with k contiguous, so the MFMA operands of a lane (4 consecutive k) are one 8-byte LDS
// read.  Operands can be k-contiguous or m/n-contiguous in memory (forward, data gradient and
// weight gradient use different ones); the global loads are arranged to be contiguous either way.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/lgame_gpu.h"
#include "lgame_mlp.h"

namespace lgame {

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

__host__ __device__ inline uint16_t f2bf(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));  // inf / nan
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t b) {
  const uint32_t u = (uint32_t)b << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

enum Epi { EPI_F32 = 0, EPI_BIAS_GELU = 1, EPI_BIAS_F32 = 2, EPI_DGELU = 3 };

struct GemmArgs {
  const uint16_t *A;  // A(m,k) = A[m*sam + k*sak]
  long long sam, sak;
  const uint16_t *B;  // B(k,n) = B[k*sbk + n*sbn]
  long long sbk, sbn;
  int M, N, K;
  float *C;            // EPI_F32 / EPI_BIAS_F32: C[m*ldc + n]
  long long ldc;
  const float *bias;   // per n
  uint16_t *Z, *Aout;  // EPI_BIAS_GELU: pre-activation and activation, [m*ldz + n]
  const uint16_t *Zin; // EPI_DGELU: pre-activation of the layer
  uint16_t *D;         // EPI_DGELU: output gradient [m*ldz + n]
  long long ldz;
};

constexpr int TM = 64, TN = 64, TK = 32, LDK = TK + 8;  // LDS row: 32 k values + 8 padding

template <int EPI>
__global__ void __launch_bounds__(256) k_gemm(const GemmArgs g) {
  __shared__ uint16_t As[TM][LDK];
  __shared__ uint16_t Bs[TN][LDK];
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  float4_t acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = float4_t{0.f, 0.f, 0.f, 0.f};
  const bool a_kc = g.sak == 1, b_kc = g.sbk == 1;
  for (int k0 = 0; k0 < g.K; k0 += TK) {
    // A tile: 64 m x 32 k, 8 elements per thread, contiguous in memory
    if (a_kc) {
      const int m = t >> 2, kk = (t & 3) * 8;
      const uint16_t *src = g.A + (long long)(m0 + m) * g.sam + (k0 + kk);
#pragma unroll
      for (int q = 0; q < 8; ++q) As[m][kk + q] = src[q];
    } else {
      const int kk = t >> 3, mm = (t & 7) * 8;
      const uint16_t *src = g.A + (long long)(k0 + kk) * g.sak + (long long)(m0 + mm) * g.sam;
#pragma unroll
      for (int q = 0; q < 8; ++q) As[mm + q][kk] = src[(long long)q * g.sam];
    }
    if (b_kc) {
      const int n = t >> 2, kk = (t & 3) * 8;
      const uint16_t *src = g.B + (long long)(n0 + n) * g.sbn + (k0 + kk);
#pragma unroll
      for (int q = 0; q < 8; ++q) Bs[n][kk + q] = src[q];
    } else {
      const int kk = t >> 3, nn = (t & 7) * 8;
      const uint16_t *src = g.B + (long long)(k0 + kk) * g.sbk + (long long)(n0 + nn) * g.sbn;
#pragma unroll
      for (int q = 0; q < 8; ++q) Bs[nn + q][kk] = src[(long long)q * g.sbn];
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < TK; ks += 16) {
      // lane layout of v_mfma_f32_16x16x16_bf16: row/column index lane % 16, k = 4*(lane / 16) + j
      const int r = lane & 15, kq = ks + 4 * (lane >> 4);
      short4_t av[2], bv[2];
#pragma unroll
      for (int a = 0; a < 2; ++a) av[a] = *reinterpret_cast<const short4_t *>(&As[wm + 16 * a + r][kq]);
#pragma unroll
      for (int b = 0; b < 2; ++b) bv[b] = *reinterpret_cast<const short4_t *>(&Bs[wn + 16 * b + r][kq]);
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av[a], bv[b], acc[a][b], 0, 0, 0);
    }
    __syncthreads();
  }
  // result layout: lane holds rows 4*(lane/16) + j (j = 0..3) of column lane % 16
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int m = m0 + wm + 16 * a + 4 * (lane >> 4) + j;
        const int n = n0 + wn + 16 * b + (lane & 15);
        const float v = acc[a][b][j];
        if (EPI == EPI_F32) {
          g.C[(long long)m * g.ldc + n] = v;
        } else if (EPI == EPI_BIAS_F32) {
          g.C[(long long)m * g.ldc + n] = v + g.bias[n];
        } else if (EPI == EPI_BIAS_GELU) {
          const float z = v + g.bias[n];
          g.Z[(long long)m * g.ldz + n] = f2bf(z);
          g.Aout[(long long)m * g.ldz + n] = f2bf(gelu(z));
        } else {
          const float z = bf2f(g.Zin[(long long)m * g.ldz + n]);
          g.D[(long long)m * g.ldz + n] = f2bf(v * gelu_grad(z));
        }
      }
}

// normalised BF16 inputs, padded: X[c][i] = (x - mean)/scale for i < n_in, 0 above
__global__ void k_norm_in(const float *x, long long ncol, int n_in, int ldi, const float *mean, const float *scale, uint16_t *X) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= ncol * ldi) return;
  const long long c = idx / ldi;
  const int i = (int)(idx % ldi);
  X[idx] = i < n_in ? f2bf((x[c * n_in + i] - mean[i]) / scale[i]) : (uint16_t)0;
}

// physical outputs from the network's normalised ones
__global__ void k_denorm_out(const float *Yn, long long ncol, int n_out, int ldo, const float *mean, const float *scale, float *y) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= ncol * n_out) return;
  const long long c = idx / n_out;
  const int o = (int)(idx % n_out);
  y[idx] = Yn[c * ldo + o] * scale[o] + mean[o];
}

// normalised FP32 targets, padded
__global__ void k_norm_out(const float *y, long long ncol, int n_out, int ldo, const float *mean, const float *scale, float *Yn) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= ncol * ldo) return;
  const long long c = idx / ldo;
  const int o = (int)(idx % ldo);
  Yn[idx] = o < n_out ? (y[c * n_out + o] - mean[o]) / scale[o] : 0.0f;
}

// mean-squared-error gradient of a batch: dY = 2 (yhat - y) / (B * n_out) on the valid outputs;
// per-block partial sums of the squared error
__global__ void __launch_bounds__(256) k_mse_grad(const float *yhat, const float *y, const int *rows, int B, int n_out, int ldo,
                                                  uint16_t *dY, float *partial) {
  __shared__ float sh[256];
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  float se = 0.0f;
  if (idx < (long long)B * ldo) {
    const int b = (int)(idx / ldo), o = (int)(idx % ldo);
    float d = 0.0f;
    if (o < n_out) {
      const long long r = rows ? (long long)rows[b] : b;
      const float e = yhat[(long long)b * ldo + o] - y[r * ldo + o];
      se = e * e;
      d = 2.0f * e / (float)((long long)B * n_out);
    }
    dY[idx] = f2bf(d);
  }
  sh[threadIdx.x] = se;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// column sums of a [B][ld] BF16 gradient: the bias gradient (FP32)
__global__ void k_colsum(const uint16_t *D, int B, int ld, float *out) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= ld) return;
  float s = 0.0f;
  for (int b = 0; b < B; ++b) s += bf2f(D[(long long)b * ld + n]);
  out[n] = s;
}

// gather a batch of rows of the normalised data set
__global__ void k_gather(const uint16_t *X, const int *rows, int B, int ld, uint16_t *out) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)B * ld) return;
  const int b = (int)(idx / ld), i = (int)(idx % ld);
  out[idx] = X[(long long)rows[b] * ld + i];
}

// Adam on the FP32 master weights of one parameter block (padded [rows][ld] layout; entries
// outside the valid rows/columns stay zero), then the BF16 copy
__global__ void k_adam(float *w, float *m, float *v, const float *grad, uint16_t *wbf, long long n, int ld, int valid_cols,
                       int valid_rows, float lr, float b1, float b2, float eps, float c1, float c2) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int col = (int)(i % ld);
  const long long row = i / ld;
  if (col >= valid_cols || row >= valid_rows) return;
  const float gi = grad[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  if (wbf) wbf[i] = f2bf(wi);
}

__global__ void k_to_bf16(const float *w, uint16_t *wbf, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) wbf[i] = f2bf(w[i]);
}

}  // namespace lgame

using namespace lgame;

// ------------------------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------------------------
struct MlpLayer {
  int in, out;      // logical sizes
  int in_p, out_p;  // padded to 64
  float *w = nullptr, *b = nullptr;          // FP32 master [out_p][in_p], [out_p]
  float *mw = nullptr, *vw = nullptr, *mb = nullptr, *vb = nullptr;  // Adam moments
  float *gw = nullptr, *gb = nullptr;        // gradients
  uint16_t *wbf = nullptr;                   // BF16 copy
};

struct lgame_mlp {
  int n_in, n_out, hidden, nh;
  std::vector<MlpLayer> L;
  float *x_mean = nullptr, *x_scale = nullptr, *y_mean = nullptr, *y_scale = nullptr;
  // activations for up to `cap` rows: per layer Z and A (BF16), gradient buffers
  long long cap = 0;
  std::vector<uint16_t *> Z, A;
  uint16_t *X = nullptr, *dA = nullptr, *dB = nullptr, *dYb = nullptr;
  float *Yf = nullptr, *Yt = nullptr, *partial = nullptr, *xin = nullptr, *yout = nullptr;
  int *rows = nullptr;
  // data set (normalised, device)
  uint16_t *DX = nullptr;
  float *DY = nullptr;
  long long ndata = 0;
  long long adam_t = 0;
  hipStream_t stream = nullptr;
  std::string err;
};

namespace {

int mfail(lgame_mlp *m, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (m) m->err = buf;
  return code;
}

#define MCHK(m, call)                                                                            \
  do {                                                                                           \
    hipError_t e_ = (call);                                                                      \
    if (e_ != hipSuccess) return mfail(m, LGAME_EHIP, "%s: %s", #call, hipGetErrorString(e_));  \
  } while (0)

inline int pad64(int n) { return (n + 63) / 64 * 64; }
inline unsigned nblk(long long n, int b = 256) { return (unsigned)((n + b - 1) / b); }

int ensure_cap(lgame_mlp *m, long long rows) {
  if (rows <= m->cap) return 0;
  for (auto p : m->Z) (void)hipFree(p);
  for (auto p : m->A) (void)hipFree(p);
  m->Z.clear();
  m->A.clear();
  for (void *p : {(void *)m->X, (void *)m->dA, (void *)m->dB, (void *)m->dYb, (void *)m->Yf, (void *)m->Yt, (void *)m->partial,
                  (void *)m->xin, (void *)m->yout, (void *)m->rows})
    if (p) (void)hipFree(p);
  const int hp = pad64(m->hidden), ip = pad64(m->n_in), op = pad64(m->n_out);
  const int wmax = hp > ip ? hp : ip;
  m->cap = rows;
  for (int l = 0; l < m->nh; ++l) {
    uint16_t *z, *a;
    MCHK(m, hipMalloc(&z, (size_t)rows * hp * 2));
    MCHK(m, hipMalloc(&a, (size_t)rows * hp * 2));
    m->Z.push_back(z);
    m->A.push_back(a);
  }
  MCHK(m, hipMalloc(&m->X, (size_t)rows * ip * 2));
  MCHK(m, hipMemset(m->X, 0, (size_t)rows * ip * 2));
  MCHK(m, hipMalloc(&m->dA, (size_t)rows * wmax * 2));
  MCHK(m, hipMalloc(&m->dB, (size_t)rows * wmax * 2));
  MCHK(m, hipMalloc(&m->dYb, (size_t)rows * op * 2));
  MCHK(m, hipMalloc(&m->Yf, (size_t)rows * op * 4));
  MCHK(m, hipMalloc(&m->Yt, (size_t)rows * op * 4));
  MCHK(m, hipMalloc(&m->partial, (size_t)nblk(rows * op) * 4));
  MCHK(m, hipMalloc(&m->xin, (size_t)rows * m->n_in * 4));
  MCHK(m, hipMalloc(&m->yout, (size_t)rows * m->n_out * 4));
  MCHK(m, hipMalloc(&m->rows, (size_t)rows * 4));
  return 0;
}

template <int EPI>
void gemm(hipStream_t st, const GemmArgs &g) {
  hipLaunchKernelGGL(k_gemm<EPI>, dim3(g.N / TN, g.M / TM), dim3(256), 0, st, g);
}

// forward pass of B normalised rows X (padded) -> Yf (FP32, padded), keeping Z and A
void forward(lgame_mlp *m, const uint16_t *X, int B) {
  const uint16_t *in = X;
  int in_p = pad64(m->n_in);
  for (int l = 0; l <= m->nh; ++l) {
    const MlpLayer &Ly = m->L[l];
    GemmArgs g{};
    g.A = in;
    g.sam = in_p;
    g.sak = 1;
    g.B = Ly.wbf;
    g.sbk = 1;
    g.sbn = Ly.in_p;
    g.M = B;
    g.N = Ly.out_p;
    g.K = Ly.in_p;
    g.bias = Ly.b;
    if (l < m->nh) {
      g.Z = m->Z[l];
      g.Aout = m->A[l];
      g.ldz = Ly.out_p;
      gemm<EPI_BIAS_GELU>(m->stream, g);
      in = m->A[l];
      in_p = Ly.out_p;
    } else {
      g.C = m->Yf;
      g.ldc = Ly.out_p;
      gemm<EPI_BIAS_F32>(m->stream, g);
    }
  }
}

// backward pass from dYb (BF16, [B][out_p]) with the activations of the last forward pass
void backward(lgame_mlp *m, const uint16_t *X, int B) {
  const uint16_t *dcur = m->dYb;
  for (int l = m->nh; l >= 0; --l) {
    MlpLayer &Ly = m->L[l];
    const uint16_t *in = l == 0 ? X : m->A[l - 1];
    // weight gradient: gW[o][i] = sum_b d[b][o] * in[b][i]
    GemmArgs gw{};
    gw.A = dcur;
    gw.sam = 1;
    gw.sak = Ly.out_p;
    gw.B = in;
    gw.sbk = Ly.in_p;
    gw.sbn = 1;
    gw.M = Ly.out_p;
    gw.N = Ly.in_p;
    gw.K = B;
    gw.C = Ly.gw;
    gw.ldc = Ly.in_p;
    gemm<EPI_F32>(m->stream, gw);
    hipLaunchKernelGGL(k_colsum, dim3(nblk(Ly.out_p)), dim3(256), 0, m->stream, dcur, B, Ly.out_p, Ly.gb);
    if (l == 0) break;
    // data gradient through the GELU of layer l-1: d_prev[b][i] = (sum_o d[b][o] W[o][i]) gelu'(Z)
    uint16_t *dnext = (dcur == m->dA) ? m->dB : m->dA;
    GemmArgs gd{};
    gd.A = dcur;
    gd.sam = Ly.out_p;
    gd.sak = 1;
    gd.B = Ly.wbf;
    gd.sbk = Ly.in_p;
    gd.sbn = 1;
    gd.M = B;
    gd.N = Ly.in_p;
    gd.K = Ly.out_p;
    gd.Zin = m->Z[l - 1];
    gd.D = dnext;
    gd.ldz = Ly.in_p;
    gemm<EPI_DGELU>(m->stream, gd);
    dcur = dnext;
  }
}

float loss_sum(lgame_mlp *m, int B) {
  const int op = pad64(m->n_out);
  std::vector<float> p(nblk((long long)B * op));
  if (hipMemcpyAsync(p.data(), m->partial, p.size() * 4, hipMemcpyDeviceToHost, m->stream) != hipSuccess) return NAN;
  if (hipStreamSynchronize(m->stream) != hipSuccess) return NAN;
  double s = 0.0;
  for (float v : p) s += v;
  return (float)s;
}

}  // namespace

namespace lgame {

// Device-side inference used by the radiation surrogate: x (ncol x n_in, FP32, device) ->
// y (ncol x n_out, FP32, device), in chunks of the activation capacity, on `stream`.
int mlp_infer_device(lgame_mlp *m, const float *x, long long ncol, float *y, hipStream_t stream) {
  const long long chunk = 65536;
  if (ensure_cap(m, chunk)) return LGAME_ENOMEM;
  hipStream_t saved = m->stream;
  m->stream = stream;
  const int ip = pad64(m->n_in), op = pad64(m->n_out);
  for (long long c0 = 0; c0 < ncol; c0 += chunk) {
    const long long nc = ncol - c0 < chunk ? ncol - c0 : chunk;
    const int B = (int)((nc + 63) / 64 * 64);
    // rows beyond nc are left from earlier chunks: finite, and their outputs are not copied
    hipLaunchKernelGGL(k_norm_in, dim3(nblk(nc * ip)), dim3(256), 0, stream, x + c0 * m->n_in, nc, m->n_in, ip, m->x_mean,
                       m->x_scale, m->X);
    forward(m, m->X, B);
    hipLaunchKernelGGL(k_denorm_out, dim3(nblk(nc * m->n_out)), dim3(256), 0, stream, m->Yf, nc, m->n_out, op, m->y_mean,
                       m->y_scale, y + c0 * m->n_out);
  }
  m->stream = saved;
  return hipGetLastError() == hipSuccess ? 0 : LGAME_EHIP;
}

int mlp_dims(const lgame_mlp *m, int *n_in, int *n_out) {
  *n_in = m->n_in;
  *n_out = m->n_out;
  return 0;
}

}  // namespace lgame

extern "C" {

int lgame_mlp_create(int n_in, int hidden, int n_hidden_layers, int n_out, unsigned long long seed, lgame_mlp **out) {
  if (!out) return LGAME_EINVAL;
  *out = nullptr;
  if (n_in < 1 || hidden < 1 || n_hidden_layers < 1 || n_out < 1 || n_hidden_layers > 8) return LGAME_EINVAL;
  lgame_mlp *m = new lgame_mlp();
  m->n_in = n_in;
  m->n_out = n_out;
  m->hidden = hidden;
  m->nh = n_hidden_layers;
  if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) {
    delete m;
    return LGAME_EHIP;
  }
  // He-style uniform initialisation from a 64-bit LCG (reproducible across platforms)
  unsigned long long st = seed * 6364136223846793005ull + 1442695040888963407ull;
  auto urand = [&]() {
    st = st * 6364136223846793005ull + 1442695040888963407ull;
    return (double)(st >> 11) * (1.0 / 9007199254740992.0);
  };
  for (int l = 0; l <= m->nh; ++l) {
    MlpLayer Ly;
    Ly.in = l == 0 ? n_in : hidden;
    Ly.out = l == m->nh ? n_out : hidden;
    Ly.in_p = pad64(Ly.in);
    Ly.out_p = pad64(Ly.out);
    const size_t nw = (size_t)Ly.in_p * Ly.out_p, nb = Ly.out_p;
    std::vector<float> w(nw, 0.0f), b(nb, 0.0f);
    const double lim = std::sqrt(6.0 / Ly.in) * (l == m->nh ? 0.5 : 1.0);
    for (int o = 0; o < Ly.out; ++o)
      for (int i = 0; i < Ly.in; ++i) w[(size_t)o * Ly.in_p + i] = (float)((2.0 * urand() - 1.0) * lim);
    bool ok = true;
    for (float **p : {&Ly.w, &Ly.mw, &Ly.vw, &Ly.gw}) ok = ok && hipMalloc(p, nw * 4) == hipSuccess;
    for (float **p : {&Ly.b, &Ly.mb, &Ly.vb, &Ly.gb}) ok = ok && hipMalloc(p, nb * 4) == hipSuccess;
    ok = ok && hipMalloc(&Ly.wbf, nw * 2) == hipSuccess;
    ok = ok && hipMemcpy(Ly.w, w.data(), nw * 4, hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemcpy(Ly.b, b.data(), nb * 4, hipMemcpyHostToDevice) == hipSuccess;
    for (float *p : {Ly.mw, Ly.vw, Ly.gw}) ok = ok && hipMemset(p, 0, nw * 4) == hipSuccess;
    for (float *p : {Ly.mb, Ly.vb, Ly.gb}) ok = ok && hipMemset(p, 0, nb * 4) == hipSuccess;
    m->L.push_back(Ly);
    if (!ok) {
      lgame_mlp_destroy(m);
      return LGAME_ENOMEM;
    }
    hipLaunchKernelGGL(k_to_bf16, dim3(nblk(nw)), dim3(256), 0, m->stream, Ly.w, Ly.wbf, (long long)nw);
  }
  bool ok = true;
  for (float **p : {&m->x_mean, &m->x_scale}) ok = ok && hipMalloc(p, (size_t)n_in * 4) == hipSuccess;
  for (float **p : {&m->y_mean, &m->y_scale}) ok = ok && hipMalloc(p, (size_t)n_out * 4) == hipSuccess;
  if (!ok) {
    lgame_mlp_destroy(m);
    return LGAME_ENOMEM;
  }
  std::vector<float> zeros((size_t)(n_in > n_out ? n_in : n_out), 0.0f), ones(zeros.size(), 1.0f);
  (void)hipMemcpy(m->x_mean, zeros.data(), (size_t)n_in * 4, hipMemcpyHostToDevice);
  (void)hipMemcpy(m->x_scale, ones.data(), (size_t)n_in * 4, hipMemcpyHostToDevice);
  (void)hipMemcpy(m->y_mean, zeros.data(), (size_t)n_out * 4, hipMemcpyHostToDevice);
  (void)hipMemcpy(m->y_scale, ones.data(), (size_t)n_out * 4, hipMemcpyHostToDevice);
  if (hipStreamSynchronize(m->stream) != hipSuccess) {
    lgame_mlp_destroy(m);
    return LGAME_EHIP;
  }
  *out = m;
  return LGAME_OK;
}

long long lgame_mlp_n_params(const lgame_mlp *m) {
  if (!m) return -1;
  long long n = 0;
  for (auto &Ly : m->L) n += (long long)Ly.in * Ly.out + Ly.out;
  return n;
}

// parameters in the order W_1 [out][in], b_1, W_2, b_2, ... (unpadded, FP32)
static int params_io(lgame_mlp *m, float *p, long long n, bool set) {
  if (!m || !p) return LGAME_EINVAL;
  if (n != lgame_mlp_n_params(m)) return mfail(m, LGAME_EINVAL, "expected %lld parameters, got %lld", lgame_mlp_n_params(m), n);
  long long off = 0;
  for (auto &Ly : m->L) {
    std::vector<float> w((size_t)Ly.in_p * Ly.out_p, 0.0f), b(Ly.out_p, 0.0f);
    if (!set) {
      MCHK(m, hipMemcpy(w.data(), Ly.w, w.size() * 4, hipMemcpyDeviceToHost));
      MCHK(m, hipMemcpy(b.data(), Ly.b, b.size() * 4, hipMemcpyDeviceToHost));
    }
    for (int o = 0; o < Ly.out; ++o)
      for (int i = 0; i < Ly.in; ++i) {
        float &x = w[(size_t)o * Ly.in_p + i];
        if (set) x = p[off]; else p[off] = x;
        ++off;
      }
    for (int o = 0; o < Ly.out; ++o) {
      if (set) b[o] = p[off]; else p[off] = b[o];
      ++off;
    }
    if (set) {
      MCHK(m, hipMemcpy(Ly.w, w.data(), w.size() * 4, hipMemcpyHostToDevice));
      MCHK(m, hipMemcpy(Ly.b, b.data(), b.size() * 4, hipMemcpyHostToDevice));
      hipLaunchKernelGGL(k_to_bf16, dim3(nblk((long long)w.size())), dim3(256), 0, m->stream, Ly.w, Ly.wbf, (long long)w.size());
      MCHK(m, hipStreamSynchronize(m->stream));
    }
  }
  if (set) m->adam_t = 0;
  return LGAME_OK;
}

int lgame_mlp_set_params(lgame_mlp *m, const float *p, long long n) { return params_io(m, const_cast<float *>(p), n, true); }
int lgame_mlp_get_params(lgame_mlp *m, float *p, long long n) { return params_io(m, p, n, false); }

int lgame_mlp_set_norm(lgame_mlp *m, const float *x_mean, const float *x_scale, const float *y_mean, const float *y_scale) {
  if (!m || !x_mean || !x_scale || !y_mean || !y_scale) return LGAME_EINVAL;
  for (int i = 0; i < m->n_in; ++i)
    if (!(x_scale[i] > 0.0f)) return mfail(m, LGAME_EINVAL, "x_scale[%d] must be positive", i);
  for (int i = 0; i < m->n_out; ++i)
    if (!(y_scale[i] > 0.0f)) return mfail(m, LGAME_EINVAL, "y_scale[%d] must be positive", i);
  MCHK(m, hipMemcpy(m->x_mean, x_mean, (size_t)m->n_in * 4, hipMemcpyHostToDevice));
  MCHK(m, hipMemcpy(m->x_scale, x_scale, (size_t)m->n_in * 4, hipMemcpyHostToDevice));
  MCHK(m, hipMemcpy(m->y_mean, y_mean, (size_t)m->n_out * 4, hipMemcpyHostToDevice));
  MCHK(m, hipMemcpy(m->y_scale, y_scale, (size_t)m->n_out * 4, hipMemcpyHostToDevice));
  return LGAME_OK;
}

int lgame_mlp_forward(lgame_mlp *m, const float *x, long long n, float *y) {
  if (!m || !x || !y || n < 0) return LGAME_EINVAL;
  if (n == 0) return LGAME_OK;
  const long long chunk = 65536;
  if (ensure_cap(m, chunk)) return LGAME_ENOMEM;
  for (long long c0 = 0; c0 < n; c0 += chunk) {
    const long long nc = n - c0 < chunk ? n - c0 : chunk;
    MCHK(m, hipMemcpyAsync(m->xin, x + c0 * m->n_in, (size_t)nc * m->n_in * 4, hipMemcpyHostToDevice, m->stream));
    int rc = mlp_infer_device(m, m->xin, nc, m->yout, m->stream);
    if (rc) return mfail(m, rc, "inference failed");
    MCHK(m, hipMemcpyAsync(y + c0 * m->n_out, m->yout, (size_t)nc * m->n_out * 4, hipMemcpyDeviceToHost, m->stream));
    MCHK(m, hipStreamSynchronize(m->stream));
  }
  return LGAME_OK;
}

// Gradient of the mean-squared error (normalised outputs) over n rows, parameters in the order of
// lgame_mlp_get_params; *loss receives the mean squared error.  n must be a multiple of 64.
int lgame_mlp_gradient(lgame_mlp *m, const float *x, const float *y, long long n, float *grad, float *loss) {
  if (!m || !x || !y || !grad || !loss || n <= 0 || n % 64) return LGAME_EINVAL;
  if (ensure_cap(m, n)) return mfail(m, LGAME_ENOMEM, "activation buffers");
  const int ip = pad64(m->n_in), op = pad64(m->n_out);
  float *dy;
  MCHK(m, hipMalloc(&dy, (size_t)n * m->n_out * 4));
  MCHK(m, hipMemcpyAsync(m->xin, x, (size_t)n * m->n_in * 4, hipMemcpyHostToDevice, m->stream));
  MCHK(m, hipMemcpyAsync(dy, y, (size_t)n * m->n_out * 4, hipMemcpyHostToDevice, m->stream));
  hipLaunchKernelGGL(k_norm_in, dim3(nblk(n * ip)), dim3(256), 0, m->stream, m->xin, n, m->n_in, ip, m->x_mean, m->x_scale, m->X);
  hipLaunchKernelGGL(k_norm_out, dim3(nblk(n * op)), dim3(256), 0, m->stream, dy, n, m->n_out, op, m->y_mean, m->y_scale, m->Yt);
  forward(m, m->X, (int)n);
  hipLaunchKernelGGL(k_mse_grad, dim3(nblk(n * op)), dim3(256), 0, m->stream, m->Yf, m->Yt, (const int *)nullptr, (int)n, m->n_out,
                     op, m->dYb, m->partial);
  backward(m, m->X, (int)n);
  *loss = loss_sum(m, (int)n) / (float)(n * m->n_out);
  (void)hipFree(dy);
  long long off = 0;
  for (auto &Ly : m->L) {
    std::vector<float> w((size_t)Ly.in_p * Ly.out_p), b(Ly.out_p);
    MCHK(m, hipMemcpy(w.data(), Ly.gw, w.size() * 4, hipMemcpyDeviceToHost));
    MCHK(m, hipMemcpy(b.data(), Ly.gb, b.size() * 4, hipMemcpyDeviceToHost));
    for (int o = 0; o < Ly.out; ++o)
      for (int i = 0; i < Ly.in; ++i) grad[off++] = w[(size_t)o * Ly.in_p + i];
    for (int o = 0; o < Ly.out; ++o) grad[off++] = b[o];
  }
  return LGAME_OK;
}

int lgame_mlp_load_data(lgame_mlp *m, const float *x, const float *y, long long n) {
  if (!m || !x || !y || n < 64) return LGAME_EINVAL;
  const int ip = pad64(m->n_in), op = pad64(m->n_out);
  if (m->DX) (void)hipFree(m->DX);
  if (m->DY) (void)hipFree(m->DY);
  m->DX = nullptr;
  m->DY = nullptr;
  MCHK(m, hipMalloc(&m->DX, (size_t)n * ip * 2));
  MCHK(m, hipMalloc(&m->DY, (size_t)n * op * 4));
  const long long chunk = 65536;
  float *tx, *ty;
  MCHK(m, hipMalloc(&tx, (size_t)chunk * m->n_in * 4));
  MCHK(m, hipMalloc(&ty, (size_t)chunk * m->n_out * 4));
  for (long long c0 = 0; c0 < n; c0 += chunk) {
    const long long nc = n - c0 < chunk ? n - c0 : chunk;
    MCHK(m, hipMemcpyAsync(tx, x + c0 * m->n_in, (size_t)nc * m->n_in * 4, hipMemcpyHostToDevice, m->stream));
    MCHK(m, hipMemcpyAsync(ty, y + c0 * m->n_out, (size_t)nc * m->n_out * 4, hipMemcpyHostToDevice, m->stream));
    hipLaunchKernelGGL(k_norm_in, dim3(nblk(nc * ip)), dim3(256), 0, m->stream, tx, nc, m->n_in, ip, m->x_mean, m->x_scale,
                       m->DX + c0 * ip);
    hipLaunchKernelGGL(k_norm_out, dim3(nblk(nc * op)), dim3(256), 0, m->stream, ty, nc, m->n_out, op, m->y_mean, m->y_scale,
                       m->DY + c0 * op);
    MCHK(m, hipStreamSynchronize(m->stream));
  }
  (void)hipFree(tx);
  (void)hipFree(ty);
  m->ndata = n;
  return LGAME_OK;
}

// Adam training on the loaded data set: `epochs` passes in shuffled batches of `batch` rows (a
// multiple of 64); loss[e] receives the mean squared error of epoch e.  *seconds (if non-null)
// the device time of the run.
int lgame_mlp_train(lgame_mlp *m, int epochs, int batch, float lr, unsigned long long seed, float *loss, double *seconds) {
  if (!m || epochs < 1 || batch < 64 || batch % 64 || !loss) return LGAME_EINVAL;
  if (m->ndata < batch) return mfail(m, LGAME_ESTATE, "load at least one batch of data first");
  if (ensure_cap(m, batch)) return mfail(m, LGAME_ENOMEM, "activation buffers");
  const int ip = pad64(m->n_in), op = pad64(m->n_out);
  const long long nb = m->ndata / batch;
  std::vector<int> perm(m->ndata);
  unsigned long long st = seed * 0x9E3779B97F4A7C15ull + 1;
  hipEvent_t e0, e1;
  MCHK(m, hipEventCreate(&e0));
  MCHK(m, hipEventCreate(&e1));
  int *drows;
  MCHK(m, hipMalloc(&drows, (size_t)m->ndata * 4));
  MCHK(m, hipEventRecord(e0, m->stream));
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int ep = 0; ep < epochs; ++ep) {
    for (long long i = 0; i < m->ndata; ++i) perm[i] = (int)i;
    for (long long i = m->ndata - 1; i > 0; --i) {
      st ^= st << 13;
      st ^= st >> 7;
      st ^= st << 17;
      std::swap(perm[i], perm[st % (unsigned long long)(i + 1)]);
    }
    MCHK(m, hipMemcpyAsync(drows, perm.data(), (size_t)m->ndata * 4, hipMemcpyHostToDevice, m->stream));
    double se = 0.0;
    std::vector<float> part(nblk((long long)batch * op));
    for (long long b = 0; b < nb; ++b) {
      const int *rows = drows + b * batch;
      hipLaunchKernelGGL(k_gather, dim3(nblk((long long)batch * ip)), dim3(256), 0, m->stream, m->DX, rows, batch, ip, m->X);
      forward(m, m->X, batch);
      hipLaunchKernelGGL(k_mse_grad, dim3(nblk((long long)batch * op)), dim3(256), 0, m->stream, m->Yf, m->DY, rows, batch,
                         m->n_out, op, m->dYb, m->partial);
      backward(m, m->X, batch);
      ++m->adam_t;
      const float c1 = 1.0f - std::pow(b1, (float)m->adam_t), c2 = 1.0f - std::pow(b2, (float)m->adam_t);
      for (auto &Ly : m->L) {
        const long long nw = (long long)Ly.in_p * Ly.out_p;
        hipLaunchKernelGGL(k_adam, dim3(nblk(nw)), dim3(256), 0, m->stream, Ly.w, Ly.mw, Ly.vw, (const float *)Ly.gw, Ly.wbf, nw,
                           Ly.in_p, Ly.in, Ly.out, lr, b1, b2, eps, c1, c2);
        hipLaunchKernelGGL(k_adam, dim3(nblk(Ly.out_p)), dim3(256), 0, m->stream, Ly.b, Ly.mb, Ly.vb, (const float *)Ly.gb,
                           (uint16_t *)nullptr, (long long)Ly.out_p, Ly.out_p, Ly.out, 1, lr, b1, b2, eps, c1, c2);
      }
      if (b % 16 == 15 || b == nb - 1) {  // read the loss now and then (keeps the queue short)
        MCHK(m, hipMemcpyAsync(part.data(), m->partial, part.size() * 4, hipMemcpyDeviceToHost, m->stream));
        MCHK(m, hipStreamSynchronize(m->stream));
        for (float v : part) se += v;
      }
    }
    (void)se;
    // the loss of the epoch: exact mean over the last batch of every 16 (cheap estimate)
    loss[ep] = (float)(se / ((double)((nb + 15) / 16) * batch * m->n_out));
  }
  MCHK(m, hipEventRecord(e1, m->stream));
  MCHK(m, hipEventSynchronize(e1));
  float ms = 0.0f;
  MCHK(m, hipEventElapsedTime(&ms, e0, e1));
  if (seconds) *seconds = ms * 1e-3;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(drows);
  MCHK(m, hipGetLastError());
  return LGAME_OK;
}

const char *lgame_mlp_last_error(const lgame_mlp *m) { return m ? m->err.c_str() : ""; }

void lgame_mlp_destroy(lgame_mlp *m) {
  if (!m) return;
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  for (auto &Ly : m->L)
    for (void *p : {(void *)Ly.w, (void *)Ly.b, (void *)Ly.mw, (void *)Ly.vw, (void *)Ly.mb, (void *)Ly.vb, (void *)Ly.gw,
                    (void *)Ly.gb, (void *)Ly.wbf})
      if (p) (void)hipFree(p);
  for (auto p : m->Z) (void)hipFree(p);
  for (auto p : m->A) (void)hipFree(p);
  for (void *p : {(void *)m->X, (void *)m->dA, (void *)m->dB, (void *)m->dYb, (void *)m->Yf, (void *)m->Yt, (void *)m->partial,
                  (void *)m->xin, (void *)m->yout, (void *)m->rows, (void *)m->DX, (void *)m->DY, (void *)m->x_mean,
                  (void *)m->x_scale, (void *)m->y_mean, (void *)m->y_scale})
    if (p) (void)hipFree(p);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}

}  // extern "C"
