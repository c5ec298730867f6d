// PPPM long-range electrostatics of point charges in an orthorhombic periodic box
// (get_p3mlr_fn, deepmd_jax/utils.py:174-222), the long-range term of DPLR.
//
// Charges are spread onto an M0 x M1 x M2 grid with the triangular-shaped cloud (27 points
// per charge, FP64 atomics), the grid goes through a real-to-complex FFT (hipFFT), and
//   E = sum_k K(k) |S(k)|^2,  K = (14.399645 2 pi / V) exp(-k^2 / 4 beta^2) w(k)^2 / (S(k) k^2)
// over the full spectrum (K(0) = 0; w the TSC window prod sinc^3(k h / 2), S its alias sum
// (prod (1 - sin^2 + 2/15 sin^4))^2).  The half spectrum counts every k_z other than 0 and
// M2/2 twice.  The gradient of E with respect to the grid is phi = 2 IFFT(K S) (unnormalised
// complex-to-real transform of the half spectrum), and the force on a charge is
// -q sum over its 27 points of phi times the gradient of the TSC weight.  Everything is FP64
// (the reference evaluates in FP32); the energy sum is a two-pass deterministic reduction.
// The box is read from device memory (it may move on the device under NPT), so the influence
// function is formed for every evaluation.  With charge positions that scale with the box, the
// strain derivative is that of K alone: dK/de_ab = K (-delta_ab + k_a k_b / 2 beta^2 +
// 2 k_a k_b / k^2) (the TSC factors depend on k h, fixed under the scaling), and the kernel also
// sums the reciprocal virial W_ab = -sum_k |S|^2 dK/de_ab.
#include <hipfft/hipfft.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "internal.h"

namespace dpmd {

#define DPMD_FFT(x)                                                        \
  do {                                                                     \
    hipfftResult r_ = (x);                                                 \
    if (r_ != HIPFFT_SUCCESS) {                                            \
      set_error("hipFFT error %d at %s:%d (%s)", (int)r_, __FILE__, __LINE__, #x); \
      return DPMD_E_HIP;                                                   \
    }                                                                      \
  } while (0)

namespace {

constexpr int kPppmBlocks = 256;
constexpr double kCoulomb = 14.399645;  // e^2 / (4 pi eps0) in eV Angstrom

struct PppmGeom {
  int M[3];
  const double* box;  // device: the three box lengths
  double beta;
};

// folded integer wave number of index n on an M-point axis: n for n < M/2, n - M above
// (M even: n = M/2 folds to -M/2), as (n - M/2) mod M - M/2
__device__ __forceinline__ double fold_k(int n, int M) {
  const double h = 0.5 * M;
  double v = fmod((double)n - h, (double)M);
  if (v < 0) v += M;
  return v - h;
}

__global__ void pppm_kfac_kernel(PppmGeom g, double* kfac) {
  const double box[3] = {g.box[0], g.box[1], g.box[2]};
  const int M2h = g.M[2] / 2 + 1;
  const size_t total = (size_t)g.M[0] * g.M[1] * M2h;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int l = (int)(t % M2h);
  const int j = (int)((t / M2h) % g.M[1]);
  const int i = (int)(t / ((size_t)M2h * g.M[1]));
  const int idx[3] = {i, j, l};
  double k2 = 0.0, w3 = 1.0, sk = 1.0;
  for (int d = 0; d < 3; ++d) {
    const double kd = 2.0 * M_PI / box[d] * fold_k(idx[d], g.M[d]);
    k2 += kd * kd;
    const double z = kd * (box[d] / g.M[d]) * 0.5;
    const double sz = sin(z);
    if (z != 0.0) {
      const double sc = sin(z) / z;
      w3 *= sc * sc * sc;
    }
    sk *= 1.0 - sz * sz + 2.0 / 15.0 * sz * sz * sz * sz;
  }
  sk *= sk;
  const double V = box[0] * box[1] * box[2];
  kfac[t] = k2 != 0.0 ? (kCoulomb * 2.0 * M_PI / V) * exp(-k2 / (4.0 * g.beta * g.beta)) * w3 * w3 / (sk * k2) : 0.0;
}

__global__ void pppm_zero_kernel(double* x, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    x[i] = 0.0;
}

// TSC weights of one axis: grid point c + o - 1 (o = 0, 1, 2) with r = u - rint(u)
struct Tsc {
  int c[3];
  double w[3][3], dw[3][3];  // [axis][offset]
};

__device__ __forceinline__ void tsc(const PppmGeom& g, const double* box, const double* x, Tsc& t) {
  for (int d = 0; d < 3; ++d) {
    double xm = fmod(x[d], box[d]);
    if (xm < 0) xm += box[d];
    const double u = xm / box[d] * g.M[d];
    const double c = rint(u);
    const double r = u - c;
    t.c[d] = isfinite(c) ? (int)c : 0;  // a non-finite position (bit 16 ends the run) bins to 0
    t.w[d][0] = 0.5 * (r - 0.5) * (r - 0.5);
    t.w[d][1] = 0.75 - r * r;
    t.w[d][2] = 0.5 * (r + 0.5) * (r + 0.5);
    const double s = g.M[d] / box[d];  // du / dx
    t.dw[d][0] = (r - 0.5) * s;
    t.dw[d][1] = -2.0 * r * s;
    t.dw[d][2] = (r + 0.5) * s;
  }
}

__device__ __forceinline__ int wrap_idx(int v, int M) {
  v %= M;
  return v < 0 ? v + M : v;
}

__global__ void pppm_spread_kernel(PppmGeom g, const double* xyz, const double* q, int n, double* grid) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Tsc t;
  const double box[3] = {g.box[0], g.box[1], g.box[2]};
  tsc(g, box, xyz + (size_t)i * 3, t);
  const double qi = q[i];
  for (int a = 0; a < 3; ++a) {
    const int ix = wrap_idx(t.c[0] + a - 1, g.M[0]);
    for (int b = 0; b < 3; ++b) {
      const int iy = wrap_idx(t.c[1] + b - 1, g.M[1]);
      const double wab = t.w[0][a] * t.w[1][b] * qi;
      for (int c = 0; c < 3; ++c) {
        const int iz = wrap_idx(t.c[2] + c - 1, g.M[2]);
        atomicAdd(&grid[((size_t)ix * g.M[1] + iy) * g.M[2] + iz], wab * t.w[2][c]);
      }
    }
  }
}

// per block: sums over the half spectrum of w K |S|^2 (energy) and of the reciprocal virial
// W_ab (00, 11, 22, 01, 02, 12); S <- 2 K S for the potential.  part[7 block + c]
__global__ void pppm_energy_kernel(PppmGeom g, hipfftDoubleComplex* spec, const double* kfac, double* part) {
  __shared__ double red[256];
  const double box[3] = {g.box[0], g.box[1], g.box[2]};
  const double ib2 = 1.0 / (2.0 * g.beta * g.beta);
  const int M2h = g.M[2] / 2 + 1;
  const size_t total = (size_t)g.M[0] * g.M[1] * M2h;
  double acc[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
    const int l = (int)(t % M2h);
    const int j = (int)((t / M2h) % g.M[1]);
    const int i = (int)(t / ((size_t)M2h * g.M[1]));
    const double w = (l == 0 || (g.M[2] % 2 == 0 && l == g.M[2] / 2)) ? 1.0 : 2.0;
    const double K = kfac[t];
    const hipfftDoubleComplex s = spec[t];
    const double e = w * K * (s.x * s.x + s.y * s.y);
    acc[0] += e;
    if (K != 0.0) {
      const double k0 = 2.0 * M_PI / box[0] * fold_k(i, g.M[0]);
      const double k1 = 2.0 * M_PI / box[1] * fold_k(j, g.M[1]);
      const double k2v = 2.0 * M_PI / box[2] * fold_k(l, g.M[2]);
      const double f = ib2 + 2.0 / (k0 * k0 + k1 * k1 + k2v * k2v);
      acc[1] -= e * (-1.0 + k0 * k0 * f);
      acc[2] -= e * (-1.0 + k1 * k1 * f);
      acc[3] -= e * (-1.0 + k2v * k2v * f);
      acc[4] -= e * k0 * k1 * f;
      acc[5] -= e * k0 * k2v * f;
      acc[6] -= e * k1 * k2v * f;
    }
    spec[t].x = 2.0 * K * s.x;
    spec[t].y = 2.0 * K * s.y;
  }
  for (int c = 0; c < 7; ++c) {
    red[threadIdx.x] = acc[c];
    __syncthreads();
    for (int k = blockDim.x / 2; k > 0; k >>= 1) {
      if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x * 7 + c] = red[0];
    __syncthreads();
  }
}

// energy and the 3 x 3 reciprocal virial (row-major) from the block partials
__global__ void pppm_sum_kernel(const double* part, int nb, double* out) {
  const int c = threadIdx.x;
  if (c >= 7 || blockIdx.x != 0) return;
  double s = 0.0;
  for (int b = 0; b < nb; ++b) s += part[b * 7 + c];
  const int dst[7][2] = {{0, -1}, {1, -1}, {5, -1}, {9, -1}, {2, 4}, {3, 7}, {6, 8}};
  out[dst[c][0]] = s;
  if (dst[c][1] >= 0) out[dst[c][1]] = s;
}

// F_i = -q_i sum_{27 points} phi(m) grad W(x_i, m); added to force (n x 3, FP64)
__global__ void pppm_gather_kernel(PppmGeom g, const double* xyz, const double* q, int n, const double* phi,
                                   double* force) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Tsc t;
  const double box[3] = {g.box[0], g.box[1], g.box[2]};
  tsc(g, box, xyz + (size_t)i * 3, t);
  double f0 = 0.0, f1 = 0.0, f2 = 0.0;
  for (int a = 0; a < 3; ++a) {
    const int ix = wrap_idx(t.c[0] + a - 1, g.M[0]);
    for (int b = 0; b < 3; ++b) {
      const int iy = wrap_idx(t.c[1] + b - 1, g.M[1]);
      for (int c = 0; c < 3; ++c) {
        const int iz = wrap_idx(t.c[2] + c - 1, g.M[2]);
        const double p = phi[((size_t)ix * g.M[1] + iy) * g.M[2] + iz];
        f0 += p * t.dw[0][a] * t.w[1][b] * t.w[2][c];
        f1 += p * t.w[0][a] * t.dw[1][b] * t.w[2][c];
        f2 += p * t.w[0][a] * t.w[1][b] * t.dw[2][c];
      }
    }
  }
  const double qi = q[i];
  force[(size_t)i * 3 + 0] -= qi * f0;
  force[(size_t)i * 3 + 1] -= qi * f1;
  force[(size_t)i * 3 + 2] -= qi * f2;
}

}  // namespace

Pppm::~Pppm() { release(); }

void Pppm::release() {
  if (fwd) hipfftDestroy(fwd);
  if (inv) hipfftDestroy(inv);
  fwd = inv = 0;
  for (void* p : {(void*)grid, (void*)spec, (void*)kfac, (void*)part})
    if (p) hipFree(p);
  grid = nullptr;
  spec = nullptr;
  kfac = nullptr;
  part = nullptr;
  M[0] = M[1] = M[2] = 0;
}

int32_t Pppm::setup(const int* grid_size, double beta_) {
  for (int d = 0; d < 3; ++d)
    if (grid_size[d] < 2 || grid_size[d] > 4096) {
      set_error("PPPM grid size %d outside [2, 4096]", grid_size[d]);
      return DPMD_E_ARG;
    }
  if (!(beta_ > 0)) { set_error("PPPM needs beta > 0"); return DPMD_E_ARG; }
  beta = beta_;
  if (grid_size[0] == M[0] && grid_size[1] == M[1] && grid_size[2] == M[2]) return DPMD_OK;
  release();
  for (int d = 0; d < 3; ++d) M[d] = grid_size[d];
  const size_t nr = (size_t)M[0] * M[1] * M[2], nc = (size_t)M[0] * M[1] * (M[2] / 2 + 1);
  DPMD_HIP(hipMalloc(&grid, sizeof(double) * nr));
  DPMD_HIP(hipMalloc(&spec, sizeof(hipfftDoubleComplex) * nc));  // void* member
  DPMD_HIP(hipMalloc(&kfac, sizeof(double) * nc));
  DPMD_HIP(hipMalloc(&part, sizeof(double) * 7 * kPppmBlocks));
  DPMD_FFT(hipfftPlan3d(&fwd, M[0], M[1], M[2], HIPFFT_D2Z));
  DPMD_FFT(hipfftPlan3d(&inv, M[0], M[1], M[2], HIPFFT_Z2D));
  return DPMD_OK;
}

int32_t Pppm::compute(hipStream_t st, const double* d_box, const double* xyz, const double* q, int n,
                      double* d_out, double* force) {
  if (!fwd) { set_error("PPPM used before setup"); return DPMD_E_STATE; }
  PppmGeom g;
  for (int d = 0; d < 3; ++d) g.M[d] = M[d];
  g.box = d_box;
  g.beta = beta;
  const size_t nr = (size_t)M[0] * M[1] * M[2];
  const size_t nc = (size_t)M[0] * M[1] * (M[2] / 2 + 1);
  hipLaunchKernelGGL(pppm_kfac_kernel, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, st, g, kfac);
  hipLaunchKernelGGL(pppm_zero_kernel, dim3((unsigned)std::min<size_t>(4096, (nr + 255) / 256)), dim3(256), 0, st,
                     grid, nr);
  if (n > 0)
    hipLaunchKernelGGL(pppm_spread_kernel, dim3((n + 127) / 128), dim3(128), 0, st, g, xyz, q, n, grid);
  DPMD_FFT(hipfftSetStream(fwd, st));
  DPMD_FFT(hipfftSetStream(inv, st));
  hipfftDoubleComplex* sp = static_cast<hipfftDoubleComplex*>(spec);
  DPMD_FFT(hipfftExecD2Z(fwd, grid, sp));
  hipLaunchKernelGGL(pppm_energy_kernel, dim3(kPppmBlocks), dim3(256), 0, st, g, sp, kfac, part);
  hipLaunchKernelGGL(pppm_sum_kernel, dim3(1), dim3(64), 0, st, part, kPppmBlocks, d_out);
  DPMD_FFT(hipfftExecZ2D(inv, sp, grid));
  if (n > 0 && force)
    hipLaunchKernelGGL(pppm_gather_kernel, dim3((n + 127) / 128), dim3(128), 0, st, g, xyz, q, n, grid, force);
  DPMD_HIP(hipGetLastError());
  return DPMD_OK;
}

}  // namespace dpmd

using namespace dpmd;

extern "C" int32_t dpmd_pppm_compute(int32_t device, const double* box3, double beta, const int32_t* grid_size,
                                     int32_t n, const double* xyz, const double* q, double* energy,
                                     double* force, double* virial) {
  if (!box3 || !grid_size || !energy || (n > 0 && (!xyz || !q)) || n < 0) {
    set_error("null argument");
    return DPMD_E_ARG;
  }
  for (int d = 0; d < 3; ++d)
    if (!(box3[d] > 0)) { set_error("box lengths must be positive"); return DPMD_E_ARG; }
  DPMD_HIP(hipSetDevice(device));
  hipStream_t st;
  DPMD_HIP(hipStreamCreate(&st));
  Pppm p;
  double *d_x = nullptr, *d_q = nullptr, *d_f = nullptr, *d_e = nullptr, *d_b = nullptr;
  int32_t rc = DPMD_OK;
  const size_t nn = std::max(n, 1);
  const int gs[3] = {grid_size[0], grid_size[1], grid_size[2]};
  do {
    if (hipMalloc(&d_x, sizeof(double) * 3 * nn) != hipSuccess || hipMalloc(&d_q, sizeof(double) * nn) != hipSuccess ||
        hipMalloc(&d_f, sizeof(double) * 3 * nn) != hipSuccess || hipMalloc(&d_e, sizeof(double) * 10) != hipSuccess ||
        hipMalloc(&d_b, sizeof(double) * 3) != hipSuccess) {
      set_error("PPPM: device allocation failed");
      rc = DPMD_E_HIP;
      break;
    }
    if (n > 0) {
      hipMemcpyAsync(d_x, xyz, sizeof(double) * 3 * n, hipMemcpyHostToDevice, st);
      hipMemcpyAsync(d_q, q, sizeof(double) * n, hipMemcpyHostToDevice, st);
    }
    hipMemsetAsync(d_f, 0, sizeof(double) * 3 * nn, st);
    hipMemcpyAsync(d_b, box3, sizeof(double) * 3, hipMemcpyHostToDevice, st);
    if ((rc = p.setup(gs, beta)) != DPMD_OK) break;
    if ((rc = p.compute(st, d_b, d_x, d_q, n, d_e, d_f)) != DPMD_OK) break;
    hipMemcpyAsync(energy, d_e, sizeof(double), hipMemcpyDeviceToHost, st);
    if (virial) hipMemcpyAsync(virial, d_e + 1, sizeof(double) * 9, hipMemcpyDeviceToHost, st);
    if (force && n > 0) hipMemcpyAsync(force, d_f, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, st);
    if (hipStreamSynchronize(st) != hipSuccess) {
      set_error("PPPM: stream synchronisation failed");
      rc = DPMD_E_HIP;
    }
  } while (0);
  for (void* ptr_ : {(void*)d_x, (void*)d_q, (void*)d_f, (void*)d_e, (void*)d_b})
    if (ptr_) hipFree(ptr_);
  p.release();
  hipStreamDestroy(st);
  return rc;
}
