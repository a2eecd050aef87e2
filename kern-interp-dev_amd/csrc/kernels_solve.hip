// Solve-phase kernels and matrix-free FP64 operators.
//
// Sweeps restate SkelFactorization::solve (skel_factorization.cpp:595-778):
//   up   (levels L-1..0):  x[r] -= T^T x[s];  x[s] -= L x[r]
//   diag (all boxes):      x[r] = X_rr^{-1} x[r]
//   down (levels 0..L-1):  x[r] -= U x[s];   x[s] -= T x[r]
// One workgroup per compressed box; boxes of a level own disjoint DOFs, so a
// level is one launch.  L is stored transposed (r x k) so that both sweeps
// read every factor with coalesced loads.
//
// The matvec is the FP64 residual operator K mu for the whole boundary (the
// reference forms it densely in solve_err, linear_solve.cpp:339-426; here it
// is an N-body loop over LDS-staged source tiles), and diag3d restates the
// row-sum diagonals (kernel.cpp:124-301).
#include <cmath>

#include "ki_internal.h"

namespace ki {

namespace {

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <typename T>
__global__ void __launch_bounds__(256)
sweep_kernel(const SweepDesc* __restrict__ desc, const int* __restrict__ idx,
             T* const* __restrict__ facbufs, T* __restrict__ x, int up) {
  const SweepDesc d = desc[blockIdx.x];
  const int k = d.k, r = d.r;
  extern __shared__ char smem[];
  T* xs = reinterpret_cast<T*>(smem);
  T* xr = xs + k;
  const int* si = idx + d.s_idx;
  const int* ri = idx + d.r_idx;
  const T* F = facbufs[d.buf];
  const T* Tm = F + d.offT;   // k x r
  const T* LT = F + d.offL;   // r x k (L transposed)
  const T* U = F + d.offU;    // r x k
  for (int i = threadIdx.x; i < k; i += 256) xs[i] = x[si[i]];
  for (int j = threadIdx.x; j < r; j += 256) xr[j] = x[ri[j]];
  __syncthreads();
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (up) {
    // x[r] -= T^T x[s]: column j of T dotted with xs (wave per column)
    for (int j = wid; j < r; j += 4) {
      const T* c = Tm + (int64_t)k * j;
      T p = 0;
      for (int i = lane; i < k; i += 64) p += c[i] * xs[i];
      p = wave_sum(p);
      if (lane == 0) xr[j] -= p;
    }
    __syncthreads();
    // x[s] -= L x[r]: (L x_r)_i = column i of L^T dotted with xr
    for (int i = wid; i < k; i += 4) {
      const T* c = LT + (int64_t)r * i;
      T p = 0;
      for (int j = lane; j < r; j += 64) p += c[j] * xr[j];
      p = wave_sum(p);
      if (lane == 0) xs[i] -= p;
    }
  } else {
    // x[r] -= U x[s]: thread per row j, coalesced over j
    for (int j = threadIdx.x; j < r; j += 256) {
      T p = 0;
      for (int i = 0; i < k; i++) p += U[j + (int64_t)r * i] * xs[i];
      xr[j] -= p;
    }
    __syncthreads();
    // x[s] -= T x[r]: thread per row i of T
    for (int i = threadIdx.x; i < k; i += 256) {
      T p = 0;
      for (int j = 0; j < r; j++) p += Tm[i + (int64_t)k * j] * xr[j];
      xs[i] -= p;
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < k; i += 256) x[si[i]] = xs[i];
  for (int j = threadIdx.x; j < r; j += 256) x[ri[j]] = xr[j];
}

// x[r] = X^{-1} x[r] with X = P L U (LAPACK getrf layout), one box per WG.
template <typename T>
__global__ void __launch_bounds__(256)
diag_solve_kernel(const SweepDesc* __restrict__ desc,
                  const int* __restrict__ idx, T* const* __restrict__ facbufs,
                  int* const* __restrict__ pivbufs, T* __restrict__ x) {
  const SweepDesc d = desc[blockIdx.x];
  const int r = d.r;
  if (r == 0) return;
  extern __shared__ char smem[];
  T* xr = reinterpret_cast<T*>(smem);
  const int* ri = idx + d.r_idx;
  const T* X = facbufs[d.buf] + d.offX;
  const int* piv = pivbufs[d.buf] + d.offPiv;
  for (int j = threadIdx.x; j < r; j += 256) xr[j] = x[ri[j]];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 0; q < r; q++) {
      const int p = piv[q];
      if (p != q) {
        const T t = xr[q];
        xr[q] = xr[p];
        xr[p] = t;
      }
    }
  }
  __syncthreads();
  for (int q = 0; q < r; q++) {  // unit lower, column-oriented
    const T v = xr[q];
    const T* c = X + (int64_t)r * q;
    for (int t = q + 1 + threadIdx.x; t < r; t += 256) xr[t] -= c[t] * v;
    __syncthreads();
  }
  for (int q = r - 1; q >= 0; q--) {  // upper
    const T* c = X + (int64_t)r * q;
    const T v = xr[q] / c[q];
    __syncthreads();
    if (threadIdx.x == 0) xr[q] = v;
    for (int t = threadIdx.x; t < q; t += 256) xr[t] -= c[t] * v;
    __syncthreads();
  }
  for (int j = threadIdx.x; j < r; j += 256) x[ri[j]] = xr[j];
}

// ---------------------------------------------------------------------------
// FP64 matrix-free operator y = K mu over the whole boundary.
// ---------------------------------------------------------------------------
constexpr int MT = 256;

// Targets are the boundary points themselves (self = 1, the residual
// operator) or arbitrary domain points with zero normal (self = 0, the
// forward evaluation of Kernel::apply_forward).
__global__ void __launch_bounds__(MT)
matvec_kernel(DevGeom g, const double* __restrict__ tpts,
              const double* __restrict__ tnrm, int ntgt, int self,
              const double* __restrict__ mu, double* __restrict__ y) {
  const double PI = 3.14159265358979323846;
  const int dom = g.dom, sol = g.sol, n = g.npts;
  const int i = blockIdx.x * MT + threadIdx.x;
  __shared__ double sx[MT][3], sn[MT][3], sw[MT], sm[MT][3];
  double tx[3] = {0, 0, 0}, tn[3] = {0, 0, 0};
  if (i < ntgt)
    for (int a = 0; a < dom; a++) {
      tx[a] = tpts[(int64_t)i * dom + a];
      tn[a] = tnrm ? tnrm[(int64_t)i * dom + a] : 0.0;
    }
  double acc[3] = {0, 0, 0};
  for (int j0 = 0; j0 < n; j0 += MT) {
    const int j = j0 + threadIdx.x;
    __syncthreads();
    if (j < n) {
      for (int a = 0; a < 3; a++) {
        sx[threadIdx.x][a] = a < dom ? g.pts[(int64_t)j * dom + a] : 0.0;
        sn[threadIdx.x][a] = a < dom ? g.nrm[(int64_t)j * dom + a] : 0.0;
        sm[threadIdx.x][a] = a < sol ? mu[(int64_t)j * sol + a] : 0.0;
      }
      sw[threadIdx.x] = g.w[j];
    }
    __syncthreads();
    const int cnt = min(MT, n - j0);
    if (i >= ntgt) continue;
    for (int q = 0; q < cnt; q++) {
      if (self && j0 + q == i) continue;
      const double r0 = tx[0] - sx[q][0], r1 = tx[1] - sx[q][1],
                   r2 = tx[2] - sx[q][2];
      const double rr = r0 * r0 + r1 * r1 + r2 * r2;
      const double dot = r0 * sn[q][0] + r1 * sn[q][1] + r2 * sn[q][2];
      const double w = sw[q];
      if (sol == 1) {
        if (dom == 2)
          acc[0] += -w * (1.0 / (2 * PI)) * dot / rr * sm[q][0];
        else {
          const double ir = 1.0 / sqrt(rr);
          acc[0] += -w * (1.0 / (4 * PI)) * dot * (ir * ir * ir) * sm[q][0];
        }
      } else {
        const double rm = r0 * sm[q][0] + r1 * sm[q][1] + r2 * sm[q][2];
        const double nm = sn[q][0] * sm[q][0] + sn[q][1] * sm[q][1] +
                          sn[q][2] * sm[q][2];
        double f;
        if (dom == 2) {
          f = w * (1 / PI) * dot / (rr * rr);
        } else {
          const double ir = 1.0 / sqrt(rr);
          const double ir2 = ir * ir;
          f = w * (-3.0 / (4.0 * PI)) * dot * (ir2 * ir2 * ir);
        }
        const double rv[3] = {r0, r1, r2};
        for (int a = 0; a < sol; a++) acc[a] += f * rv[a] * rm + w * tn[a] * nm;
      }
    }
  }
  if (i >= ntgt) return;
  if (!self) {
    for (int a = 0; a < sol; a++) y[(int64_t)i * sol + a] = acc[a];
    return;
  }
  // self term
  const double w = g.w[i];
  if (sol == 1) {
    double dg;
    if (dom == 2)
      dg = 0.5 + 0.5 * g.curv[i] * w * (1.0 / (2 * PI));
    else
      dg = g.diag[i];
    y[i] = acc[0] + dg * mu[i];
  } else if (dom == 2) {
    const double c = g.curv[i];
    for (int a = 0; a < 2; a++) {
      double s = acc[a];
      for (int b = 0; b < 2; b++) {
        const int isd = (a + b + 1) % 2;
        const double dg = -0.5 * isd + (1 - 2 * isd) * 0.5 * c * w * (1 / PI) *
                              tn[(a + 1) % 2] * tn[(b + 1) % 2] +
                          w * tn[a] * tn[b];
        s += dg * mu[(int64_t)i * 2 + b];
      }
      y[(int64_t)i * 2 + a] = s;
    }
  } else {
    for (int a = 0; a < 3; a++) {
      double s = acc[a];
      for (int b = 0; b < 3; b++)
        s += g.diag[(int64_t)i * 9 + a * 3 + b] * mu[(int64_t)i * 3 + b];
      y[(int64_t)i * 3 + a] = s;
    }
  }
}

__global__ void __launch_bounds__(MT) diag3d_kernel(DevGeom g) {
  const double PI = 3.14159265358979323846;
  const int n = g.npts;
  const int i = blockIdx.x * MT + threadIdx.x;
  __shared__ double sx[MT][3], sn[MT][3], sw[MT];
  double tx[3] = {0, 0, 0}, tn[3] = {0, 0, 0};
  if (i < n)
    for (int a = 0; a < 3; a++) {
      tx[a] = g.pts[(int64_t)i * 3 + a];
      tn[a] = g.nrm[(int64_t)i * 3 + a];
    }
  double s[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  double wn[3] = {0, 0, 0};
  for (int j0 = 0; j0 < n; j0 += MT) {
    const int j = j0 + threadIdx.x;
    __syncthreads();
    if (j < n) {
      for (int a = 0; a < 3; a++) {
        sx[threadIdx.x][a] = g.pts[(int64_t)j * 3 + a];
        sn[threadIdx.x][a] = g.nrm[(int64_t)j * 3 + a];
      }
      sw[threadIdx.x] = g.w[j];
    }
    __syncthreads();
    const int cnt = min(MT, n - j0);
    if (i >= n) continue;
    for (int q = 0; q < cnt; q++) {
      if (j0 + q == i) continue;
      const double r0 = tx[0] - sx[q][0], r1 = tx[1] - sx[q][1],
                   r2 = tx[2] - sx[q][2];
      const double rr = r0 * r0 + r1 * r1 + r2 * r2;
      const double dot = r0 * sn[q][0] + r1 * sn[q][1] + r2 * sn[q][2];
      const double w = sw[q];
      const double ir = 1.0 / sqrt(rr);
      if (g.sol == 1) {
        s[0] += -w * (1.0 / (4 * PI)) * dot * (ir * ir * ir);
      } else {
        const double ir2 = ir * ir;
        const double f = w * (-3.0 / (4.0 * PI)) * dot * (ir2 * ir2 * ir);
        const double rv[3] = {r0, r1, r2};
        for (int a = 0; a < 3; a++)
          for (int b = 0; b < 3; b++) s[a * 3 + b] += f * rv[a] * rv[b];
        for (int b = 0; b < 3; b++) wn[b] += w * sn[q][b];
      }
    }
  }
  if (i >= n) return;
  if (g.sol == 1) {
    g.diag[i] = 1.0 - s[0];
  } else {
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++)
        g.diag[(int64_t)i * 9 + a * 3 + b] =
            (a == b ? 1.0 : 0.0) - (s[a * 3 + b] + tn[a] * wn[b]);
  }
}

template <typename TO, typename TI>
__global__ void convert_kernel(const TI* in, TO* out, int64_t n) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x)
    out[e] = (TO)in[e];
}

template <typename TI, typename TO>
__global__ void gather_kernel(const TI* x, const int* idx, int n, TO* out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) out[e] = (TO)x[idx[e]];
}

template <typename TI, typename TO>
__global__ void scatter_vec_kernel(const TI* in, const int* idx, int n,
                                   TO* x) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) x[idx[e]] = (TO)in[e];
}

}  // namespace

template <typename T>
void launch_sweep(const SweepDesc* d, int nd, const int* idx,
                  T* const* facbufs, T* x, int up, hipStream_t s,
                  int max_n) {
  if (nd == 0) return;
  hipLaunchKernelGGL(sweep_kernel<T>, dim3(nd), dim3(256),
                     sizeof(T) * max_n, s, d, idx, facbufs, x, up);
  KI_HIP(hipGetLastError());
}

template <typename T>
void launch_diag_solve(const SweepDesc* d, int nd, const int* idx,
                       T* const* facbufs, int* const* pivbufs, T* x,
                       hipStream_t s, int max_r) {
  if (nd == 0) return;
  hipLaunchKernelGGL(diag_solve_kernel<T>, dim3(nd), dim3(256),
                     sizeof(T) * max_r, s, d, idx, facbufs, pivbufs, x);
  KI_HIP(hipGetLastError());
}

void launch_matvec(const DevGeom& g, const double* mu, double* y,
                   hipStream_t s) {
  hipLaunchKernelGGL(matvec_kernel, dim3((g.npts + MT - 1) / MT), dim3(MT), 0,
                     s, g, g.pts, g.nrm, g.npts, 1, mu, y);
  KI_HIP(hipGetLastError());
}

void launch_forward(const DevGeom& g, const double* tpts, int ntgt,
                    const double* mu, double* y, hipStream_t s) {
  if (ntgt <= 0) return;
  hipLaunchKernelGGL(matvec_kernel, dim3((ntgt + MT - 1) / MT), dim3(MT), 0,
                     s, g, tpts, nullptr, ntgt, 0, mu, y);
  KI_HIP(hipGetLastError());
}

void launch_diag3d(DevGeom& g, hipStream_t s) {
  hipLaunchKernelGGL(diag3d_kernel, dim3((g.npts + MT - 1) / MT), dim3(MT), 0,
                     s, g);
  KI_HIP(hipGetLastError());
}

template <typename TO, typename TI>
void launch_convert(const TI* in, TO* out, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  const int gx = (int)std::min<int64_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL((convert_kernel<TO, TI>), dim3(gx), dim3(256), 0, s, in,
                     out, n);
  KI_HIP(hipGetLastError());
}

template <typename TI, typename TO>
void launch_gather(const TI* x, const int* idx, int n, TO* out,
                   hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL((gather_kernel<TI, TO>), dim3((n + 255) / 256), dim3(256),
                     0, s, x, idx, n, out);
  KI_HIP(hipGetLastError());
}

template <typename TI, typename TO>
void launch_scatter_vec(const TI* in, const int* idx, int n, TO* x,
                        hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL((scatter_vec_kernel<TI, TO>), dim3((n + 255) / 256),
                     dim3(256), 0, s, in, idx, n, x);
  KI_HIP(hipGetLastError());
}

template void launch_gather<float, double>(const float*, const int*, int,
                                           double*, hipStream_t);
template void launch_gather<double, double>(const double*, const int*, int,
                                            double*, hipStream_t);
template void launch_scatter_vec<double, float>(const double*, const int*,
                                                int, float*, hipStream_t);
template void launch_scatter_vec<double, double>(const double*, const int*,
                                                 int, double*, hipStream_t);

template void launch_sweep<float>(const SweepDesc*, int, const int*,
                                  float* const*, float*, int, hipStream_t,
                                  int);
template void launch_sweep<double>(const SweepDesc*, int, const int*,
                                   double* const*, double*, int, hipStream_t,
                                   int);
template void launch_diag_solve<float>(const SweepDesc*, int, const int*,
                                       float* const*, int* const*, float*,
                                       hipStream_t, int);
template void launch_diag_solve<double>(const SweepDesc*, int, const int*,
                                        double* const*, int* const*, double*,
                                        hipStream_t, int);
template void launch_convert<float, double>(const double*, float*, int64_t,
                                            hipStream_t);
template void launch_convert<double, float>(const float*, double*, int64_t,
                                            hipStream_t);
template void launch_convert<double, double>(const double*, double*, int64_t,
                                             hipStream_t);

}  // namespace ki
