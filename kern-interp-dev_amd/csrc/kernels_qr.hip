// Batched column-pivoted Householder QR for the interpolative decomposition.
//
// Restates the ID of the reference (ki_mat.cpp:525-568): LAPACK dgeqp3
// followed by "rank k = first i >= 1 with |R_ii| < tol |R_00|" (0 when no
// such i: the box stays uncompressed).  The pivoting is dgeqp3's: pick the
// column of largest partial norm (first index on ties, as idamax), keep the
// norms current by downdating with the pivot row, and recompute a norm from
// the column when the downdate has lost too many digits (the dlaqp2 guard
// temp2 <= sqrt(eps)).
//
// One workgroup factors one box; a whole tree level is one launch.  Column
// dot products and norms are wavefront reductions; each wavefront owns a
// column at a time and streams it with coalesced loads.  In "early" mode the
// factorization stops at the step whose |R_ii| falls under the threshold
// (steps past the rank cannot change the skeleton, R11 or R12); "full" mode
// runs every step like dgeqp3 and is kept as the faithful plain method.
#include <cfloat>
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
struct Eps;
template <>
struct Eps<float> {
  static constexpr float v = FLT_EPSILON * 0.5f;
};
template <>
struct Eps<double> {
  static constexpr double v = DBL_EPSILON * 0.5;
};

// Block-wide sum; every thread gets the result.
template <typename T, int NT>
__device__ __forceinline__ T block_sum(T v, T* red) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  T s = 0;
#pragma unroll
  for (int w = 0; w < NT / 64; w++) s += red[w];
  return s;
}

// Block-wide argmax of x[lo..hi) (first index among equal maxima).
template <typename T, int NT>
__device__ __forceinline__ int block_argmax(const T* x, int lo, int hi,
                                            T* redv, int* redi) {
  T best = -1;
  int bi = hi;
  for (int j = lo + threadIdx.x; j < hi; j += NT) {
    const T v = x[j];
    if (v > best) {
      best = v;
      bi = j;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    const T ov = __shfl_xor(best, o, 64);
    const int oi = __shfl_xor(bi, o, 64);
    if (ov > best || (ov == best && oi < bi)) {
      best = ov;
      bi = oi;
    }
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) {
    redv[wid] = best;
    redi[wid] = bi;
  }
  __syncthreads();
  T b = redv[0];
  int i = redi[0];
  for (int w = 1; w < NT / 64; w++) {
    if (redv[w] > b || (redv[w] == b && redi[w] < i)) {
      b = redv[w];
      i = redi[w];
    }
  }
  return i;
}

template <typename T, int NT>
__global__ void __launch_bounds__(NT)
cpqr_wg(const QrDesc* __restrict__ desc, T* __restrict__ a_all,
        int* __restrict__ piv_all, T* __restrict__ ws_all,
        int* __restrict__ rank_out, double tol, int full) {
  const QrDesc d = desc[blockIdx.x];
  const int m = d.m, n = d.n;
  T* A = a_all + d.a_off;
  int* piv = piv_all + d.piv_off;
  T* vn1 = ws_all + d.ws_off;
  T* vn2 = vn1 + n;
  __shared__ T red[NT / 64];
  __shared__ int redi[NT / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  constexpr int NW = NT / 64;
  const T tol3z = sqrt(Eps<T>::v);

  for (int j = wid; j < n; j += NW) {
    const T* c = A + (int64_t)m * j;
    T s = 0;
    for (int i = lane; i < m; i += 64) s += c[i] * c[i];
    s = wave_sum(s);
    if (lane == 0) {
      vn1[j] = sqrt(s);
      vn2[j] = vn1[j];
    }
  }
  for (int j = threadIdx.x; j < n; j += NT) piv[j] = j;
  __syncthreads();

  const int kmax = min(min(m, n), d.max_rank);
  int rank = 0;
  T thresh = 0;
  for (int i = 0; i < kmax; i++) {
    int p = block_argmax<T, NT>(vn1, i, n, red, redi);
    if (p < i || p >= n) p = i;  // non-finite norms: keep the column order
    if (p != i) {
      T* ci = A + (int64_t)m * i;
      T* cp = A + (int64_t)m * p;
      for (int r = threadIdx.x; r < m; r += NT) {
        const T t = ci[r];
        ci[r] = cp[r];
        cp[r] = t;
      }
      if (threadIdx.x == 0) {
        T t = vn1[i];
        vn1[i] = vn1[p];
        vn1[p] = t;
        t = vn2[i];
        vn2[i] = vn2[p];
        vn2[p] = t;
        const int ti = piv[i];
        piv[i] = piv[p];
        piv[p] = ti;
      }
    }
    __syncthreads();
    // Householder reflector for column i, rows i..m-1 (dlarfg).
    T* v = A + (int64_t)m * i;
    const T alpha = v[i];
    T s = 0;
    for (int r = i + 1 + threadIdx.x; r < m; r += NT) s += v[r] * v[r];
    const T xnorm = sqrt(block_sum<T, NT>(s, red));
    T beta, tau;
    if (xnorm == 0) {
      tau = 0;
      beta = alpha;
    } else {
      beta = -copysign(sqrt(alpha * alpha + xnorm * xnorm), alpha);
      tau = (beta - alpha) / beta;
      const T scal = (T)1 / (alpha - beta);
      for (int r = i + 1 + threadIdx.x; r < m; r += NT) v[r] *= scal;
    }
    // Rank rule of the reference: thresh = |tol * R00|; first i >= 1 with
    // |R_ii| < thresh is the rank.
    bool stop = false;
    if (i == 0) {
      thresh = fabs((T)tol * beta);
    } else if (rank == 0 && fabs(beta) < thresh) {
      rank = i;
      stop = !full;
    }
    __syncthreads();
    if (threadIdx.x == 0) v[i] = beta;
    if (stop) break;
    // Apply H = I - tau v v^T to columns i+1..n-1 and downdate their norms.
    for (int j = i + 1 + wid; j < n; j += NW) {
      T* c = A + (int64_t)m * j;
      const T ci = c[i];
      T part = 0;
      for (int r = i + 1 + lane; r < m; r += 64) part += v[r] * c[r];
      const T w = tau * (ci + wave_sum(part));
      const T cnew = ci - w;
      T s2 = 0;
      for (int r = i + 1 + lane; r < m; r += 64) {
        const T u = c[r] - w * v[r];
        c[r] = u;
        s2 += u * u;
      }
      if (lane == 0) c[i] = cnew;
      const T v1 = vn1[j];
      if (v1 != 0) {
        T t = fabs(cnew) / v1;
        t = (T)1 - t * t;
        t = t > 0 ? t : 0;
        const T q = v1 / vn2[j];
        const T t2 = t * q * q;
        const T nn = sqrt(wave_sum(s2));
        if (lane == 0) {
          if (t2 <= tol3z) {
            vn1[j] = i + 1 < m ? nn : 0;
            vn2[j] = vn1[j];
          } else {
            vn1[j] = v1 * sqrt(t);
          }
        }
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) rank_out[blockIdx.x] = rank;
}

// T = R11^{-1} R12, in place over R12 (one thread per column of R12).
template <typename T>
__global__ void __launch_bounds__(256)
idsolve_kernel(const QrDesc* __restrict__ desc, T* __restrict__ a_all,
               const int* __restrict__ rank) {
  const QrDesc d = desc[blockIdx.y];
  const int k = rank[blockIdx.y];
  const int m = d.m, n = d.n;
  const int j = k + blockIdx.x * blockDim.x + threadIdx.x;
  if (k <= 0 || j >= n) return;
  T* A = a_all + d.a_off;
  T* b = A + (int64_t)m * j;
  for (int ii = k - 1; ii >= 0; ii--) {
    T x = b[ii];
    for (int l = ii + 1; l < k; l++) x -= A[ii + (int64_t)m * l] * b[l];
    b[ii] = x / A[ii + (int64_t)m * ii];
  }
}

}  // namespace

template <typename T>
void launch_cpqr(const QrDesc* d_desc, int ndesc, int max_n, T* a, int* piv,
                 T* ws, int* rank, double tol, hipStream_t s) {
  (void)max_n;
  if (ndesc == 0) return;
  hipLaunchKernelGGL((cpqr_wg<T, 256>), dim3(ndesc), dim3(256), 0, s, d_desc,
                     a, piv, ws, rank, tol, 0);
  KI_HIP(hipGetLastError());
}

template <typename T>
void launch_idsolve(const QrDesc* d_desc, int ndesc, int max_n, T* a,
                    const int* rank, hipStream_t s) {
  if (ndesc == 0 || max_n == 0) return;
  dim3 grid((max_n + 255) / 256, ndesc);
  hipLaunchKernelGGL(idsolve_kernel<T>, grid, dim3(256), 0, s, d_desc, a,
                     rank);
  KI_HIP(hipGetLastError());
}

template void launch_cpqr<float>(const QrDesc*, int, int, float*, int*,
                                 float*, int*, double, hipStream_t);
template void launch_cpqr<double>(const QrDesc*, int, int, double*, int*,
                                  double*, int*, double, hipStream_t);
template void launch_idsolve<float>(const QrDesc*, int, int, float*,
                                    const int*, hipStream_t);
template void launch_idsolve<double>(const QrDesc*, int, int, double*,
                                     const int*, hipStream_t);

}  // namespace ki
