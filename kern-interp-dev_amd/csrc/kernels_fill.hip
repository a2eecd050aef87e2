// Green's-function block generator.
//
// One launch fills every block of a tree level: each workgroup owns one
// column of one block and walks its rows, so stores are coalesced (column-
// major output) and the column's source/target point is a wavefront-uniform
// scalar load.  Entry formulas restate the reference's kernels
// (kernel/kernel.cpp:33-121): 2D Laplace / Stokes double layer with the
// curvature self term, 3D Laplace / Stokes double layer whose self term comes
// from the precomputed row-sum diagonal (kernel.cpp:124-301), and the
// Stokes rank-one completion w n_t n_s^T.  Proxy rows follow
// kernel.cpp:646-796 (box->proxy rows first, then proxy->box rows).
//
// Precision: the displacement r = t - s is always formed in FP64 (point
// coordinates are O(1) and neighbours are close, so an FP32 subtraction would
// lose digits); everything after it runs in the output type T.
#include <cmath>

#include "ki_internal.h"

namespace ki {

namespace {

struct Pt {
  double x[3];
  double n[3];
  double w, c;
};

__device__ __forceinline__ void load_pt(const DevGeom& g, int p, Pt& o) {
  const int d = g.dom;
  for (int i = 0; i < 3; i++) {
    o.x[i] = i < d ? g.pts[(int64_t)p * d + i] : 0.0;
    o.n[i] = i < d ? g.nrm[(int64_t)p * d + i] : 0.0;
  }
  o.w = g.w[p];
  o.c = g.curv ? g.curv[p] : 0.0;
}

__device__ __forceinline__ void proxy_pt(const ProxySpec& px, int dom,
                                         const double* c, double R, int q,
                                         Pt& o) {
  const double PI = 3.14159265358979323846;
  o.c = 0.0;
  if (dom == 2) {
    const double ang = 2 * PI * q * (1.0 / px.npts);
    const double ca = cos(ang), sa = sin(ang);
    o.x[0] = c[0] + R * ca;
    o.x[1] = c[1] + R * sa;
    o.x[2] = 0;
    o.n[0] = ca;
    o.n[1] = sa;
    o.n[2] = 0;
    o.w = 2.0 * PI * R / px.npts;
    return;
  }
  if (px.kind == 0) {
    const int i = q / px.nphi, j = q % px.nphi;
    const double th = 2 * PI * i * (1.0 / px.ntheta);
    const double ph = px.gl_phi[j];
    const double sp = sin(ph), cp = cos(ph), ct = cos(th), st = sin(th);
    o.x[0] = c[0] + R * sp * ct;
    o.x[1] = c[1] + R * sp * st;
    o.x[2] = c[2] + R * cp;
    o.n[0] = sp * ct;
    o.n[1] = sp * st;
    o.n[2] = cp;
    o.w = R * R * (px.gl_w[j] * sp) * (2 * PI / px.ntheta);
  } else {
    const double z = 1.0 - (2.0 * q + 1.0) / px.npts;
    const double rho = sqrt(fmax(0.0, 1.0 - z * z));
    const double th = PI * (3.0 - sqrt(5.0)) * q;
    const double u0 = rho * cos(th), u1 = rho * sin(th);
    o.x[0] = c[0] + R * u0;
    o.x[1] = c[1] + R * u1;
    o.x[2] = c[2] + R * z;
    o.n[0] = u0;
    o.n[1] = u1;
    o.n[2] = z;
    o.w = 4.0 * PI * R * R / px.npts;
  }
}

// K(target, source) for one (row component a, column component b) pair.
template <typename T>
__device__ __forceinline__ T entry(const DevGeom& g, const Pt& t, int a,
                                   const Pt& s, int b, int sidx, bool self_ok) {
  const double PI = 3.14159265358979323846;
  const double r0 = t.x[0] - s.x[0], r1 = t.x[1] - s.x[1],
               r2 = t.x[2] - s.x[2];
  const bool zero = self_ok && r0 == 0.0 && r1 == 0.0 && r2 == 0.0;
  if (g.dom == 2) {
    if (g.sol == 1) {
      if (zero) return (T)(0.5 + 0.5 * s.c * s.w * (1.0 / (2 * PI)));
      const T x0 = (T)r0, x1 = (T)r1;
      const T dot = x0 * (T)s.n[0] + x1 * (T)s.n[1];
      return -(T)s.w * (T)(1.0 / (2 * PI)) * dot / (x0 * x0 + x1 * x1);
    }
    if (zero) {
      const int isd = (a + b + 1) % 2;
      return (T)(-0.5 * isd + (1 - 2 * isd) * 0.5 * s.c * s.w * (1 / PI) *
                                  s.n[(a + 1) % 2] * s.n[(b + 1) % 2] +
                 s.w * t.n[a] * s.n[b]);
    }
    const T x0 = (T)r0, x1 = (T)r1;
    const T rr = x0 * x0 + x1 * x1;
    const T dot = x0 * (T)s.n[0] + x1 * (T)s.n[1];
    const T ra = a == 0 ? x0 : x1, rb = b == 0 ? x0 : x1;
    return (T)s.w * (T)(1 / PI) * dot / (rr * rr) * ra * rb +
           (T)s.w * (T)t.n[a] * (T)s.n[b];
  }
  if (g.sol == 1) {
    if (zero) return (T)g.diag[sidx];
    const T x0 = (T)r0, x1 = (T)r1, x2 = (T)r2;
    const T rr = x0 * x0 + x1 * x1 + x2 * x2;
    const T dot = x0 * (T)s.n[0] + x1 * (T)s.n[1] + x2 * (T)s.n[2];
    const T ir = (T)1 / sqrt(rr);
    return -(T)s.w * (T)(1.0 / (4 * PI)) * dot * (ir * ir * ir);
  }
  if (zero) return (T)g.diag[(int64_t)sidx * 9 + a * 3 + b];
  const T x0 = (T)r0, x1 = (T)r1, x2 = (T)r2;
  const T rr = x0 * x0 + x1 * x1 + x2 * x2;
  const T dot = x0 * (T)s.n[0] + x1 * (T)s.n[1] + x2 * (T)s.n[2];
  const T ir = (T)1 / sqrt(rr);
  const T ir2 = ir * ir;
  const T ra = a == 0 ? x0 : (a == 1 ? x1 : x2);
  const T rb = b == 0 ? x0 : (b == 1 ? x1 : x2);
  return (T)s.w * (T)(-3.0 / (4.0 * PI)) * dot * (ir2 * ir2 * ir) * ra * rb +
         (T)s.w * (T)t.n[a] * (T)s.n[b];
}

// One workgroup fills a tile of FR rows x FC columns of one block: each
// thread resolves its row once (segment, DOF, point or proxy point), the
// tile's column points sit in LDS, and the thread then produces its row's
// FC entries; for every column the 256 threads store 256 consecutive rows.
constexpr int FR = 256, FC = 16;

template <typename T>
__global__ void __launch_bounds__(FR)
fill_kernel(DevGeom g, ProxySpec px, const FillDesc* __restrict__ desc,
            const int* __restrict__ idx, T* __restrict__ out) {
  const FillDesc& d = desc[blockIdx.z];
  const int c0 = blockIdx.x * FC;
  if (c0 >= d.ncols) return;
  int m = 0;
  for (int sg = 0; sg < d.nseg; sg++) m += d.seg_rows[sg];
  const int i = blockIdx.y * FR + threadIdx.x;
  if ((int)blockIdx.y * FR >= m) return;
  const int sol = g.sol;
  const int nc = min(FC, d.ncols - c0);
  __shared__ Pt cpts[FC];
  __shared__ int cdofs[FC];
  if ((int)threadIdx.x < nc) {
    const int cd = idx[d.col_idx + c0 + threadIdx.x];
    cdofs[threadIdx.x] = cd;
    load_pt(g, cd / sol, cpts[threadIdx.x]);
  }
  __syncthreads();
  if (i >= m) return;
  int sg = 0, base = 0;
  while (i >= base + d.seg_rows[sg]) base += d.seg_rows[sg++];
  const int kind = d.seg_kind[sg], ii = i - base, nr = d.seg_rows[sg];
  Pt rp;
  int rcomp, rpt = 0;
  bool to_proxy = false;
  if (kind == SEG_K || kind == SEG_KT) {
    const int rd = idx[d.seg_idx[sg] + ii];
    rpt = rd / sol;
    rcomp = rd % sol;
    load_pt(g, rpt, rp);
  } else {
    const int half = nr / 2;
    to_proxy = ii < half;
    const int q = to_proxy ? ii : ii - half;
    proxy_pt(px, g.dom, d.center, d.radius, q / sol, rp);
    rcomp = q % sol;
  }
  T* row = out + d.out_off + i;
  for (int jj = 0; jj < nc; jj++) {
    const int cd = cdofs[jj];
    const int cpi = cd / sol, cpar = cd % sol;
    T v;
    if (kind == SEG_K) {
      v = entry<T>(g, rp, rcomp, cpts[jj], cpar, cpi, true);
    } else if (kind == SEG_KT) {
      v = entry<T>(g, cpts[jj], cpar, rp, rcomp, rpt, true);
    } else if (to_proxy) {  // box -> proxy: target proxy, source column DOF
      Pt src = cpts[jj];
      src.c = 0.0;
      v = entry<T>(g, rp, rcomp, src, cpar, cpi, false);
    } else {                // proxy -> box: target column DOF, source proxy
      v = entry<T>(g, cpts[jj], cpar, rp, rcomp, 0, false);
    }
    row[(int64_t)d.ld * (c0 + jj)] = v;
  }
}

}  // namespace

template <typename T>
void launch_fill(const DevGeom& g, const ProxySpec& px, const FillDesc* d_desc,
                 int ndesc, int max_cols, int max_rows, const int* d_idx,
                 T* out, hipStream_t s) {
  if (ndesc == 0 || max_cols == 0 || max_rows == 0) return;
  KI_CHECK(ndesc <= 65535, 2, "too many blocks in one fill launch");
  KI_CHECK((max_rows + FR - 1) / FR <= 65535, 2, "block too tall to fill");
  dim3 grid((max_cols + FC - 1) / FC, (max_rows + FR - 1) / FR, ndesc);
  hipLaunchKernelGGL(fill_kernel<T>, grid, dim3(FR), 0, s, g, px, d_desc,
                     d_idx, out);
  KI_HIP(hipGetLastError());
}

template void launch_fill<float>(const DevGeom&, const ProxySpec&,
                                 const FillDesc*, int, int, int, const int*,
                                 float*, hipStream_t);
template void launch_fill<double>(const DevGeom&, const ProxySpec&,
                                  const FillDesc*, int, int, int, const int*,
                                  double*, hipStream_t);

}  // namespace ki
