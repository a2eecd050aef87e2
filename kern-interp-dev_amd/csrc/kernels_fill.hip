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

template <typename T>
__global__ void __launch_bounds__(256)
fill_kernel(DevGeom g, ProxySpec px, const FillDesc* __restrict__ desc,
            const int* __restrict__ idx, T* __restrict__ out) {
  const FillDesc& d = desc[blockIdx.y];
  const int j = blockIdx.x;
  if (j >= d.ncols) return;
  const int sol = g.sol;
  const int cdof = idx[d.col_idx + j];
  const int cp = cdof / sol, cpar = cdof % sol;
  Pt cpt;
  load_pt(g, cp, cpt);
  T* col = out + d.out_off + (int64_t)d.ld * j;
  int row0 = 0;
  for (int sgi = 0; sgi < d.nseg; sgi++) {
    const int kind = d.seg_kind[sgi];
    const int nr = d.seg_rows[sgi];
    const int* rl = idx + d.seg_idx[sgi];
    for (int i = threadIdx.x; i < nr; i += blockDim.x) {
      T v;
      if (kind == SEG_K) {
        const int rd = rl[i];
        Pt tp;
        load_pt(g, rd / sol, tp);
        v = entry<T>(g, tp, rd % sol, cpt, cpar, cp, true);
      } else if (kind == SEG_KT) {
        const int rd = rl[i];
        Pt sp;
        load_pt(g, rd / sol, sp);
        v = entry<T>(g, cpt, cpar, sp, rd % sol, rd / sol, true);
      } else {
        const int half = nr / 2;
        Pt pp;
        if (i < half) {  // box -> proxy: target proxy point, source column DOF
          proxy_pt(px, g.dom, d.center, d.radius, i / sol, pp);
          Pt src = cpt;
          src.c = 0.0;
          v = entry<T>(g, pp, i % sol, src, cpar, cp, false);
        } else {         // proxy -> box: target column DOF, source proxy point
          const int q = i - half;
          proxy_pt(px, g.dom, d.center, d.radius, q / sol, pp);
          v = entry<T>(g, cpt, cpar, pp, q % sol, 0, false);
        }
      }
      col[row0 + i] = v;
    }
    row0 += nr;
  }
}

}  // namespace

template <typename T>
void launch_fill(const DevGeom& g, const ProxySpec& px, const FillDesc* d_desc,
                 int ndesc, int max_cols, const int* d_idx, T* out,
                 hipStream_t s) {
  if (ndesc == 0 || max_cols == 0) return;
  KI_CHECK(ndesc <= 65535, 2, "too many blocks in one fill launch");
  dim3 grid(max_cols, ndesc);
  hipLaunchKernelGGL(fill_kernel<T>, grid, dim3(256), 0, s, g, px, d_desc,
                     d_idx, out);
  KI_HIP(hipGetLastError());
}

template void launch_fill<float>(const DevGeom&, const ProxySpec&,
                                 const FillDesc*, int, int, const int*,
                                 float*, hipStream_t);
template void launch_fill<double>(const DevGeom&, const ProxySpec&,
                                  const FillDesc*, int, int, const int*,
                                  double*, hipStream_t);

}  // namespace ki
