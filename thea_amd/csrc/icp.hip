// Iterative closest point on the GPU: Thea's Icp3<double>::align (Algorithms/Icp3.hpp:193-263) for
// K problems that share one target BVH (ShapeAlign tries up to 24 starting rotations of the same
// source points, Tools/ShapeAlign/ShapeAlign.cpp:343-414).
//
// Per iteration and problem, as Thea does:
//   inc  = alignOneStep(from_points, to_points)      means, cross-covariance, 3x3 SVD (Icp3.hpp:288-375)
//   new  = inc * old
//   err  = measureError(inc, from_points, to)        to_points = nearest(from_points), then the error
//                                                    of inc on that mapping (Icp3.hpp:378-465)
//   from_points = new * from                         (only if another iteration follows)
// so each fit uses the correspondences of the positions one iteration earlier, exactly as in Thea.
//
// Device side: the nearest points of all active problems' positions come from one closest-point
// launch. On a double point BVH (ShapeAlign's BvhN<Vector3d, 3, double>, ShapeAlign.cpp:339) the
// positions stay double and the search is in double, as in Thea; on a float BVH (a mesh, or float
// points) the positions are rounded to float and the correspondences are the float closest points.
// Means, covariance and error are double sums over each problem's points,
// reduced in a fixed order (per-thread sequential sums of every 256th point, an LDS tree per block,
// then the blocks in order), so a problem's result does not depend on the other problems in the batch. The 3x3 SVD and
// the convergence logic run on the host.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

#include "common.hpp"
#include "kernels_common.hpp"
#include "query.hpp"
#include "query64.hpp"

namespace thg {

constexpr int ICP_THREADS = 256;
constexpr int ICP_PER_THREAD = 16;
constexpr int64_t ICP_CHUNK = (int64_t)ICP_THREADS * ICP_PER_THREAD;  // points per block
constexpr int ICP_MAXC = 9;                                            // sums per block at most

// Row-major 3x4 transform [L | t] in double.
struct Xf {
  double m[12];
};

__host__ __device__ inline void xf_apply(const double *x, double px, double py, double pz, double &ox, double &oy,
                                         double &oz) {
  // L p + t, each row as ((L_i0 p0 + L_i1 p1) + L_i2 p2) + t_i (no contraction)
  ox = ((x[0] * px + x[1] * py) + x[2] * pz) + x[3];
  oy = ((x[4] * px + x[5] * py) + x[6] * pz) + x[7];
  oz = ((x[8] * px + x[9] * py) + x[10] * pz) + x[11];
}

// src[k n + j] = init[k] * from[j] (the source points of problem k, as the caller would transform
// them before calling Icp3::align).
__global__ void __launch_bounds__(256) k_icp_init(int64_t n, int K, const double *__restrict__ from,
                                                  const double *__restrict__ init, double *__restrict__ src) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n * K) return;
  const int k = (int)(i / n);
  const int64_t j = i - (int64_t)k * n;
  double x = from[3 * j], y = from[3 * j + 1], z = from[3 * j + 2];
  if (init) xf_apply(init + 12 * k, x, y, z, x, y, z);
  src[3 * i] = x;
  src[3 * i + 1] = y;
  src[3 * i + 2] = z;
}

// Positions of the active problems (slot s is problem prob[s], its points [s n, (s + 1) n)):
// pos = cur[k] * src[k] in double, rounded to T (float for a float BVH, double for a double one).
template <typename T>
__global__ void __launch_bounds__(256) k_icp_pos(int64_t n, int A, const int32_t *__restrict__ prob,
                                                 const double *__restrict__ src, const double *__restrict__ cur,
                                                 T *__restrict__ pos) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n * A) return;
  const int s = (int)(i / n);
  const int64_t j = i - (int64_t)s * n;
  const int k = prob[s];
  const double *p = src + 3 * ((int64_t)k * n + j);
  double x, y, z;
  xf_apply(cur + 12 * k, p[0], p[1], p[2], x, y, z);
  pos[3 * i] = (T)x;
  pos[3 * i + 1] = (T)y;
  pos[3 * i + 2] = (T)z;
}

// Sums over the points of each active problem, with q the correspondences by problem (q[k n + j]).
// MODE 0: wt p (3), wt q (3), wt (1). MODE 1: the cross-covariance sum of (wr dp)(wr dq)^T with the
// up component removed (9). MODE 2: wr^2 |L dp - dq|^2 (1). aux (12 per problem): up in [0..2] and
// has_up in [3] (MODE 1), L row-major in [0..8] (MODE 2); sums7: the MODE 0 sums (means).
template <int MODE, typename QT>
__global__ void __launch_bounds__(ICP_THREADS) k_icp_sums(int64_t n, const int32_t *__restrict__ prob,
                                                          const double *__restrict__ src,
                                                          const double *__restrict__ cur,
                                                          const QT *__restrict__ q, const double *__restrict__ wt,
                                                          const double *__restrict__ wr,
                                                          const double *__restrict__ sums7,
                                                          const double *__restrict__ aux, int nb,
                                                          double *__restrict__ partial) {
  constexpr int NC = MODE == 0 ? 7 : (MODE == 1 ? 9 : 1);
  __shared__ double s_red[ICP_THREADS * NC];
  const int s = blockIdx.x / nb;  // slot
  const int b = blockIdx.x % nb;  // block within the problem
  const int k = prob[s];
  const double *xf = cur + 12 * k;
  double pm[3] = {0, 0, 0}, qm[3] = {0, 0, 0};
  if (MODE != 0) {
    const double *sm = sums7 + 7 * k;
    for (int a = 0; a < 3; ++a) {
      pm[a] = sm[a] / sm[6];
      qm[a] = sm[3 + a] / sm[6];
    }
  }
  double acc[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) acc[c] = 0;
  // thread t sums points b CHUNK + t + 256 i (i = 0..15) in order: a wave's loads are contiguous
  const int64_t j0 = (int64_t)b * ICP_CHUNK + threadIdx.x;
  for (int t = 0; t < ICP_PER_THREAD; ++t) {
    const int64_t j = j0 + (int64_t)t * ICP_THREADS;
    if (j >= n) break;
    const double *sp = src + 3 * ((int64_t)k * n + j);
    double p[3];
    xf_apply(xf, sp[0], sp[1], sp[2], p[0], p[1], p[2]);
    const QT *qq = q + 3 * ((int64_t)k * n + j);
    const double qd[3] = {(double)qq[0], (double)qq[1], (double)qq[2]};
    if (MODE == 0) {
      const double w = wt ? wt[j] : 1.0;
      for (int a = 0; a < 3; ++a) {
        acc[a] += wt ? w * p[a] : p[a];
        acc[3 + a] += wt ? w * qd[a] : qd[a];
      }
      acc[6] += w;
    } else if (MODE == 1) {
      double dp[3], dq[3];
      for (int a = 0; a < 3; ++a) {
        dp[a] = p[a] - pm[a];
        dq[a] = qd[a] - qm[a];
      }
      const double *up = aux + 12 * k;
      if (up[3] != 0) {  // remove the component along the up vector (Icp3.hpp:333-337)
        const double ep = (dp[0] * up[0] + dp[1] * up[1]) + dp[2] * up[2];
        const double eq = (dq[0] * up[0] + dq[1] * up[1]) + dq[2] * up[2];
        for (int a = 0; a < 3; ++a) {
          dp[a] = dp[a] - ep * up[a];
          dq[a] = dq[a] - eq * up[a];
        }
      }
      if (wr) {
        for (int a = 0; a < 3; ++a) {
          dp[a] *= wr[j];
          dq[a] *= wr[j];
        }
      }
      for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 3; ++c) acc[3 * a + c] += dp[a] * dq[c];
    } else {
      const double *L = aux + 12 * k;
      double e = 0;
      double r[3];
      for (int a = 0; a < 3; ++a) {
        const double dp0 = p[0] - pm[0], dp1 = p[1] - pm[1], dp2 = p[2] - pm[2];
        r[a] = ((L[3 * a] * dp0 + L[3 * a + 1] * dp1) + L[3 * a + 2] * dp2) - (qd[a] - qm[a]);
      }
      e = (r[0] * r[0] + r[1] * r[1]) + r[2] * r[2];
      acc[0] += wr ? (wr[j] * wr[j]) * e : e;
    }
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) s_red[c * ICP_THREADS + threadIdx.x] = acc[c];
  __syncthreads();
  for (int w = ICP_THREADS / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      for (int c = 0; c < NC; ++c) s_red[c * ICP_THREADS + threadIdx.x] += s_red[c * ICP_THREADS + threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < NC) partial[((int64_t)s * nb + b) * NC + threadIdx.x] = s_red[threadIdx.x * ICP_THREADS];
}

// out[prob[s] * stride + c] = sum over the problem's blocks, in block order.
__global__ void k_icp_final(int A, int nb, int nc, const int32_t *__restrict__ prob, const double *__restrict__ partial,
                            int stride, double *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A * nc) return;
  const int s = i / nc, c = i % nc;
  double acc = 0;
  for (int b = 0; b < nb; ++b) acc += partial[((int64_t)s * nb + b) * nc + c];
  out[prob[s] * stride + c] = acc;
}

//--------------------------------------------------------------------------------------------------
// Host: 3x3 SVD (two-sided Jacobi) and the ICP loop
//--------------------------------------------------------------------------------------------------
struct M3 {
  double a[3][3];
};

static M3 m3_identity() {
  M3 r{};
  for (int i = 0; i < 3; ++i) r.a[i][i] = 1;
  return r;
}

static M3 m3_mul(const M3 &x, const M3 &y) {
  M3 r{};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.a[i][j] = (x.a[i][0] * y.a[0][j] + x.a[i][1] * y.a[1][j]) + x.a[i][2] * y.a[2][j];
  return r;
}

// Plane rotation [c s; -s c] acting on rows/columns p, q.
struct Rot {
  double c = 1, s = 0;
};

static Rot rot_transpose(const Rot &r) { return Rot{r.c, -r.s}; }
static Rot rot_mul(const Rot &x, const Rot &y) {  // x * y
  return Rot{x.c * y.c - x.s * y.s, x.c * y.s + x.s * y.c};
}
// rows p, q of A <- R * rows
static void apply_left(M3 &A, int p, int q, const Rot &r) {
  for (int j = 0; j < 3; ++j) {
    const double x = A.a[p][j], y = A.a[q][j];
    A.a[p][j] = r.c * x + r.s * y;
    A.a[q][j] = -r.s * x + r.c * y;
  }
}
// columns p, q of A <- columns * R
static void apply_right(M3 &A, int p, int q, const Rot &r) {
  for (int i = 0; i < 3; ++i) {
    const double x = A.a[i][p], y = A.a[i][q];
    A.a[i][p] = r.c * x - r.s * y;
    A.a[i][q] = r.s * x + r.c * y;
  }
}

// Rotation G = [c s; -s c] with G^T [x y; y z] G diagonal: t = s / c solves y t^2 - (x - z) t - y = 0;
// the root of smaller magnitude, -2y / ((x - z) + sign(x - z) sqrt((x - z)^2 + 4 y^2)), keeps the
// rotation under 45 degrees (identity if y is 0).
static Rot sym_jacobi(double x, double y, double z) {
  if (y == 0) return Rot{};
  const double dxz = x - z;
  const double r = std::hypot(dxz, 2 * y);
  const double t = -2 * y / (dxz >= 0 ? dxz + r : dxz - r);
  const double c = 1 / std::sqrt(1 + t * t);
  return Rot{c, t * c};
}

// A = U diag(s) V^T by two-sided Jacobi sweeps over the pairs (1,0), (2,0), (2,1) until every
// off-diagonal pair is below 2 eps times the largest diagonal entry; then signs made positive and
// the singular values sorted in decreasing order. Returns U V^T, the orthogonal factor Thea takes
// (Icp3.hpp:349-350). With an exactly zero row and column (points projected off an axis-aligned up
// vector) that axis is never rotated, so U V^T keeps it fixed.
static M3 svd_uvt(const M3 &in) {
  M3 A = in, U = m3_identity(), V = m3_identity();
  const double eps = 2 * DBL_EPSILON;
  for (int sweep = 0; sweep < 100; ++sweep) {
    bool changed = false;
    for (int p = 1; p < 3; ++p)
      for (int q = 0; q < p; ++q) {
        double dmax = std::max(std::fabs(A.a[p][p]), std::fabs(A.a[q][q]));
        const double thr = std::max(DBL_MIN, eps * dmax);
        if (std::max(std::fabs(A.a[p][q]), std::fabs(A.a[q][p])) <= thr) continue;
        changed = true;
        // make the 2x2 block symmetric with a left rotation, then diagonalise it symmetrically
        const double t = A.a[p][p] + A.a[q][q], d = A.a[q][p] - A.a[p][q];
        Rot r1;
        if (std::fabs(d) >= DBL_MIN) {
          const double u = t / d, tmp = std::sqrt(1 + u * u);
          r1.s = 1 / tmp;
          r1.c = u / tmp;
        }
        // block after r1 (rows p, q): [c s; -s c] * [app apq; aqp aqq]
        const double bpp = r1.c * A.a[p][p] + r1.s * A.a[q][p], bpq = r1.c * A.a[p][q] + r1.s * A.a[q][q];
        const double bqq = -r1.s * A.a[p][q] + r1.c * A.a[q][q];
        const Rot jr = sym_jacobi(bpp, bpq, bqq);
        const Rot jl = rot_mul(rot_transpose(jr), r1);
        apply_left(A, p, q, jl);
        apply_right(U, p, q, rot_transpose(jl));
        apply_right(A, p, q, jr);
        apply_right(V, p, q, jr);
      }
    if (!changed) break;
  }
  double sv[3];
  for (int i = 0; i < 3; ++i) {
    sv[i] = A.a[i][i];
    if (sv[i] < 0) {
      sv[i] = -sv[i];
      for (int r = 0; r < 3; ++r) U.a[r][i] = -U.a[r][i];
    }
  }
  // sort decreasing (permuting U and V columns together leaves U V^T unchanged)
  M3 R{};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      R.a[i][j] = (U.a[i][0] * V.a[j][0] + U.a[i][1] * V.a[j][1]) + U.a[i][2] * V.a[j][2];
  return R;
}

static Xf xf_identity() {
  Xf x{};
  x.m[0] = x.m[5] = x.m[10] = 1;
  return x;
}

// a * b (b applied first): L = La Lb, t = La tb + ta (AffineTransformN.hpp:154-157).
static Xf xf_compose(const Xf &a, const Xf &b) {
  Xf r{};
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j)
      r.m[4 * i + j] = (a.m[4 * i] * b.m[j] + a.m[4 * i + 1] * b.m[4 + j]) + a.m[4 * i + 2] * b.m[8 + j];
    r.m[4 * i + 3] = ((a.m[4 * i] * b.m[3] + a.m[4 * i + 1] * b.m[7]) + a.m[4 * i + 2] * b.m[11]) + a.m[4 * i + 3];
  }
  return r;
}

struct IcpProblem {
  Xf old_tr, new_tr, cur;  // cur: the transform giving the current from_points
  double old_err = 0, new_err = 0;
  int it = 0;
  bool done = false;
  Xf result;
  double result_err = 0;
  int32_t iterations = 0;
};

// One ICP run for K problems sharing `to`. from: n double3 (device); init: K x 12 (device, nullable);
// weights: n doubles each (device, nullable). Writes K x 12 transforms, K errors, K iteration counts.
void icp_align(const Bvh &to, const double *from, int64_t n, int K, const double *init, double frac_thr, int min_it,
               int max_it, bool has_up, const double up_in[3], bool kabsch, const double *wt, const double *wr,
               double *out_tr, double *out_err, int32_t *out_iters, hipStream_t s) {
  require(n > 0 && K > 0, "empty ICP problem");
  require(to.n > 0, "the target of an ICP alignment has no elements (Icp3.hpp:249-250)");
  require((double)n * K < 2.0e9, "ICP batch too large (n * K must stay below 2^31 queries)");
  double up[3] = {0, 0, 0};
  if (has_up) {  // Icp3::setUpVector normalises (Icp3.hpp:68)
    const double nn = std::sqrt((up_in[0] * up_in[0] + up_in[1] * up_in[1]) + up_in[2] * up_in[2]);
    require(nn > 0, "zero up vector");
    for (int a = 0; a < 3; ++a) up[a] = up_in[a] / nn;
  }
  const int nb = (int)((n + ICP_CHUNK - 1) / ICP_CHUNK);
  AsyncBuf src(n * K * 24, s), cur(K * 96, s), aux(K * 96, s), sums7(K * 56, s), cov(K * 72, s), errs(K * 8, s);
  // qpts: the correspondences of problem k at [k n, (k + 1) n) (kept across iterations whatever the
  // active set); qtmp: the nearest points of one launch, by slot
  // A double point BVH (BvhN<Vector3d>, ShapeAlign.cpp:339) is searched in double with the positions
  // kept in double; a float BVH with the positions rounded to float.
  const bool f64 = to.f64;
  const int64_t pb = f64 ? 24 : 12;  // bytes per position / correspondence
  AsyncBuf pos(n * K * pb, s), qpts(n * K * pb, s), qtmp(n * K * pb, s), idx(n * K * 4, s), prob(K * 4, s);
  AsyncBuf partial((size_t)K * nb * ICP_MAXC * 8, s);
  hipLaunchKernelGGL(k_icp_init, dim3(ceil_div(n * K, 256)), dim3(256), 0, s, n, K, from, init, src.as<double>());
  launch_check("k_icp_init");

  std::vector<IcpProblem> P(K);
  for (auto &p : P) { p.old_tr = p.new_tr = p.cur = xf_identity(); }
  std::vector<double> h_cur(12 * K), h_aux(12 * K), h_sums(7 * K), h_cov(9 * K), h_err(K);
  std::vector<int32_t> active;

  auto upload_cur = [&]() {
    for (int k = 0; k < K; ++k) std::copy(P[k].cur.m, P[k].cur.m + 12, &h_cur[12 * k]);
    THG_HIP(hipMemcpyAsync(cur.p, h_cur.data(), 96 * K, hipMemcpyHostToDevice, s));
  };
  auto set_active = [&]() {
    active.clear();
    for (int k = 0; k < K; ++k)
      if (!P[k].done) active.push_back(k);
    if (!active.empty())
      THG_HIP(hipMemcpyAsync(prob.p, active.data(), 4 * active.size(), hipMemcpyHostToDevice, s));
  };
  // nearest points of the active problems' current positions
  auto nearest = [&]() {
    const int A = (int)active.size();
    QueryScratch ws;
    if (f64) {
      hipLaunchKernelGGL(k_icp_pos<double>, dim3(ceil_div(n * A, 256)), dim3(256), 0, s, n, A, prob.as<int32_t>(),
                         src.as<double>(), cur.as<double>(), pos.as<double>());
      launch_check("k_icp_pos");
      launch_closest64(to, pos.as<double>(), n * A, -1.0, idx.as<int32_t>(), qtmp.as<double>(), nullptr, nullptr, ws, s);
    } else {
      hipLaunchKernelGGL(k_icp_pos<float>, dim3(ceil_div(n * A, 256)), dim3(256), 0, s, n, A, prob.as<int32_t>(),
                         src.as<double>(), cur.as<double>(), pos.as<float>());
      launch_check("k_icp_pos");
      launch_closest(to, pos.as<float>(), n * A, -1.0, 0, idx.as<int32_t>(), qtmp.as<float>(), nullptr, nullptr, ws, s);
    }
    for (int sl = 0; sl < A; ++sl)
      THG_HIP(hipMemcpyAsync((char *)qpts.p + pb * (int64_t)active[sl] * n, (char *)qtmp.p + pb * (int64_t)sl * n, n * pb,
                             hipMemcpyDeviceToDevice, s));
  };
  auto sums = [&](int mode, double *dst, int stride) {
    const int A = (int)active.size();
    const int nc = mode == 0 ? 7 : (mode == 1 ? 9 : 1);
    const dim3 g((unsigned)(A * nb));
#define THG_SUMS(M, QT)                                                                                            \
  hipLaunchKernelGGL((k_icp_sums<M, QT>), g, dim3(ICP_THREADS), 0, s, n, prob.as<int32_t>(), src.as<double>(),      \
                     cur.as<double>(), qpts.as<QT>(), wt, wr, sums7.as<double>(), aux.as<double>(), nb,              \
                     partial.as<double>())
    if (f64) {
      if (mode == 0) THG_SUMS(0, double); else if (mode == 1) THG_SUMS(1, double); else THG_SUMS(2, double);
    } else {
      if (mode == 0) THG_SUMS(0, float); else if (mode == 1) THG_SUMS(1, float); else THG_SUMS(2, float);
    }
#undef THG_SUMS
    launch_check("k_icp_sums");
    hipLaunchKernelGGL(k_icp_final, dim3(ceil_div((int64_t)A * nc, 64)), dim3(64), 0, s, A, nb, nc, prob.as<int32_t>(),
                       partial.as<double>(), stride, dst);
    launch_check("k_icp_final");
  };
  // measureError(tr, ...) with the current to_points (Icp3.hpp:422-465): error of L = linear part
  // of trs[k], translation term |tr p_mean - q_mean|^2 added on the host.
  auto measure = [&](const std::vector<Xf> &trs) {
    sums(0, sums7.as<double>(), 7);
    for (int k : active) {
      std::fill(&h_aux[12 * k], &h_aux[12 * k] + 12, 0.0);
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) h_aux[12 * k + 3 * i + j] = trs[k].m[4 * i + j];
    }
    THG_HIP(hipMemcpyAsync(aux.p, h_aux.data(), 96 * K, hipMemcpyHostToDevice, s));
    sums(2, errs.as<double>(), 1);
    THG_HIP(hipMemcpyAsync(h_sums.data(), sums7.p, 56 * K, hipMemcpyDeviceToHost, s));
    THG_HIP(hipMemcpyAsync(h_err.data(), errs.p, 8 * K, hipMemcpyDeviceToHost, s));
    THG_HIP(hipStreamSynchronize(s));
    std::vector<double> e(K, 0);
    for (int k : active) {
      const double *sm = &h_sums[7 * k];
      const double pm[3] = {sm[0] / sm[6], sm[1] / sm[6], sm[2] / sm[6]};
      const double qm[3] = {sm[3] / sm[6], sm[4] / sm[6], sm[5] / sm[6]};
      double tp[3];
      xf_apply(trs[k].m, pm[0], pm[1], pm[2], tp[0], tp[1], tp[2]);
      const double d0 = tp[0] - qm[0], d1 = tp[1] - qm[1], d2 = tp[2] - qm[2];
      e[k] = ((d0 * d0 + d1 * d1) + d2 * d2) + h_err[k];
    }
    return e;
  };

  // initial mapping and error (Icp3.hpp:210-219)
  set_active();
  upload_cur();
  nearest();
  {
    std::vector<Xf> ident(K, xf_identity());
    std::vector<double> e = measure(ident);
    for (int k = 0; k < K; ++k) {
      P[k].old_err = P[k].new_err = e[k];
      if (e[k] <= DBL_MIN) {
        P[k].done = true;
        P[k].result = P[k].old_tr;
        P[k].result_err = e[k];
      }
    }
  }
  for (int it = 0; it < max_it; ++it) {
    set_active();
    if (active.empty()) break;
    // alignOneStep with the mapping of the last measurement (Icp3.hpp:288-375)
    sums(0, sums7.as<double>(), 7);
    for (int k : active) {
      std::fill(&h_aux[12 * k], &h_aux[12 * k] + 12, 0.0);
      for (int a = 0; a < 3; ++a) h_aux[12 * k + a] = up[a];
      h_aux[12 * k + 3] = has_up ? 1.0 : 0.0;
    }
    THG_HIP(hipMemcpyAsync(aux.p, h_aux.data(), 96 * K, hipMemcpyHostToDevice, s));
    sums(1, cov.as<double>(), 9);
    THG_HIP(hipMemcpyAsync(h_cov.data(), cov.p, 72 * K, hipMemcpyDeviceToHost, s));
    THG_HIP(hipMemcpyAsync(h_sums.data(), sums7.p, 56 * K, hipMemcpyDeviceToHost, s));
    THG_HIP(hipStreamSynchronize(s));
    std::vector<Xf> inc(K, xf_identity());
    for (int k : active) {
      M3 C{};
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) C.a[i][j] = h_cov[9 * k + 3 * i + j];
      M3 R = svd_uvt(C);
      if (kabsch)  // V U^T, the minimiser of sum |R dp - dq|^2
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < i; ++j) std::swap(R.a[i][j], R.a[j][i]);
      const double *sm = &h_sums[7 * k];
      const double pm[3] = {sm[0] / sm[6], sm[1] / sm[6], sm[2] / sm[6]};
      const double qm[3] = {sm[3] / sm[6], sm[4] / sm[6], sm[5] / sm[6]};
      Xf x{};
      for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) x.m[4 * i + j] = R.a[i][j];
        // trans = q_mean - rot * p_mean (Icp3.hpp:363)
        x.m[4 * i + 3] = qm[i] - ((R.a[i][0] * pm[0] + R.a[i][1] * pm[1]) + R.a[i][2] * pm[2]);
      }
      inc[k] = x;
      P[k].new_tr = xf_compose(x, P[k].old_tr);
      P[k].iterations = it + 1;
    }
    // measureError(inc, from_points, to): new mapping at the same positions, error of inc on it
    if (it > 0) nearest();  // at it == 0 the positions are the initial ones: same mapping
    std::vector<double> e = measure(inc);
    bool moved = false;
    for (int k : active) {
      IcpProblem &p = P[k];
      p.new_err = e[k];
      if (it < max_it - 1) {
        const double frac = (p.old_err - p.new_err) / p.old_err;
        if (it >= min_it && frac < frac_thr) {
          p.done = true;
          if (frac > 0) { p.result = p.new_tr; p.result_err = p.new_err; }
          else { p.result = p.old_tr; p.result_err = p.old_err; }
          continue;
        }
        p.cur = p.new_tr;  // from_points = new_tr * from (Icp3.hpp:251-252)
        moved = true;
      }
      p.old_err = p.new_err;
      p.old_tr = p.new_tr;
    }
    if (moved) upload_cur();
  }
  for (int k = 0; k < K; ++k) {
    IcpProblem &p = P[k];
    if (!p.done) { p.result = p.new_tr; p.result_err = p.new_err; }
    std::copy(p.result.m, p.result.m + 12, out_tr + 12 * k);
    if (out_err) out_err[k] = p.result_err;
    if (out_iters) out_iters[k] = p.iterations;
  }
}

}  // namespace thg

using namespace thg;

template <typename F> static int icp_guarded(F f) {
  try {
    f();
    return THG_OK;
  } catch (const Status &e) {
    set_last_error(e.msg);
    return e.code;
  } catch (const std::bad_alloc &) {
    set_last_error("host allocation failed");
    return THG_E_OOM;
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return THG_E_DEVICE;
  }
}

extern "C" int thg_icp_align(thg_bvh to, const double *from, int64_t n, int32_t num_problems, const double *initial,
                             const thg_icp_opts *opts, int32_t mem, void *stream, double *transforms, double *errors,
                             int32_t *iterations) {
  return icp_guarded([&] {
    require(to != nullptr && transforms != nullptr, "to/transforms is NULL");
    require(n >= 0 && num_problems >= 1, "bad point or problem count");
    require(n == 0 || from != nullptr, "from is NULL");
    require(mem == THG_MEM_DEVICE || mem == THG_MEM_HOST, "mem must be THG_MEM_DEVICE or THG_MEM_HOST");
    Bvh *b = reinterpret_cast<Bvh *>(to);
    THG_HIP(hipSetDevice(b->ctx->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    // Icp3's constructor defaults (Icp3.hpp:59-66)
    double thr = 0.0001;
    int min_it = 3, max_it = 10;
    bool has_up = false, kabsch = false;
    double up[3] = {0, 0, 0};
    const double *wt = nullptr, *wr = nullptr;
    if (opts) {
      if (opts->fractional_error_threshold >= 0) thr = opts->fractional_error_threshold;
      if (opts->min_iterations >= 0) min_it = opts->min_iterations;
      if (opts->max_iterations >= 0) max_it = opts->max_iterations;
      has_up = opts->has_up != 0;
      kabsch = opts->kabsch != 0;
      for (int a = 0; a < 3; ++a) up[a] = opts->up[a];
      wt = opts->translation_weights;
      wr = opts->rotation_weights;
    }
    const int K = num_problems;
    if (n == 0) {  // Icp3.hpp:201-205: identity, error 0
      for (int k = 0; k < K; ++k) {
        for (int i = 0; i < 12; ++i) transforms[12 * k + i] = (i % 5 == 0) ? 1.0 : 0.0;
        if (errors) errors[k] = 0;
        if (iterations) iterations[k] = 0;
      }
      return;
    }
    AsyncBuf hf, hi, hwt, hwr;
    auto stage = [&](AsyncBuf &buf, const double *p, size_t count) -> const double * {
      if (!p || mem == THG_MEM_DEVICE) return p;
      buf.ensure(count * 8, s);
      THG_HIP(hipMemcpyAsync(buf.p, p, count * 8, hipMemcpyHostToDevice, s));
      return buf.as<double>();
    };
    const double *d_from = stage(hf, from, (size_t)n * 3);
    const double *d_init = stage(hi, initial, (size_t)K * 12);
    const double *d_wt = stage(hwt, wt, (size_t)n);
    const double *d_wr = stage(hwr, wr, (size_t)n);
    icp_align(*b, d_from, n, K, d_init, thr, min_it, max_it, has_up, up, kabsch, d_wt, d_wr, transforms, errors,
              iterations, s);
    THG_HIP(hipStreamSynchronize(s));
  });
}
