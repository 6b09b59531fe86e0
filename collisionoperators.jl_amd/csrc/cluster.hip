// cluster.hip -- stiff clusters of the implicit step (extension; the reference
// has no counterpart).
//
// Markers closer than max(w_i, w_j) / theta form clusters.  Inside a cluster
// the pair terms w U(d) (G_a - G_g) make the midpoint system stiff (a pair's
// stiffness grows like dt w |Hess L| / |d|); once there are more such modes
// than Anderson window columns the plain iteration stalls.  Every FP64 map
// therefore replaces each cluster's rows by the exact solution of the
// cluster's own midpoint equations, everything outside the cluster held at the
// map's values (nonlinear elimination):
//   m_i = v0_i + dt/2 (ext_i + sum_{j in C} w_j U(m_i - m_j) (g_j - g_i)),
//   g_i = grad L(m_i) - 2 lambda (m_i - v0_i) / w_i,     x_i = 2 m_i - v0_i,
// ext_i = the map's vdot_i minus the cluster's own terms at the iterate, and L
// (= M^-1 r), lambda those of the map.  A fixed point of the modified map is
// one of the plain map, so the step's root is unchanged.
//
// Kernels:
//   k_cell_count / k_cell_scan / k_cell_fill: markers bucketed by mesh cell;
//   k_near_cells: one workgroup per cell tests the cell's pairs and, for
//     markers within reach of a cell edge, the pairs across it (FP64);
//   k_cluster_solve: one wavefront per cluster (<= 32 markers, 64 unknowns):
//     a pair is one unknown angle (grid of 64 angles, one per lane, then
//     bisection); larger clusters take the homotopy in the coupling strength t
//     from t = 0 (m = v0 + dt/2 ext) to t = 1, Newton with the analytic
//     Jacobian and partial-pivot LU in LDS at every point; a warm start from
//     the cluster's previous root first, and Newton from the eigen-line
//     configurations of Hess L if the path folds.
#include "collop_launch.hpp"

namespace collop {
namespace {

constexpr int kClN = 2 * kClusterMax;  // unknowns per cluster

// ---- cell buckets ---------------------------------------------------------------

// x[0 .. n) <- value (a kernel, not a memset, so that the step stays capturable)
__global__ void k_fill_int(int* __restrict__ x, int n, int value) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = value;
}

__global__ void k_cell_count(MeshView m, const double* __restrict__ v, int N, int* __restrict__ cell_out,
                             int* __restrict__ cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double x = v[i], y = v[N + i];
  int c = -1;
  if (in_box(x, y, m.lo1, m.hi1, m.lo2, m.hi2)) {
    const int a = cell_of(m.bp1, m.nb1, m.uniform1, m.lo1, m.inv_h1, x);
    const int b = cell_of(m.bp2, m.nb2, m.uniform2, m.lo2, m.inv_h2, y);
    c = a + m.ne1 * b;
    atomicAdd(&cnt[c], 1);
  }
  cell_out[i] = c;
}

// Exclusive scan of ncell counts in one workgroup: start[c], start[ncell] = total;
// cnt is reset to 0 for the fill pass.
__global__ __launch_bounds__(1024) void k_cell_scan(int* __restrict__ cnt, int ncell, int* __restrict__ start) {
  __shared__ int part[1024];
  const int t = threadIdx.x;
  const int per = (ncell + 1023) / 1024;
  const int b = t * per, e = min(ncell, b + per);
  int s = 0;
  for (int c = b; c < e; ++c) s += cnt[c];
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const int x = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += x;
    __syncthreads();
  }
  int run = t ? part[t - 1] : 0;
  for (int c = b; c < e; ++c) {
    start[c] = run;
    run += cnt[c];
    cnt[c] = 0;
  }
  if (t == 1023) start[ncell] = part[1023];
}

__global__ void k_cell_fill(const int* __restrict__ cell_in, int N, const int* __restrict__ start,
                            int* __restrict__ fill, int* __restrict__ sorted) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int c = cell_in[i];
  if (c >= 0) sorted[start[c] + atomicAdd(&fill[c], 1)] = i;
}

__device__ __forceinline__ void push_link(double q, int a, int b, NearLink* out, int* count, int cap) {
  const int k = atomicAdd(count, 1);
  if (k < cap) out[k] = NearLink{a < b ? a : b, a < b ? b : a, q};
}

__device__ __forceinline__ bool near_pair(const double* __restrict__ v, const double* __restrict__ w, int N, int a,
                                          int b, double th2, double& q) {
  const double d1 = v[a] - v[b], d2 = v[N + a] - v[N + b];
  q = d1 * d1 + d2 * d2;
  const double wm = fmax(w[a], w[b]);
  return q >= 1e-24 && th2 * q < wm * wm;
}

// One workgroup per cell: the cell's own pairs, then its pairs with the right,
// upper, upper-right and upper-left neighbours for markers within r = max w /
// theta of the shared edge or corner (each unordered pair is tested once).
__global__ __launch_bounds__(256) void k_near_cells(MeshView m, const double* __restrict__ v,
                                                    const double* __restrict__ w, int N, double th2, double r,
                                                    const int* __restrict__ start, const int* __restrict__ sorted,
                                                    NearLink* __restrict__ out, int* __restrict__ count, int cap) {
  const int c = blockIdx.x;
  const int ci = c % m.ne1, cj = c / m.ne1;
  const int b0 = start[c], n = start[c + 1] - b0;
  if (n == 0) return;
  // own pairs: flattened upper triangle (p, q), p < q
  const long long npair = (long long)n * (n - 1) / 2;
  for (long long t = threadIdx.x; t < npair; t += blockDim.x) {
    // row p of the triangle: t = p n - p (p + 1) / 2 + (q - p - 1)
    int p = (int)((2.0 * n - 1.0 - sqrt((2.0 * n - 1.0) * (2.0 * n - 1.0) - 8.0 * (double)t)) * 0.5);
    while (p > 0 && (long long)p * n - (long long)p * (p + 1) / 2 > t) --p;
    while ((long long)(p + 1) * n - (long long)(p + 1) * (p + 2) / 2 <= t) ++p;
    const int q = (int)(t - ((long long)p * n - (long long)p * (p + 1) / 2)) + p + 1;
    const int a = sorted[b0 + p], b = sorted[b0 + q];
    double qd;
    if (near_pair(v, w, N, a, b, th2, qd)) push_link(qd, a, b, out, count, cap);
  }
  // neighbours across edges: (di, dj) in {(1,0), (0,1), (1,1), (-1,1)}
  const double x0 = m.bp1[ci], x1 = m.bp1[ci + 1], y1 = m.bp2[cj + 1];
  for (int s = threadIdx.x; s < n; s += blockDim.x) {
    const int a = sorted[b0 + s];
    const double x = v[a], y = v[N + a];
    const bool nr = x1 - x < r, nu = y1 - y < r, nl = x - x0 < r;
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) {
      const int di = nb == 0 ? 1 : (nb == 1 ? 0 : (nb == 2 ? 1 : -1));
      const int dj = nb == 0 ? 0 : 1;
      const bool reach = nb == 0 ? nr : (nb == 1 ? nu : (nb == 2 ? (nr && nu) : (nl && nu)));
      const int ni = ci + di, nj = cj + dj;
      if (!reach || ni < 0 || ni >= m.ne1 || nj >= m.ne2) continue;
      const int cn = ni + m.ne1 * nj;
      for (int u = start[cn]; u < start[cn + 1]; ++u) {
        const int b = sorted[u];
        double qd;
        if (near_pair(v, w, N, a, b, th2, qd)) push_link(qd, a, b, out, count, cap);
      }
    }
  }
}

// cl_row[idx[r]] = r for the cluster members (or -1: reset of the previous members)
__global__ void k_cluster_rows(const int* __restrict__ idx, int total, int set, int* __restrict__ row) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < total) row[idx[r]] = set ? r : -1;
}

// ---- the cluster solve: one wavefront per cluster --------------------------------

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
  return x;
}
__device__ __forceinline__ double wave_min(double x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x = fmin(x, __shfl_xor(x, off, 64));
  return x;
}

struct ClusterLds {
  double sb1[kMaxBp], sb2[kMaxBp];
  double v0[kClN], ext[kClN], m[kClN], mt0[kClN], mp[kClN], mn[kClN], F[kClN], dl[kClN], g[kClN];
  double H[3 * kClusterMax], w[kClusterMax];
  double J[kClN][kClN + 1];
};

// grad L and Hess L (H11, H12, H22) at (x, y) with L the coefficient array; false outside the box
__device__ bool grad_hess_dev(const MeshView& M, const double* sb1, const double* sb2, const double* __restrict__ L,
                              double x, double y, double* g, double* H) {
  Located lc;
  if (!locate(M, sb1, sb2, x, y, lc)) return false;
  double f1[3], df1[3], f2[3], df2[3], dd1[3], dd2[3];
  basis3_grad(M.C1, lc.i, lc.xi1, f1, df1);
  basis3_grad(M.C2, lc.j, lc.xi2, f2, df2);
  // second derivatives of the quadratic Bernstein polynomials: (2, -4, 2)
  const double* c1 = M.C1 + 9 * lc.i;
  const double* c2 = M.C2 + 9 * lc.j;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    dd1[j] = 2.0 * c1[3 * j] - 4.0 * c1[3 * j + 1] + 2.0 * c1[3 * j + 2];
    dd2[j] = 2.0 * c2[3 * j] - 4.0 * c2[3 * j + 1] + 2.0 * c2[3 * j + 2];
  }
  const int s1 = M.bs1[lc.i], s2 = M.bs2[lc.j];
  double gx = 0.0, gy = 0.0, hxx = 0.0, hxy = 0.0, hyy = 0.0;
#pragma unroll
  for (int j2 = 0; j2 < 3; ++j2)
#pragma unroll
    for (int j1 = 0; j1 < 3; ++j1) {
      const double c = L[(s1 + j1) + (s2 + j2) * M.nd1];
      gx += c * df1[j1] * f2[j2];
      gy += c * f1[j1] * df2[j2];
      hxx += c * dd1[j1] * f2[j2];
      hxy += c * df1[j1] * df2[j2];
      hyy += c * f1[j1] * dd2[j2];
    }
  const double i1 = 1.0 / lc.h1, i2 = 1.0 / lc.h2;
  g[0] = gx * i1;
  g[1] = gy * i2;
  H[0] = hxx * i1 * i1;
  H[1] = hxy * i1 * i2;
  H[2] = hyy * i2 * i2;
  return true;
}

// Phi(x; t) into S.F (and the Jacobian into S.J): wave-uniform result, false if
// a marker left the box or two coincide.
__device__ bool cl_eval(ClusterLds& S, const ClusterArgs& A, int k, double lam, const double* x, double t,
                        bool jac) {
  const int lane = threadIdx.x, n = 2 * k;
  bool ok = true;
  if (lane < k) {
    double g[2], H[3];
    ok = grad_hess_dev(A.m, S.sb1, S.sb2, A.L, x[2 * lane], x[2 * lane + 1], g, H);
    const double iw = 2.0 * lam / S.w[lane];
    S.g[2 * lane] = g[0] - iw * (x[2 * lane] - S.v0[2 * lane]);
    S.g[2 * lane + 1] = g[1] - iw * (x[2 * lane + 1] - S.v0[2 * lane + 1]);
    S.H[3 * lane] = H[0];
    S.H[3 * lane + 1] = H[1];
    S.H[3 * lane + 2] = H[2];
  }
  if (__any(!ok)) return false;
  __syncthreads();
  const double s = t * 0.5 * A.dt;
  if (lane < k) {
    const int i = lane;
    double acc0 = 0.0, acc1 = 0.0;
    double Jii[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
    const double Bi[2][2] = {{S.H[3 * i] - 2.0 * lam / S.w[i], S.H[3 * i + 1]},
                             {S.H[3 * i + 1], S.H[3 * i + 2] - 2.0 * lam / S.w[i]}};
    for (int j = 0; j < k; ++j) {
      if (j == i) {
        if (jac) S.J[2 * i][2 * j] = S.J[2 * i][2 * j + 1] = S.J[2 * i + 1][2 * j] = S.J[2 * i + 1][2 * j + 1] = 0.0;
        continue;
      }
      const double d0 = x[2 * i] - x[2 * j], d1 = x[2 * i + 1] - x[2 * j + 1];
      const double y0 = S.g[2 * j] - S.g[2 * i], y1 = S.g[2 * j + 1] - S.g[2 * i + 1];
      const double qd = d0 * d0 + d1 * d1;
      if (!(qd >= 1e-24)) ok = false;
      const double r = sqrt(qd), r3 = qd * r;
      const double dy = d0 * y0 + d1 * y1;
      const double wj = S.w[j];
      acc0 += wj * (y0 / r - d0 * dy / r3);
      acc1 += wj * (y1 / r - d1 * dy / r3);
      if (jac) {
        const double dd[2] = {d0, d1}, yy[2] = {y0, y1};
        const double Bj[2][2] = {{S.H[3 * j] - 2.0 * lam / wj, S.H[3 * j + 1]},
                                 {S.H[3 * j + 1], S.H[3 * j + 2] - 2.0 * lam / wj}};
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int b = 0; b < 2; ++b) {
            const double Dd = -(yy[a] * dd[b] + dd[a] * yy[b] + (a == b ? dy : 0.0)) / r3 +
                              3.0 * dd[a] * dy * dd[b] / (r3 * qd);
            const double U0 = ((a == 0 ? 1.0 : 0.0) - dd[a] * dd[0] / qd) / r;
            const double U1 = ((a == 1 ? 1.0 : 0.0) - dd[a] * dd[1] / qd) / r;
            const double UBi = U0 * Bi[0][b] + U1 * Bi[1][b];
            const double UBj = U0 * Bj[0][b] + U1 * Bj[1][b];
            Jii[a][b] -= s * wj * (Dd - UBi);
            S.J[2 * i + a][2 * j + b] = -s * wj * (UBj - Dd);
          }
      }
    }
    S.F[2 * i] = x[2 * i] - S.v0[2 * i] - 0.5 * A.dt * S.ext[2 * i] - s * acc0;
    S.F[2 * i + 1] = x[2 * i + 1] - S.v0[2 * i + 1] - 0.5 * A.dt * S.ext[2 * i + 1] - s * acc1;
    if (jac) {
      S.J[2 * i][2 * i] = 1.0 + Jii[0][0];
      S.J[2 * i][2 * i + 1] = Jii[0][1];
      S.J[2 * i + 1][2 * i] = Jii[1][0];
      S.J[2 * i + 1][2 * i + 1] = 1.0 + Jii[1][1];
    }
  }
  (void)n;
  const bool all = !__any(!ok);
  __syncthreads();
  return all;
}

__device__ double cl_norm(const ClusterLds& S, int n) {
  const int lane = threadIdx.x;
  const double f = lane < n ? S.F[lane] : 0.0;
  return sqrt(wave_sum(f * f));
}

// S.dl <- solve(S.J, -S.F): partial-pivot LU, rows on lanes.  False if singular.
__device__ bool cl_lu(ClusterLds& S, int n) {
  const int lane = threadIdx.x;
  if (lane < n) S.dl[lane] = -S.F[lane];
  __syncthreads();
  for (int c = 0; c < n; ++c) {
    double a = (lane >= c && lane < n) ? fabs(S.J[lane][c]) : -1.0;
    int p = lane;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const double ao = __shfl_xor(a, off, 64);
      const int po = __shfl_xor(p, off, 64);
      if (ao > a || (ao == a && po < p)) {
        a = ao;
        p = po;
      }
    }
    if (!(a > 0.0)) return false;
    if (p != c) {
      for (int j = lane; j < n; j += 64) {
        const double t = S.J[c][j];
        S.J[c][j] = S.J[p][j];
        S.J[p][j] = t;
      }
      if (lane == 0) {
        const double t = S.dl[c];
        S.dl[c] = S.dl[p];
        S.dl[p] = t;
      }
    }
    __syncthreads();
    if (lane > c && lane < n) {
      const double f = S.J[lane][c] / S.J[c][c];
      if (f != 0.0) {
        for (int j = c + 1; j < n; ++j) S.J[lane][j] -= f * S.J[c][j];
        S.dl[lane] -= f * S.dl[c];
      }
    }
    __syncthreads();
  }
  for (int r = n - 1; r >= 0; --r) {
    const double xr = S.dl[r] / S.J[r][r];
    __syncthreads();
    if (lane < r) S.dl[lane] -= S.J[lane][r] * xr;
    if (lane == r) S.dl[r] = xr;
    __syncthreads();
  }
  return true;
}

// Newton on Phi(x; t) from x (in/out, in LDS).  A step is shortened so that no
// two markers move by more than half their distance relative to each other
// (factor a); it fails on a singular Jacobian or a residual not below
// (1 - a / 10) of the last one (unless already within 1e3 tol).
__device__ bool cl_newton(ClusterLds& S, const ClusterArgs& A, int k, double lam, double* x, double t, int maxit,
                          double tol, int& its) {
  const int lane = threadIdx.x, n = 2 * k;
  if (!cl_eval(S, A, k, lam, x, t, false)) return false;
  double nf = cl_norm(S, n);
  for (int it = 0; it < maxit; ++it) {
    if (nf < tol) return true;
    if (!cl_eval(S, A, k, lam, x, t, true)) return false;
    ++its;
    if (!cl_lu(S, n)) return false;
    double a = 1.0;
    if (lane < k) {
      for (int j = lane + 1; j < k; ++j) {
        const double r0 = S.dl[2 * lane] - S.dl[2 * j], r1 = S.dl[2 * lane + 1] - S.dl[2 * j + 1];
        const double d0 = x[2 * lane] - x[2 * j], d1 = x[2 * lane + 1] - x[2 * j + 1];
        const double rr = r0 * r0 + r1 * r1, dd = d0 * d0 + d1 * d1;
        if (rr > 0.25 * dd) a = fmin(a, 0.5 * sqrt(dd / rr));
      }
    }
    a = wave_min(a);
    if (lane < n) S.mn[lane] = x[lane] + a * S.dl[lane];
    __syncthreads();
    if (!cl_eval(S, A, k, lam, S.mn, t, false)) return false;
    const double nn = cl_norm(S, n);
    if (!(nn <= (1.0 - 0.1 * a) * nf) && nf > 1e3 * tol) return false;
    if (lane < n) x[lane] = S.mn[lane];
    __syncthreads();
    nf = nn;
  }
  return nf < tol;
}

// ---- two-marker clusters: one unknown ---------------------------------------------
// With W = w_g + w_a the weighted centre c of (m_g, m_a) is explicit and
// d = m_g - m_a solves d = a + dt/2 W U(d) (g_a - g_g), a = v0_g - v0_a +
// dt/2 (ext_g - ext_a).  U(d) y is orthogonal to d, so d . (d - a) = 0: d lies
// on the circle over a, d = |a| cos(psi) e(phi_a + psi), |psi| < pi/2, and
//   F(psi) = -|a|^2 / 2 sin(2 psi) + dt/2 W e_perp . (g_a - g_g) = 0.
// The root taken is the one nearest psi_ref: sign changes of F on a fixed grid
// of 64 angles (one per lane), then bisection.
struct PairGeo {
  double c0, c1, a0, a1, na, pa, W;
};

__device__ PairGeo pair_geo(const ClusterLds& S, double dt) {
  PairGeo G;
  const double wg = S.w[0], wa = S.w[1];
  G.W = wg + wa;
  G.c0 = (wg * S.v0[0] + wa * S.v0[2]) / G.W + 0.5 * dt * (wg * S.ext[0] + wa * S.ext[2]) / G.W;
  G.c1 = (wg * S.v0[1] + wa * S.v0[3]) / G.W + 0.5 * dt * (wg * S.ext[1] + wa * S.ext[3]) / G.W;
  G.a0 = S.v0[0] - S.v0[2] + 0.5 * dt * (S.ext[0] - S.ext[2]);
  G.a1 = S.v0[1] - S.v0[3] + 0.5 * dt * (S.ext[1] - S.ext[3]);
  G.na = sqrt(G.a0 * G.a0 + G.a1 * G.a1);
  G.pa = atan2(G.a1, G.a0);
  return G;
}

__device__ __forceinline__ void pair_mid(const PairGeo& G, const ClusterLds& S, double psi, double* m) {
  const double ph = G.pa + psi, rho = G.na * cos(psi);
  const double d0 = rho * cos(ph), d1 = rho * sin(ph);
  m[0] = G.c0 + (S.w[1] / G.W) * d0;
  m[1] = G.c1 + (S.w[1] / G.W) * d1;
  m[2] = G.c0 - (S.w[0] / G.W) * d0;
  m[3] = G.c1 - (S.w[0] / G.W) * d1;
}

// F(psi) for this lane's psi; false outside the box
__device__ bool pair_F(const PairGeo& G, const ClusterLds& S, const ClusterArgs& A, double lam, double psi,
                       double& f) {
  double m[4], g[4], H[3];
  pair_mid(G, S, psi, m);
  for (int i = 0; i < 2; ++i) {
    if (!grad_hess_dev(A.m, S.sb1, S.sb2, A.L, m[2 * i], m[2 * i + 1], &g[2 * i], H)) return false;
    g[2 * i] -= 2.0 * lam * (m[2 * i] - S.v0[2 * i]) / S.w[i];
    g[2 * i + 1] -= 2.0 * lam * (m[2 * i + 1] - S.v0[2 * i + 1]) / S.w[i];
  }
  const double ph = G.pa + psi;
  f = -0.5 * G.na * G.na * sin(2.0 * psi) + 0.5 * A.dt * G.W * (-sin(ph) * (g[2] - g[0]) + cos(ph) * (g[3] - g[1]));
  return true;
}

// The pair's root into S.m (m_g, m_a interleaved); wave-uniform result
__device__ bool pair_solve(ClusterLds& S, const ClusterArgs& A, double lam, double psi_ref) {
  const int lane = threadIdx.x;
  const PairGeo G = pair_geo(S, A.dt);
  if (!(G.na > 0.0)) return false;
  const double pi = 3.141592653589793;
  const double psi_j = -0.5 * pi + pi * (lane + 0.5) / 64.0;
  double fj = 0.0;
  const bool okj = pair_F(G, S, A, lam, psi_j, fj);
  if (__any(!okj)) return false;
  const double fn = __shfl_down(fj, 1, 64);
  // candidate bracket [psi_j, psi_j+1]: a sign change (or a zero at psi_j)
  double dd = 1e300;
  if (lane < 63 && !((fj < 0.0) == (fn < 0.0) && fj != 0.0)) {
    const double md = 0.5 * (psi_j + (-0.5 * pi + pi * (lane + 1.5) / 64.0));
    dd = fabs(md - psi_ref);
  }
  int best = lane;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double od = __shfl_xor(dd, off, 64);
    const int ob = __shfl_xor(best, off, 64);
    if (od < dd || (od == dd && ob < best)) {
      dd = od;
      best = ob;
    }
  }
  if (!(dd < 1e300)) return false;
  double lo = -0.5 * pi + pi * (best + 0.5) / 64.0, hi = -0.5 * pi + pi * (best + 1.5) / 64.0;
  double flo = __shfl(fj, best, 64);
  for (int it = 0; it < 200 && hi - lo > 1e-15; ++it) {
    const double md = 0.5 * (lo + hi);
    double fm = 0.0;
    const bool okm = pair_F(G, S, A, lam, md, fm);
    if (!okm) return false;  // every lane evaluates the same point
    if ((fm < 0.0) == (flo < 0.0) && fm != 0.0) {
      lo = md;
      flo = fm;
    } else {
      hi = md;
    }
  }
  double m[4];
  pair_mid(G, S, 0.5 * (lo + hi), m);
  if (lane < 4) S.m[lane] = m[lane];
  __syncthreads();
  return true;
}

__global__ __launch_bounds__(64) void k_cluster_solve(ClusterArgs A) {
  __shared__ ClusterLds S;
  const int c = blockIdx.x, lane = threadIdx.x;
  const int o = A.off[c], k = A.off[c + 1] - o, n = 2 * k;
  const int N = A.N;
  stage_bp(A.m, S.sb1, S.sb2);
  const double lam = A.lam[0];
  // inputs: v0, w, the map's vdot (slice sum, the epilogue's order), and ext =
  // vdot minus the cluster's own terms at the iterate (v_mid, G_eff of the map)
  double vm0 = 0.0, vm1 = 0.0, ge0 = 0.0, ge1 = 0.0, dv0 = 0.0, dv1 = 0.0;
  if (lane < k) {
    const int a = A.idx[o + lane];
    S.v0[2 * lane] = A.v0[a];
    S.v0[2 * lane + 1] = A.v0[N + a];
    S.w[lane] = A.w[a];
    vm0 = A.vmid[a];
    vm1 = A.vmid[N + a];
    const double iw = 1.0 / A.w[a];
    ge0 = -(A.dS[a] + lam * A.dv[a]) * iw;
    ge1 = -(A.dS[N + a] + lam * A.dv[N + a]) * iw;
    for (int s = 0; s < A.S; ++s) {
      dv0 += A.part[(size_t)s * 2 * N + a];
      dv1 += A.part[(size_t)s * 2 * N + N + a];
    }
    S.mn[2 * lane] = vm0;
    S.mn[2 * lane + 1] = vm1;
    S.g[2 * lane] = ge0;
    S.g[2 * lane + 1] = ge1;
  }
  __syncthreads();
  if (lane < k) {
    double e0 = dv0, e1 = dv1;
    for (int j = 0; j < k; ++j) {
      if (j == lane) continue;
      const double d0 = vm0 - S.mn[2 * j], d1 = vm1 - S.mn[2 * j + 1];
      const double y0 = S.g[2 * j] - ge0, y1 = S.g[2 * j + 1] - ge1;
      const double qd = d0 * d0 + d1 * d1;
      if (!(qd >= 1e-24)) continue;
      const double dy = (d0 * y0 + d1 * y1) / qd, inv = 1.0 / sqrt(qd);
      e0 -= S.w[j] * (y0 - d0 * dy) * inv;
      e1 -= S.w[j] * (y1 - d1 * dy) * inv;
    }
    S.ext[2 * lane] = e0;
    S.ext[2 * lane + 1] = e1;
  }
  __syncthreads();
  double scale = 1.0;
  if (lane < n) scale = fabs(S.v0[lane]) + 1.0;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) scale = fmax(scale, __shfl_xor(scale, off, 64));
  const double tol_final = 1e-13 * scale, tol_path = 1e-12 * scale;
  int its = 0;
  bool ok = false;
  double* mr = A.mroot + 2 * o;
  // have_root: 1 = a root of an earlier map, 0 = none, -f = f failed solves this
  // step.  After kClusterTries failures the cluster keeps the plain rows for the
  // rest of the step.
  const int state = A.have_root[c];
  const bool skip = state <= -kClusterTries;
  if (k == 2 && !skip) {  // the one-unknown form; the reference angle is the previous root's
    double psi_ref = 0.0;
    if (state > 0) {
      const PairGeo G = pair_geo(S, A.dt);
      double r = atan2(mr[1] - mr[3], mr[0] - mr[2]) - G.pa;
      while (r > 3.141592653589793) r -= 2.0 * 3.141592653589793;
      while (r < -3.141592653589793) r += 2.0 * 3.141592653589793;
      if (fabs(r) < 0.5 * 3.141592653589793) psi_ref = r;
    }
    ok = pair_solve(S, A, lam, psi_ref);
  } else if (state > 0) {
    if (lane < n) S.m[lane] = mr[lane];
    __syncthreads();
    ok = cl_newton(S, A, k, lam, S.m, 1.0, 10, tol_final, its);
  }
  if (!ok && !skip && k > 2) {
    // the homotopy in the coupling strength from t = 0: the cluster drifts with ext alone
    if (lane < n) {
      S.m[lane] = S.v0[lane] + 0.5 * A.dt * S.ext[lane];
      S.mt0[lane] = S.m[lane];
      S.mp[lane] = S.m[lane];
    }
    __syncthreads();
    double t = 0.0, tp = 0.0, dtau = 0.05;
    bool have_prev = false, path_ok = true;
    while (t < 1.0) {
      if (its > kClusterPathIts) {
        path_ok = false;
        break;
      }
      const double t1 = fmin(1.0, t + dtau);
      // secant predictor into S.F-free scratch: S.dl holds the guess
      double guess = 0.0;
      if (lane < n) guess = have_prev ? S.m[lane] + (S.m[lane] - S.mp[lane]) * (t1 - t) / (t - tp) : S.m[lane];
      __syncthreads();
      if (lane < n) S.dl[lane] = guess;
      __syncthreads();
      // Newton works in place on its own array: copy the guess into mt0's twin (S.mn is used inside)
      double keep = lane < n ? S.m[lane] : 0.0;
      if (lane < n) S.m[lane] = S.dl[lane];
      __syncthreads();
      if (cl_newton(S, A, k, lam, S.m, t1, 8, tol_path, its)) {
        if (lane < n) S.mp[lane] = keep;
        tp = t;
        t = t1;
        have_prev = true;
        dtau = fmin(1.5 * dtau, 0.5);
      } else {
        if (lane < n) S.m[lane] = keep;
        dtau *= 0.5;
        if (dtau < 1e-7) path_ok = false;
      }
      __syncthreads();
      if (!path_ok) break;
    }
    ok = path_ok && cl_newton(S, A, k, lam, S.m, 1.0, 20, tol_final, its);
    if (!ok) {
      // In the stiff limit the cluster's own dynamics lines its markers up along
      // an eigenvector of Hess L (every pair term vanishes there): Newton from
      // the t = 0 configuration projected on the line through its weighted
      // centre along each eigenvector, the smaller eigenvalue's first.
      double W = lane < k ? S.w[lane] : 0.0;
      double cx = lane < k ? S.w[lane] * S.mt0[2 * lane] : 0.0;
      double cy = lane < k ? S.w[lane] * S.mt0[2 * lane + 1] : 0.0;
      W = wave_sum(W);
      cx = wave_sum(cx) / W;
      cy = wave_sum(cy) / W;
      double g[2], H[3];
      if (grad_hess_dev(A.m, S.sb1, S.sb2, A.L, cx, cy, g, H)) {
        const double th = 0.5 * atan2(2.0 * H[1], H[0] - H[2]);
        const double ca = cos(th), sa = sin(th);
        const double la = H[0] * ca * ca + 2.0 * H[1] * ca * sa + H[2] * sa * sa;
        const double lb = H[0] * sa * sa - 2.0 * H[1] * ca * sa + H[2] * ca * ca;
        for (int q = 0; q < 2 && !ok && its <= kClusterIts; ++q) {
          const bool first_a = la <= lb;
          const bool use_a = (q == 0) == first_a;
          const double e0 = use_a ? ca : -sa, e1 = use_a ? sa : ca;
          if (lane < k) {
            const double pr = (S.mt0[2 * lane] - cx) * e0 + (S.mt0[2 * lane + 1] - cy) * e1;
            S.m[2 * lane] = cx + pr * e0;
            S.m[2 * lane + 1] = cy + pr * e1;
          }
          __syncthreads();
          ok = cl_newton(S, A, k, lam, S.m, 1.0, 20, tol_final, its);
        }
      }
    }
  }
  // hat rows: 2 m - v0, or the plain map value v0 + dt vdot if no root was reached
  if (lane < k) {
    const int a = A.idx[o + lane];
    const size_t r = (size_t)(o + lane);
    if (ok) {
      A.x_out[2 * r] = 2.0 * S.m[2 * lane] - S.v0[2 * lane];
      A.x_out[2 * r + 1] = 2.0 * S.m[2 * lane + 1] - S.v0[2 * lane + 1];
      mr[2 * lane] = S.m[2 * lane];
      mr[2 * lane + 1] = S.m[2 * lane + 1];
    } else {
      A.x_out[2 * r] = A.v0[a] + A.dt * dv0;
      A.x_out[2 * r + 1] = A.v0[N + a] + A.dt * dv1;
    }
  }
  if (lane == 0) {
    if (!skip) {
      A.have_root[c] = ok ? 1 : (state > 0 ? -1 : state - 1);
      if (!ok) atomicAdd(A.fail_count, 1);
    }
    A.newton_its[c] += its;
  }
}

}  // namespace

void launch_cell_sort(const MeshView& m, const double* v, int N, int* cell_of, int* cnt, int* start, int* sorted,
                      hipStream_t st) {
  const int ncell = m.ne1 * m.ne2;
  hipLaunchKernelGGL(k_fill_int, dim3((ncell + 255) / 256), dim3(256), 0, st, cnt, ncell, 0);
  hipLaunchKernelGGL(k_cell_count, dim3((N + 255) / 256), dim3(256), 0, st, m, v, N, cell_of, cnt);
  hipLaunchKernelGGL(k_cell_scan, dim3(1), dim3(1024), 0, st, cnt, ncell, start);
  hipLaunchKernelGGL(k_cell_fill, dim3((N + 255) / 256), dim3(256), 0, st, cell_of, N, start, cnt, sorted);
}

void launch_near_cells(const MeshView& m, const double* v, const double* w, int N, double theta, double w_max,
                       const int* start, const int* sorted, NearLink* out, int* count, int cap, hipStream_t st) {
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, st, count, 1, 0);
  const double r = w_max / theta;
  hipLaunchKernelGGL(k_near_cells, dim3(m.ne1 * m.ne2), dim3(256), 0, st, m, v, w, N, theta * theta, r, start,
                     sorted, out, count, cap);
}

void launch_fill_int(int* x, int n, int value, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_int, dim3((n + 255) / 256), dim3(256), 0, st, x, n, value);
}

void launch_cluster_rows(const int* idx, int total, bool set, int* row, hipStream_t st) {
  if (total <= 0) return;
  hipLaunchKernelGGL(k_cluster_rows, dim3((total + 255) / 256), dim3(256), 0, st, idx, total, set ? 1 : 0, row);
}

void launch_cluster_solve(const ClusterArgs& a, int ncl, hipStream_t st) {
  if (ncl <= 0) return;
  hipLaunchKernelGGL(k_cluster_solve, dim3(ncl), dim3(64), 0, st, a);
}

}  // namespace collop
