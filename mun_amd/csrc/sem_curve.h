// Closed-curve velocity anomaly on the GPU (the reference's two parameterisations).
//
// CLAMPED6 restates Bspline_Coupling/SEM/sem_waveform/velocity.py:8-202 and the
// analytic projection of core.py:902-970:
//   curve   C(u), u in [0,1]: cubic B-spline with control points
//           P = [ctrl0..ctrl5, ctrl0, ctrl0, ctrl0] and clamped knots
//           [0,0,0,0, 1/6, 2/6, 3/6, 4/6, 5/6, 1,1,1,1]; samples u_k = k/(S-1)
//   grad    the three copies of P0 fold back into ctrl0; nodes kept where
//           |dv/dsd| > 1e-8 (v_out - v_in)
// PERIODIC restates CO2_plume/one_stage/single_cubic/sem_waveform/velocity.py:9-152
// and core.py:1075-1185:
//   curve   C(t), t in [3, K+3): uniform cubic B-spline with control points
//           [ctrl0..ctrl_{K-1}, ctrl0, ctrl1, ctrl2] and knots 0, 1, ..., K+6;
//           samples t_k = 3 + kK/S (endpoint excluded); u* stored as (t-3)/K
//   grad    closed index j folds onto ctrl (j mod K); nodes kept where
//           |dv/dsd| > 1e-8 max(1, |v_in - v_bg|)
// Both:
//   sd      nearest of S curve samples, then Newton steps on |C - X|^2 / 2
//           (parameter clipped to the curve's interval); sign negative inside
//           the sampled polygon (crossing-number test)
//   v       smooth : v_bg + (v_in - v_bg) sigmoid(-sd / tau)
//           replace: v_in where sd < 0, else v_bg
//   grad    sum over nodes of dL/dv * dv/dsd * N_j(u*) * n_out, interleaved
//           [x0, z0, x1, z1, ...]
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/sem_capi.h"

namespace sem {

constexpr int CURVE_MAX_SAMPLES = 2048;
constexpr int CURVE_MAX_CLOSED = SEM_CURVE_MAX_CTRL + 3;

struct CurveParams {
    int kind, K, S, newton, blend;
    double v_in, v_out, tau;
    const double* bg;  // per-node background or null
};

// Curve geometry shared by a workgroup (LDS).
struct CurveGeom {
    int kind, K, ncl;         // ncl: closed control count (9 or K + 3)
    double t[13];             // clamped knots (CLAMPED6)
    double P[CURVE_MAX_CLOSED][2];
    double lo, hi;            // parameter interval, Newton clip [lo, hi]
};

__device__ inline void clamped_knots(double* t) {
    for (int i = 0; i < 4; ++i) t[i] = 0.0;
    for (int i = 1; i <= 5; ++i) t[3 + i] = i * (1.0 / 6.0);
    for (int i = 9; i < 13; ++i) t[i] = 1.0;
}

// Basis values and first two derivatives of the degree-3 basis functions that
// are non-zero on the span of u, clamped knots (The NURBS Book, algorithm A2.3).
__device__ inline int clamped_ders(double u, const double* t, double ders[3][4]) {
    int s = 3;
    while (s < 8 && u >= t[s + 1]) ++s;
    double ndu[4][4], left[4], right[4];
    ndu[0][0] = 1.0;
    for (int j = 1; j <= 3; ++j) {
        left[j] = u - t[s + 1 - j];
        right[j] = t[s + j] - u;
        double saved = 0.0;
        for (int r = 0; r < j; ++r) {
            ndu[j][r] = right[r + 1] + left[j - r];
            const double tmp = ndu[r][j - 1] / ndu[j][r];
            ndu[r][j] = saved + right[r + 1] * tmp;
            saved = left[j - r] * tmp;
        }
        ndu[j][j] = saved;
    }
    for (int j = 0; j <= 3; ++j) ders[0][j] = ndu[j][3];
    for (int r = 0; r <= 3; ++r) {
        double a[2][4];
        int s1 = 0, s2 = 1;
        a[0][0] = 1.0;
        for (int k = 1; k <= 2; ++k) {
            double d = 0.0;
            const int rk = r - k, pk = 3 - k;
            if (r >= k) {
                a[s2][0] = a[s1][0] / ndu[pk + 1][rk];
                d = a[s2][0] * ndu[rk][pk];
            }
            const int j1 = (rk >= -1) ? 1 : -rk;
            const int j2 = (r - 1 <= pk) ? k - 1 : 3 - r;
            for (int j = j1; j <= j2; ++j) {
                a[s2][j] = (a[s1][j] - a[s1][j - 1]) / ndu[pk + 1][rk + j];
                d += a[s2][j] * ndu[rk + j][pk];
            }
            if (r <= pk) {
                a[s2][k] = -a[s1][k - 1] / ndu[pk + 1][r];
                d += a[s2][k] * ndu[r][pk];
            }
            ders[k][r] = d;
            const int tmp = s1;
            s1 = s2;
            s2 = tmp;
        }
    }
    for (int j = 0; j <= 3; ++j) {
        ders[1][j] *= 3.0;
        ders[2][j] *= 6.0;
    }
    return s - 3;  // closed index of the first non-zero basis function
}

// Uniform cubic B-spline on unit knot spacing: span s = floor(t), tau = t - s,
// basis functions of closed controls s-3 .. s and their derivatives in t.
__device__ inline int uniform_ders(double t, double ders[3][4]) {
    const double s = floor(t);
    const double x = t - s, y = 1.0 - x;
    const double x2 = x * x, x3 = x2 * x;
    ders[0][0] = y * y * y / 6.0;
    ders[0][1] = (3.0 * x3 - 6.0 * x2 + 4.0) / 6.0;
    ders[0][2] = (-3.0 * x3 + 3.0 * x2 + 3.0 * x + 1.0) / 6.0;
    ders[0][3] = x3 / 6.0;
    ders[1][0] = -0.5 * y * y;
    ders[1][1] = 1.5 * x2 - 2.0 * x;
    ders[1][2] = -1.5 * x2 + x + 0.5;
    ders[1][3] = 0.5 * x2;
    ders[2][0] = y;
    ders[2][1] = 3.0 * x - 2.0;
    ders[2][2] = 1.0 - 3.0 * x;
    ders[2][3] = x;
    return (int)s - 3;
}

__device__ inline int curve_ders(const CurveGeom& g, double u, double d[3][4]) {
    return g.kind == SEM_CURVE_PERIODIC ? uniform_ders(u, d) : clamped_ders(u, g.t, d);
}

__device__ inline void curve_eval(const CurveGeom& g, double u, double* C, double* C1, double* C2) {
    double d[3][4];
    const int f = curve_ders(g, u, d);
    for (int q = 0; q < 2; ++q) {
        double c0 = 0.0, c1 = 0.0, c2 = 0.0;
        for (int r = 0; r < 4; ++r) {
            c0 += d[0][r] * g.P[f + r][q];
            c1 += d[1][r] * g.P[f + r][q];
            c2 += d[2][r] * g.P[f + r][q];
        }
        C[q] = c0;
        if (C1) C1[q] = c1;
        if (C2) C2[q] = c2;
    }
}

// Thread 0 fills the geometry of model m (caller synchronises).
__device__ inline void curve_setup(CurveGeom& g, const CurveParams& p, const double* ctrl, long m) {
    g.kind = p.kind;
    g.K = p.K;
    g.ncl = p.K + 3;
    const double* c = ctrl + m * 2 * p.K;
    if (p.kind == SEM_CURVE_PERIODIC) {
        for (int j = 0; j < g.ncl; ++j) {
            g.P[j][0] = c[2 * (j % p.K)];
            g.P[j][1] = c[2 * (j % p.K) + 1];
        }
        g.lo = 3.0;
        g.hi = nextafter((double)(p.K + 3), 3.0);
    } else {
        clamped_knots(g.t);
        for (int j = 0; j < g.ncl; ++j) {
            const int k = j < 6 ? j : 0;
            g.P[j][0] = c[2 * k];
            g.P[j][1] = c[2 * k + 1];
        }
        g.lo = 0.0;
        g.hi = 1.0;
    }
}

__device__ inline double curve_sample_param(const CurveGeom& g, int k, int S) {
    // numpy.linspace's rounding: k * step + start
    if (g.kind == SEM_CURVE_PERIODIC) return (double)k * (((double)(g.K + 3) - 3.0) / (double)S) + 3.0;
    return (k == S - 1) ? 1.0 : (S > 1 ? k * (1.0 / (S - 1)) : 0.0);
}

// One workgroup per (block, model) samples the curve into LDS, then every
// thread handles nodes in a grid-stride loop.  aux = [sd | u* | n_out.x |
// n_out.z] planes of N values.
template <typename T>
__global__ __launch_bounds__(256) void curve_velocity_kernel(CurveParams p, const double* ctrl, const double* x1d,
                                                             const double* z1d, int NX, int NZ, double* c, T* c2,
                                                             double* aux) {
    __shared__ double smp[CURVE_MAX_SAMPLES][2];
    __shared__ CurveGeom g;
    const long m = blockIdx.y;
    const long N = (long)NX * NZ;
    const int S = p.S;
    if (threadIdx.x == 0) curve_setup(g, p, ctrl, m);
    __syncthreads();
    for (int k = threadIdx.x; k < S; k += blockDim.x) {
        double C[2];
        curve_eval(g, curve_sample_param(g, k, S), C, nullptr, nullptr);
        smp[k][0] = C[0];
        smp[k][1] = C[1];
    }
    __syncthreads();
    const double ts = p.tau > 1e-12 ? p.tau : 1e-12;
    for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < N; k += (long)gridDim.x * blockDim.x) {
        const double X = x1d[k % NX], Z = z1d[k / NX];
        int best = 0;
        double bd = 1e300;
        for (int q = 0; q < S; ++q) {
            const double dx = X - smp[q][0], dz = Z - smp[q][1];
            const double d2 = dx * dx + dz * dz;
            if (d2 < bd) {
                bd = d2;
                best = q;
            }
        }
        double u = curve_sample_param(g, best, S);
        for (int it = 0; it < p.newton; ++it) {
            double C[2], C1[2], C2[2];
            curve_eval(g, u, C, C1, C2);
            const double rx = C[0] - X, rz = C[1] - Z;
            const double f1 = rx * C1[0] + rz * C1[1];
            double f2 = C1[0] * C1[0] + C1[1] * C1[1] + rx * C2[0] + rz * C2[1];
            if (fabs(f2) < 1e-14) f2 = 1e-14;
            u = fmin(fmax(u - f1 / f2, g.lo), g.hi);
        }
        double Cb[2];
        curve_eval(g, u, Cb, nullptr, nullptr);
        const double rx = X - Cb[0], rz = Z - Cb[1];
        const double dist = sqrt(rx * rx + rz * rz);
        // crossing-number point-in-polygon on the closed sample polygon
        bool inside = false;
        for (int q = 0, r = S - 1; q < S; r = q++) {
            const double xq = smp[q][0], zq = smp[q][1], xr = smp[r][0], zr = smp[r][1];
            if ((zq > Z) != (zr > Z) && X < (xr - xq) * (Z - zq) / (zr - zq) + xq) inside = !inside;
        }
        const double sgn = inside ? -1.0 : 1.0;
        const double sd = sgn * dist;
        const double vb = p.bg ? p.bg[k] : p.v_out;
        double v;
        if (p.blend == SEM_BLEND_REPLACE) {
            v = sd < 0.0 ? p.v_in : vb;
        } else {
            const double x = fmin(fmax(sd / ts, -50.0), 50.0);
            const double sig = 1.0 / (1.0 + exp(x));
            v = vb + (p.v_in - vb) * sig;
        }
        if (c) c[m * N + k] = v;
        c2[m * N + k] = (T)(v * v);
        double* A = aux + m * 4 * N;
        A[k] = sd;
        A[N + k] = (g.kind == SEM_CURVE_PERIODIC) ? (u - 3.0) / (double)g.K : u;
        A[2 * N + k] = sgn * rx / (dist + 1e-15);
        A[3 * N + k] = sgn * rz / (dist + 1e-15);
    }
}

constexpr int CURVE_PROJ_THREADS = 64;

// Per-block partial sums over the closed controls (2 axes x ncl): each thread
// accumulates its nodes in its own LDS row, then rows are summed in a fixed
// order (deterministic).
__global__ __launch_bounds__(CURVE_PROJ_THREADS) void curve_project_kernel(CurveParams p, const double* dLdv,
                                                                           const double* aux, long N,
                                                                           double* partial) {
    __shared__ double acc[CURVE_PROJ_THREADS][2 * CURVE_MAX_CLOSED];
    __shared__ CurveGeom g;
    const long m = blockIdx.y;
    const int tid = threadIdx.x;
    const int ncl = p.K + 3, w2 = 2 * ncl;
    if (tid == 0) {
        g.kind = p.kind;
        g.K = p.K;
        g.ncl = ncl;
        if (p.kind != SEM_CURVE_PERIODIC) clamped_knots(g.t);
    }
    for (int j = 0; j < w2; ++j) acc[tid][j] = 0.0;
    __syncthreads();
    const double* A = aux + m * 4 * N;
    const double ts = p.tau > 1e-12 ? p.tau : 1e-12;
    for (long k = blockIdx.x * (long)blockDim.x + tid; k < N; k += (long)gridDim.x * blockDim.x) {
        const double vb = p.bg ? p.bg[k] : p.v_out;
        const double x = fmin(fmax(A[k] / ts, -50.0), 50.0);
        const double sig = 1.0 / (1.0 + exp(x));
        const double dv = (p.v_in - vb) * sig * (1.0 - sig) * (-1.0 / ts);
        const double thr =
            (p.kind == SEM_CURVE_PERIODIC) ? 1e-8 * fmax(1.0, fabs(p.v_in - vb)) : 1e-8 * (vb - p.v_in);
        if (!(fabs(dv) > thr)) continue;
        const double w = dLdv[m * N + k] * dv;
        double u = A[N + k];
        if (p.kind == SEM_CURVE_PERIODIC) {
            const double un = u - floor(u);  // mod 1
            u = fmin(fmax(3.0 + un * (double)p.K, 3.0), nextafter((double)(p.K + 3), 3.0));
        }
        double d[3][4];
        const int f = curve_ders(g, u, d);
        for (int r = 0; r < 4; ++r) {
            acc[tid][f + r] += w * d[0][r] * A[2 * N + k];
            acc[tid][ncl + f + r] += w * d[0][r] * A[3 * N + k];
        }
    }
    __syncthreads();
    for (int s = CURVE_PROJ_THREADS / 2; s > 0; s >>= 1) {
        if (tid < s)
            for (int j = 0; j < w2; ++j) acc[tid][j] += acc[tid + s][j];
        __syncthreads();
    }
    for (int j = tid; j < w2; j += blockDim.x) partial[(m * gridDim.x + blockIdx.x) * w2 + j] = acc[0][j];
}

// Fold the partials in a fixed order, collapse the appended controls, interleave.
__global__ void curve_project_finish_kernel(CurveParams p, const double* partial, int nblk, double* grad) {
    const long m = blockIdx.x;
    if (threadIdx.x != 0) return;
    const int K = p.K, ncl = K + 3, w2 = 2 * ncl;
    double h[2 * CURVE_MAX_CLOSED];
    for (int j = 0; j < w2; ++j) h[j] = 0.0;
    for (int b = 0; b < nblk; ++b)
        for (int j = 0; j < w2; ++j) h[j] += partial[(m * nblk + b) * w2 + j];
    for (int axis = 0; axis < 2; ++axis) {
        const double* a = h + ncl * axis;
        for (int j = 0; j < K; ++j) {
            double s = a[j];
            if (p.kind == SEM_CURVE_PERIODIC) {
                if (j < 3) s += a[K + j];
            } else if (j == 0) {
                s = a[0] + a[6] + a[7] + a[8];
            }
            grad[m * 2 * K + 2 * j + axis] = s;
        }
    }
}

}  // namespace sem
