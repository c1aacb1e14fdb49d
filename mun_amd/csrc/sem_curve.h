// Closed-curve velocity anomaly on the GPU (the reference's parameterisation).
//
// Restates Bspline_Coupling/SEM/sem_waveform/velocity.py:8-202 and the
// analytic projection of core.py:902-970:
//   curve   C(u), u in [0,1]: cubic B-spline with control points
//           P = [ctrl0..ctrl5, ctrl0, ctrl0, ctrl0] and clamped knots
//           [0,0,0,0, 1/6, 2/6, 3/6, 4/6, 5/6, 1,1,1,1]
//   sd      nearest of `samples` uniform curve samples, then `newton_steps`
//           Newton steps on phi(u) = |C(u) - X|^2 / 2 (clipped to [0,1]);
//           sign negative inside the sampled polygon (crossing-number test)
//   v       = v_out + (v_in - v_out) sigmoid(-sd / tau)
//   grad    g9_j = sum_nodes dL/dv * dv/dsd * N_j(u*) * n_out, then the three
//           copies of P0 fold back into ctrl0; output interleaved [x0,z0,...].
#pragma once
#include <hip/hip_runtime.h>

namespace sem {

constexpr int CURVE_MAX_SAMPLES = 2048;

__device__ inline void curve_knots(double* t) {
    for (int i = 0; i < 4; ++i) t[i] = 0.0;
    for (int i = 1; i <= 5; ++i) t[3 + i] = i * (1.0 / 6.0);
    for (int i = 9; i < 13; ++i) t[i] = 1.0;
}

// Basis values and first two derivatives of the degree-3 basis functions that
// are non-zero on the span of u (The NURBS Book, algorithm A2.3).
__device__ inline int curve_ders(double u, const double* t, double ders[3][4]) {
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
    return s - 3;  // index of the first non-zero basis function
}

__device__ inline void curve_eval(const double (*P)[2], const double* t, double u, double* C, double* C1,
                                  double* C2) {
    double d[3][4];
    const int f = curve_ders(u, t, d);
    for (int q = 0; q < 2; ++q) {
        double c0 = 0.0, c1 = 0.0, c2 = 0.0;
        for (int r = 0; r < 4; ++r) {
            c0 += d[0][r] * P[f + r][q];
            c1 += d[1][r] * P[f + r][q];
            c2 += d[2][r] * P[f + r][q];
        }
        C[q] = c0;
        if (C1) C1[q] = c1;
        if (C2) C2[q] = c2;
    }
}

// One workgroup samples the curve into LDS, then every thread handles nodes in
// a grid-stride loop.  aux = [sd | u* | n_out.x | n_out.z] planes of N values.
template <typename T>
__global__ __launch_bounds__(256) void curve_velocity_kernel(const double* ctrl, int S, int newton, double v_in,
                                                             double v_out, double tau, const double* x1d,
                                                             const double* z1d, int NX, int NZ, double* c, T* c2,
                                                             double* aux) {
    __shared__ double smp[CURVE_MAX_SAMPLES][2];
    __shared__ double t[13];
    __shared__ double P[9][2];
    const long m = blockIdx.y;
    const long N = (long)NX * NZ;
    if (threadIdx.x == 0) curve_knots(t);
    if (threadIdx.x < 18) {
        const int k = threadIdx.x / 2, q = threadIdx.x % 2;
        P[k][q] = ctrl[m * 12 + (k < 6 ? k : 0) * 2 + q];
    }
    __syncthreads();
    const double step = S > 1 ? 1.0 / (S - 1) : 0.0;
    for (int k = threadIdx.x; k < S; k += blockDim.x) {
        const double u = (k == S - 1) ? 1.0 : k * step;
        double C[2];
        curve_eval(P, t, u, C, nullptr, nullptr);
        smp[k][0] = C[0];
        smp[k][1] = C[1];
    }
    __syncthreads();
    const double ts = tau > 1e-12 ? tau : 1e-12;
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
        double u = (best == S - 1) ? 1.0 : best * step;
        for (int it = 0; it < newton; ++it) {
            double C[2], C1[2], C2[2];
            curve_eval(P, t, u, C, C1, C2);
            const double rx = C[0] - X, rz = C[1] - Z;
            const double f1 = rx * C1[0] + rz * C1[1];
            double f2 = C1[0] * C1[0] + C1[1] * C1[1] + rx * C2[0] + rz * C2[1];
            if (fabs(f2) < 1e-14) f2 = 1e-14;
            u = fmin(fmax(u - f1 / f2, 0.0), 1.0);
        }
        double Cb[2];
        curve_eval(P, t, u, Cb, nullptr, nullptr);
        const double rx = X - Cb[0], rz = Z - Cb[1];
        const double dist = sqrt(rx * rx + rz * rz);
        // crossing-number point-in-polygon on the closed sample polygon
        bool inside = false;
        for (int q = 0, p = S - 1; q < S; p = q++) {
            const double xq = smp[q][0], zq = smp[q][1], xp = smp[p][0], zp = smp[p][1];
            if ((zq > Z) != (zp > Z) && X < (xp - xq) * (Z - zq) / (zp - zq) + xq) inside = !inside;
        }
        const double sgn = inside ? -1.0 : 1.0;
        const double sd = sgn * dist;
        const double x = fmin(fmax(sd / ts, -50.0), 50.0);
        const double sig = 1.0 / (1.0 + exp(x));
        const double v = v_out + (v_in - v_out) * sig;
        if (c) c[m * N + k] = v;
        c2[m * N + k] = (T)(v * v);
        double* A = aux + m * 4 * N;
        A[k] = sd;
        A[N + k] = u;
        A[2 * N + k] = sgn * rx / (dist + 1e-15);
        A[3 * N + k] = sgn * rz / (dist + 1e-15);
    }
}

// Per-block partial sums of the 18 projection terms (9 basis functions x 2 axes).
__global__ __launch_bounds__(256) void curve_project_kernel(const double* dLdv, const double* aux, double v_in,
                                                            double v_out, double tau, long N, double* partial) {
    __shared__ double red[256][18];
    __shared__ double t[13];
    if (threadIdx.x == 0) curve_knots(t);
    __syncthreads();
    const long m = blockIdx.y;
    const double* A = aux + m * 4 * N;
    const double ts = tau > 1e-12 ? tau : 1e-12;
    const double thr = 1e-8 * (v_out - v_in);
    double acc[18];
    for (int j = 0; j < 18; ++j) acc[j] = 0.0;
    for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < N; k += (long)gridDim.x * blockDim.x) {
        const double x = fmin(fmax(A[k] / ts, -50.0), 50.0);
        const double sig = 1.0 / (1.0 + exp(x));
        const double dv = (v_in - v_out) * sig * (1.0 - sig) * (-1.0 / ts);
        if (!(fabs(dv) > thr)) continue;
        const double w = dLdv[m * N + k] * dv;
        double d[3][4];
        const int f = curve_ders(A[N + k], t, d);
        for (int r = 0; r < 4; ++r) {
            acc[f + r] += w * d[0][r] * A[2 * N + k];
            acc[9 + f + r] += w * d[0][r] * A[3 * N + k];
        }
    }
    for (int j = 0; j < 18; ++j) red[threadIdx.x][j] = acc[j];
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int j = 0; j < 18; ++j) red[threadIdx.x][j] += red[threadIdx.x + s][j];
        __syncthreads();
    }
    if (threadIdx.x < 18) partial[(m * gridDim.x + blockIdx.x) * 18 + threadIdx.x] = red[0][threadIdx.x];
}

// Fold the partials in a fixed order, collapse P6..P8 into P0, interleave.
__global__ void curve_project_finish_kernel(const double* partial, int nblk, double* grad) {
    const long m = blockIdx.x;
    if (threadIdx.x != 0) return;
    double g[18];
    for (int j = 0; j < 18; ++j) g[j] = 0.0;
    for (int b = 0; b < nblk; ++b)
        for (int j = 0; j < 18; ++j) g[j] += partial[(m * nblk + b) * 18 + j];
    for (int axis = 0; axis < 2; ++axis) {
        const double* h = g + 9 * axis;
        grad[m * 12 + 0 + axis] = h[0] + h[6] + h[7] + h[8];
        for (int j = 1; j < 6; ++j) grad[m * 12 + 2 * j + axis] = h[j];
    }
}

}  // namespace sem
