// Device kernels of the MI355X SEM solver (gfx950 / CDNA4).
//
// Discretisation (restated from Bspline_Coupling/SEM/sem_waveform/operators.py:6-42,
// 65-84, 193-242 and core.py:1197-1275):
//   element stiffness  Ke = Gx^T W Gx + Gz^T W Gz,  W = diag(w_a w_b J c^2)
//   Gx = (2/dx) kron(I, D), Gz = (2/dz) kron(D, I)   (GLL collocation: L = I)
//   lumped mass        M = J Wx(ix) Wz(jz)   (Wx, Wz: assembled 1-D GLL weights)
//   sponge             Dn = exp(-(sx + sz) dt) = Dx(ix) Dz(jz)
//
// Time stepping is in increment form.  With w^n = u^n - D u^{n-1} the
// reference recurrence u^{n+1} = D[2u^n - D u^{n-1} + a^n] becomes
//   w^{n+1} = D (w^n + a^n),   u^{n+1} = D u^n + w^{n+1},
//   a^n = dt^2 M^-1 (s_n M_src e_src - K u^n),
// which is the same recurrence in exact arithmetic but keeps the small
// increment at full FP32 precision.  The adjoint uses the same step with the
// receiver residual as source (and D = 1 in the reference mode).
//
// Work decomposition: one thread per element.  A 256-thread workgroup covers a
// 16 x 16 block of elements: 15 x 15 owned elements plus one halo column on
// the left and one halo row below.  Each thread keeps its element's 5 x 5
// values in registers and computes the 25 local contributions; element-edge
// contributions reach the owning element through a 9-value LDS exchange.
// Every node is written by exactly one thread (owner computes, no atomics), so
// results are bitwise deterministic and independent of batch composition.
#pragma once
#include <hip/hip_runtime.h>

namespace sem {

constexpr int EX = 16;            // element threads along x per workgroup
constexpr int EZ = 16;            // element threads along z per workgroup
constexpr int TX = EX - 1;        // owned elements along x per tile
constexpr int TZ = EZ - 1;        // owned elements along z per tile
constexpr int NTHREADS = EX * EZ;

enum StepMode { MODE_FWD = 0, MODE_ADJ = 1, MODE_REC = 2 };

template <typename T, int P>
struct ElemConst {
    T D[P + 1][P + 1];  // D[a][k] = l_k'(xi_a) on the reference element
    T w[P + 1];         // GLL weights
    T cx, cz;           // (2/dx)^2, (2/dz)^2
};

template <typename T, int P>
struct RecEntry {
    int ex, ez;         // element holding the receiver
    int r;              // receiver column within its set
    int owned;          // 1 in the tile that owns the element (writes the sample)
    T lx[P + 1];        // Lagrange weights along x
    T lz[P + 1];        // Lagrange weights along z
};

template <typename T, int P>
struct StepArgs {
    ElemConst<T, P> ec;
    int nex, nez, NX, NZ;
    int ntx, ntiles;
    const T* u_in;   long u_in_bs;     // u^n (or lambda_k), read with halo
    T* u_out;        long u_out_bs;    // u^{n+1} (or lambda_{k-1}), owned nodes
    T* w;            long w_bs;        // increment, updated in place
    const T* c2;     long c2_bs;
    const int* model_of;
    const T* invWx;                    // [NX]  1 / Wx
    const T* dtWz;                     // [NZ]  dt^2 / Wz
    const T* Dx;                       // [NX]  exp(-sx dt)
    const T* Dz;                       // [NZ]  exp(-sz dt)
    int damp;
    // point source (forward): a[src] += dt^2 s_n
    const int* src_node;
    const T* wavelet; long wav_bs; int n; T dt2;
    // receivers
    const int* rec_set_of;             // [batch] receiver set id
    const int* rec_ptr;                // [nsets][ntiles + 1] offsets into recs
    const RecEntry<T, P>* recs;
    T* seis;         long seis_bs; int rec_pitch; int rec_t;
    // checkpoint mirrors of (u^{n+1}, w^{n+1}), may be null
    T* ck_u;         T* ck_w;      long ck_bs;
    // adjoint: residual source, forward wavefield, gradient accumulator
    const T* resid;  long resid_bs; int res_t; T invJ;
    const T* uf;     long uf_bs;
    double* G;       long G_bs;
};

template <typename T, int P, int MODE>
__global__ __launch_bounds__(NTHREADS) void step_kernel(const StepArgs<T, P> a) {
    constexpr int NP = P + 1;
    __shared__ T xr_col[EZ][EX][NP];   // element's i = P column of contributions
    __shared__ T xr_row[EZ][EX][P];    // element's j = P row (i < P)
    __shared__ T xg_col[MODE == MODE_ADJ ? EZ : 1][EX][NP];
    __shared__ T xg_row[MODE == MODE_ADJ ? EZ : 1][EX][P];

    const int t = threadIdx.x;
    const int exl = t % EX, ezl = t / EX;
    const int tile = blockIdx.x;
    const int tx = tile % a.ntx, tz = tile / a.ntx;
    const long b = blockIdx.y;
    const int ex = tx * TX - 1 + exl;
    const int ez = tz * TZ - 1 + ezl;
    const bool valid = ex >= 0 && ex < a.nex && ez >= 0 && ez < a.nez;
    const bool owned = valid && exl >= 1 && ezl >= 1;
    const long NX = a.NX;
    const long base = (long)(P * ez) * NX + (long)(P * ex);

    T u[NP][NP];
    T r[NP][NP];
    T g[MODE == MODE_ADJ ? NP : 1][MODE == MODE_ADJ ? NP : 1];

#pragma unroll
    for (int j = 0; j < NP; ++j)
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            r[j][i] = T(0);
            u[j][i] = T(0);
        }
    if constexpr (MODE == MODE_ADJ) {
#pragma unroll
        for (int j = 0; j < NP; ++j)
#pragma unroll
            for (int i = 0; i < NP; ++i) g[j][i] = T(0);
    }

    if (valid) {
        const T* __restrict__ uin = a.u_in + b * a.u_in_bs + base;
#pragma unroll
        for (int j = 0; j < NP; ++j)
#pragma unroll
            for (int i = 0; i < NP; ++i) u[j][i] = uin[j * NX + i];
    }

    if constexpr (MODE == MODE_FWD || MODE == MODE_REC) {
        // seismogram sample of the field being read (R u at this time level)
        if (a.rec_t >= 0 && owned) {
            const int set = a.rec_set_of[b];
            const int beg = a.rec_ptr[set * (a.ntiles + 1) + tile];
            const int end = a.rec_ptr[set * (a.ntiles + 1) + tile + 1];
            for (int k = beg; k < end; ++k) {
                const RecEntry<T, P>& e = a.recs[k];
                if (e.owned && e.ex == ex && e.ez == ez) {
                    T acc = T(0);
#pragma unroll
                    for (int j = 0; j < NP; ++j) {
                        T row = T(0);
#pragma unroll
                        for (int i = 0; i < NP; ++i) row += e.lx[i] * u[j][i];
                        acc += e.lz[j] * row;
                    }
                    a.seis[b * a.seis_bs + (long)a.rec_t * a.rec_pitch + e.r] = acc;
                }
            }
        }
        if constexpr (MODE == MODE_REC) return;
    }

    if (valid) {
        const long m = a.model_of ? a.model_of[b] : b;
        const T* __restrict__ c2p = a.c2 + m * a.c2_bs + base;
        T cw[NP][NP];
#pragma unroll
        for (int j = 0; j < NP; ++j)
#pragma unroll
            for (int i = 0; i < NP; ++i) cw[j][i] = c2p[j * NX + i] * (a.ec.w[j] * a.ec.w[i]);

        // forward wavefield derivatives for the gradient correlation
        T dxu[MODE == MODE_ADJ ? NP : 1][MODE == MODE_ADJ ? NP : 1];
        T dzu[MODE == MODE_ADJ ? NP : 1][MODE == MODE_ADJ ? NP : 1];
        if constexpr (MODE == MODE_ADJ) {
            const T* __restrict__ ufp = a.uf + b * a.uf_bs + base;
            T v[NP][NP];
#pragma unroll
            for (int j = 0; j < NP; ++j)
#pragma unroll
                for (int i = 0; i < NP; ++i) v[j][i] = ufp[j * NX + i];
#pragma unroll
            for (int j = 0; j < NP; ++j)
#pragma unroll
                for (int q = 0; q < NP; ++q) {
                    T sx = T(0), sz = T(0);
#pragma unroll
                    for (int k = 0; k < NP; ++k) {
                        sx += a.ec.D[q][k] * v[j][k];
                        sz += a.ec.D[j][k] * v[k][q];
                    }
                    dxu[j][q] = sx;   // d/dxi at (j, q)
                    dzu[j][q] = sz;   // d/deta at (j, q)
                }
        }

        // x direction: rows
#pragma unroll
        for (int j = 0; j < NP; ++j) {
            T f[NP];
#pragma unroll
            for (int q = 0; q < NP; ++q) {
                T d = T(0);
#pragma unroll
                for (int k = 0; k < NP; ++k) d += a.ec.D[q][k] * u[j][k];
                f[q] = cw[j][q] * a.ec.cx * d;
                if constexpr (MODE == MODE_ADJ)
                    g[j][q] = (a.ec.w[j] * a.ec.w[q]) * a.ec.cx * (dxu[j][q] * d);
            }
#pragma unroll
            for (int i = 0; i < NP; ++i) {
                T s = T(0);
#pragma unroll
                for (int q = 0; q < NP; ++q) s += a.ec.D[q][i] * f[q];
                r[j][i] = s;
            }
        }
        // z direction: columns
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            T f[NP];
#pragma unroll
            for (int q = 0; q < NP; ++q) {
                T d = T(0);
#pragma unroll
                for (int k = 0; k < NP; ++k) d += a.ec.D[q][k] * u[k][i];
                f[q] = cw[q][i] * a.ec.cz * d;
                if constexpr (MODE == MODE_ADJ)
                    g[q][i] += (a.ec.w[q] * a.ec.w[i]) * a.ec.cz * (dzu[q][i] * d);
            }
#pragma unroll
            for (int j = 0; j < NP; ++j) {
                T s = T(0);
#pragma unroll
                for (int q = 0; q < NP; ++q) s += a.ec.D[q][j] * f[q];
                r[j][i] += s;
            }
        }

        if constexpr (MODE == MODE_ADJ) {
            // adjoint source R^T res, injected into the receiver element's slots
            const int set = a.rec_set_of[b];
            const int beg = a.rec_ptr[set * (a.ntiles + 1) + tile];
            const int end = a.rec_ptr[set * (a.ntiles + 1) + tile + 1];
            for (int k = beg; k < end; ++k) {
                const RecEntry<T, P>& e = a.recs[k];
                if (e.ex == ex && e.ez == ez) {
                    const T res = a.resid[b * a.resid_bs + (long)a.res_t * a.rec_pitch + e.r] * a.invJ;
#pragma unroll
                    for (int j = 0; j < NP; ++j)
#pragma unroll
                        for (int i = 0; i < NP; ++i) r[j][i] -= (e.lz[j] * e.lx[i]) * res;
                }
            }
        }
    }

#pragma unroll
    for (int j = 0; j < NP; ++j) xr_col[ezl][exl][j] = r[j][P];
#pragma unroll
    for (int i = 0; i < P; ++i) xr_row[ezl][exl][i] = r[P][i];
    if constexpr (MODE == MODE_ADJ) {
#pragma unroll
        for (int j = 0; j < NP; ++j) xg_col[ezl][exl][j] = g[j][P];
#pragma unroll
        for (int i = 0; i < P; ++i) xg_row[ezl][exl][i] = g[P][i];
    }
    __syncthreads();
    if (!owned) return;

    const int jmax = (ez == a.nez - 1) ? P : P - 1;
    const int imax = (ex == a.nex - 1) ? P : P - 1;
    const int src = (MODE == MODE_FWD && a.src_node) ? a.src_node[b] : -1;
    const T sterm = (src >= 0) ? a.dt2 * a.wavelet[b * a.wav_bs + a.n] : T(0);
    T* __restrict__ wp = a.w + b * a.w_bs;
    T* __restrict__ uo = a.u_out + b * a.u_out_bs;

#pragma unroll
    for (int j = 0; j < NP; ++j) {
        if (j > jmax) break;
        const int gz = P * ez + j;
        const T dz = a.dtWz[gz];
        const T Dzv = a.Dz[gz];
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            if (i > imax) break;
            T acc = r[j][i];
            if (i == 0) acc += xr_col[ezl][exl - 1][j];
            if (j == 0) acc += (i < P) ? xr_row[ezl - 1][exl][i] : xr_col[ezl - 1][exl][P];
            if (i == 0 && j == 0) acc += xr_col[ezl - 1][exl - 1][P];
            const int gx = P * ex + i;
            const long node = base + j * NX + i;
            T av = -(a.invWx[gx] * dz) * acc;
            if (node == src) av += sterm;
            const T Dn = a.damp ? a.Dx[gx] * Dzv : T(1);
            const T wn = Dn * (wp[node] + av);
            const T un = Dn * u[j][i] + wn;
            wp[node] = wn;
            uo[node] = un;
            if constexpr (MODE == MODE_FWD) {
                if (a.ck_u) {
                    a.ck_u[b * a.ck_bs + node] = un;
                    a.ck_w[b * a.ck_bs + node] = wn;
                }
            }
            if constexpr (MODE == MODE_ADJ) {
                T ga = g[j][i];
                if (i == 0) ga += xg_col[ezl][exl - 1][j];
                if (j == 0) ga += (i < P) ? xg_row[ezl - 1][exl][i] : xg_col[ezl - 1][exl][P];
                if (i == 0 && j == 0) ga += xg_col[ezl - 1][exl - 1][P];
                a.G[b * a.G_bs + node] += (double)ga;
            }
        }
    }
}

// r = seis - obs (first nrec columns of each row), misfit[b] = 0.5 * sum r^2 in double.
template <typename T>
__global__ __launch_bounds__(256) void residual_kernel(const T* seis, const T* obs, T* res, double* misfit,
                                                       const int* nrec_of, int nt, int pitch, long bs) {
    const long b = blockIdx.x;
    const int nrec = nrec_of[b];
    double acc = 0.0;
    const long total = (long)nt * pitch;
    for (long k = threadIdx.x; k < total; k += blockDim.x) {
        const int c = (int)(k % pitch);
        const T rv = (c < nrec) ? seis[b * bs + k] - obs[b * bs + k] : T(0);
        res[b * bs + k] = rv;
        acc += (double)rv * (double)rv;
    }
    __shared__ double red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) misfit[b] = 0.5 * red[0];
}

template <typename T>
__global__ void fill_kernel(T* p, T v, long n) {
    for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) p[k] = v;
}

// grad_c2 = scale * G ; grad_c = 2 sqrt(c2) grad_c2   (G has stride G_bs, outputs stride N)
template <typename T>
__global__ void grad_finalize_kernel(const double* G, long G_bs, const T* c2, const int* model_of, long c2_bs,
                                     long N, double scale, double* grad_c, double* grad_c2) {
    const long b = blockIdx.y;
    const long m = model_of ? model_of[b] : b;
    for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < N; k += (long)gridDim.x * blockDim.x) {
        const double g2 = scale * G[b * G_bs + k];
        if (grad_c2) grad_c2[b * N + k] = g2;
        if (grad_c) grad_c[b * N + k] = 2.0 * sqrt((double)c2[m * c2_bs + k]) * g2;
    }
}

// c[m, jz, ix] = base + sum_p sum_q Bz[jz, p] Bx[ix, q] coef[m, p, q]
// using the <= 4 non-zero cubic basis values per node and axis.
template <typename T>
__global__ void bspline_map_kernel(const double* coef, int nbz, int nbx, const double* base,
                                   const int* bz0, const double* bzw, const int* bx0, const double* bxw,
                                   int NX, int NZ, double* c, T* c2) {
    const long m = blockIdx.y;
    const long N = (long)NX * NZ;
    for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < N; k += (long)gridDim.x * blockDim.x) {
        const int jz = (int)(k / NX), ix = (int)(k % NX);
        const int p0 = bz0[jz], q0 = bx0[ix];
        const double* C = coef + m * nbz * nbx;
        double v = base ? base[k] : 0.0;
        for (int pp = 0; pp < 4; ++pp) {
            const int p = p0 + pp;
            if (p >= nbz) break;
            double row = 0.0;
            for (int qq = 0; qq < 4; ++qq) {
                const int q = q0 + qq;
                if (q >= nbx) break;
                row += bxw[ix * 4 + qq] * C[p * nbx + q];
            }
            v += bzw[jz * 4 + pp] * row;
        }
        if (c) c[m * N + k] = v;
        c2[m * N + k] = (T)(v * v);
    }
}

// stage 1 of g_coef = Bz^T g Bx:  t[m, jz, q] = sum_ix g[m, jz, ix] Bx[ix, q]
__global__ void bspline_proj_x_kernel(const double* g, int NX, int NZ, int nbx, const int* qlo, const int* qhi,
                                      const int* bx0, const double* bxw, double* tmp) {
    const long m = blockIdx.y;
    const long total = (long)NZ * nbx;
    for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < total; k += (long)gridDim.x * blockDim.x) {
        const int jz = (int)(k / nbx), q = (int)(k % nbx);
        const double* gr = g + m * (long)NX * NZ + (long)jz * NX;
        double s = 0.0;
        for (int ix = qlo[q]; ix < qhi[q]; ++ix) {
            const int o = q - bx0[ix];
            if (o >= 0 && o < 4) s += gr[ix] * bxw[ix * 4 + o];
        }
        tmp[m * total + k] = s;
    }
}

// stage 2: g_coef[m, p, q] = sum_jz Bz[jz, p] t[m, jz, q]
__global__ void bspline_proj_z_kernel(const double* tmp, int NZ, int nbz, int nbx, const int* plo, const int* phi,
                                      const int* bz0, const double* bzw, double* gcoef) {
    const long m = blockIdx.y;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nbz * nbx) return;
    const int p = k / nbx, q = k % nbx;
    double s = 0.0;
    for (int jz = plo[p]; jz < phi[p]; ++jz) {
        const int o = p - bz0[jz];
        if (o >= 0 && o < 4) s += bzw[jz * 4 + o] * tmp[m * (long)NZ * nbx + (long)jz * nbx + q];
    }
    gcoef[m * nbz * nbx + k] = s;
}

}  // namespace sem
