// HBM streaming ceiling of the GPU: the bandwidth plain stream kernels reach
// with the forward step's access mix, the practical target for the step
// kernels (the vendor peak, 8 TB/s on MI355X, is not reachable by any kernel).
//
//   copy        1 read + 1 write stream, 16-byte accesses
//   copy_nt     the same with non-temporal loads and stores
//   read        1 read stream
//   mix3r2w_nt  3 read + 2 write streams, two of them non-temporal: the
//               forward step's streams (u, w, c^2 in; u', w out)
//   mix3r2w     the same without non-temporal hints
//
// Buffers of 1 GiB each (4 allocated), n4 = 2^26 16-byte elements per buffer,
// one thread per element, grid = n4 / 256 exactly; every access is in bounds.
//
// Then a working-set sweep of the two mix kernels on prefixes of the same
// buffers (2^20 .. 2^26 16-byte elements each, 16 MiB .. 1 GiB, four buffers
// touched per launch), with u and u' swapped between launches as the step
// kernels do: the bandwidth plain and non-temporal streams reach when the
// working set fits the 256 MB last-level cache and when it does not.
#include <hip/hip_runtime.h>

#include <cstdio>

typedef float vf4 __attribute__((ext_vector_type(4)));

#define CK(x)                                                                         \
    do {                                                                              \
        hipError_t e_ = (x);                                                          \
        if (e_ != hipSuccess) {                                                       \
            std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
            return 1;                                                                 \
        }                                                                             \
    } while (0)

__global__ void copy_k(const vf4* __restrict__ a, vf4* __restrict__ b, long n4) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n4) b[i] = a[i];
}
__global__ void copy_nt_k(const vf4* __restrict__ a, vf4* __restrict__ b, long n4) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n4) __builtin_nontemporal_store(__builtin_nontemporal_load(a + i), b + i);
}
__global__ void read_k(const vf4* __restrict__ a, float* flag, long n4) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n4) {
        const vf4 v = a[i];
        if (v.x + v.y + v.z + v.w == 12345.0f) flag[0] = 1.0f;  // never true: inputs are zero
    }
}
__global__ void mix_nt_k(const vf4* __restrict__ u, vf4* __restrict__ w, const vf4* __restrict__ c,
                         vf4* __restrict__ u2, long n4) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n4) {
        const vf4 uu = u[i], ww = __builtin_nontemporal_load(w + i), cc = c[i];
        const vf4 wn = 0.99f * (ww + cc * uu);
        __builtin_nontemporal_store(wn, w + i);
        __builtin_nontemporal_store(0.98f * uu + wn, u2 + i);
    }
}
__global__ void mix_k(const vf4* __restrict__ u, vf4* __restrict__ w, const vf4* __restrict__ c,
                      vf4* __restrict__ u2, long n4) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n4) {
        const vf4 uu = u[i], ww = w[i], cc = c[i];
        const vf4 wn = 0.99f * (ww + cc * uu);
        w[i] = wn;
        u2[i] = 0.98f * uu + wn;
    }
}

int main() {
    const long n = 1L << 28;  // floats per buffer (1 GiB)
    const long n4 = n / 4;
    float *a, *b, *c, *d, *flag;
    CK(hipMalloc(&a, n * 4));
    CK(hipMalloc(&b, n * 4));
    CK(hipMalloc(&c, n * 4));
    CK(hipMalloc(&d, n * 4));
    CK(hipMalloc(&flag, 64));
    CK(hipMemset(a, 0, n * 4));
    CK(hipMemset(b, 0, n * 4));
    CK(hipMemset(c, 0, n * 4));
    CK(hipMemset(d, 0, n * 4));
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    const int bs = 256;
    const unsigned grid = (unsigned)(n4 / bs);
    const vf4 *A = (const vf4*)a, *C = (const vf4*)c;
    vf4 *B = (vf4*)b, *D = (vf4*)d;
    auto run = [&](const char* name, double bytes, int which) -> int {
        for (int r = -3; r < 10; ++r) {
            if (r == 0) CK(hipEventRecord(e0));
            switch (which) {
                case 0: copy_k<<<grid, bs>>>(A, B, n4); break;
                case 1: copy_nt_k<<<grid, bs>>>(A, B, n4); break;
                case 2: read_k<<<grid, bs>>>(A, flag, n4); break;
                case 3: mix_nt_k<<<grid, bs>>>(A, B, C, D, n4); break;
                default: mix_k<<<grid, bs>>>(A, B, C, D, n4); break;
            }
            CK(hipGetLastError());
        }
        CK(hipEventRecord(e1));
        CK(hipEventSynchronize(e1));
        float ms = 0.0f;
        CK(hipEventElapsedTime(&ms, e0, e1));
        const double us = ms * 1e3 / 10.0;
        std::printf("%-12s %8.1f us  %7.0f GB/s\n", name, us, bytes / (us * 1e-6) / 1e9);
        std::fflush(stdout);
        return 0;
    };
    const double B1 = (double)n * 4.0;
    if (run("copy", 2 * B1, 0) || run("copy_nt", 2 * B1, 1) || run("read", B1, 2) ||
        run("mix3r2w_nt", 5 * B1, 3) || run("mix3r2w", 5 * B1, 4))
        return 1;
    // working-set sweep: m4 16-byte elements per buffer, m4 a multiple of 256
    std::printf("%-12s %10s %10s %10s %8s\n", "sweep", "MiB/buf", "set MiB", "us", "GB/s");
    for (long m4 = 1L << 20; m4 <= n4; m4 <<= 1) {
        const unsigned g = (unsigned)(m4 / bs);
        const int reps = m4 <= (1L << 22) ? 200 : 20;
        for (int nt = 0; nt < 2; ++nt) {
            for (int r = -3; r < reps; ++r) {
                if (r == 0) CK(hipEventRecord(e0));
                vf4* U = (r & 1) ? D : (vf4*)a;
                vf4* U2 = (r & 1) ? (vf4*)a : D;
                if (nt) mix_nt_k<<<g, bs>>>(U, B, C, U2, m4);
                else mix_k<<<g, bs>>>(U, B, C, U2, m4);
                CK(hipGetLastError());
            }
            CK(hipEventRecord(e1));
            CK(hipEventSynchronize(e1));
            float ms = 0.0f;
            CK(hipEventElapsedTime(&ms, e0, e1));
            const double us = ms * 1e3 / reps;
            const double mib = (double)m4 * 16.0 / (1 << 20);
            std::printf("%-12s %10.0f %10.0f %10.2f %8.0f\n", nt ? "mix3r2w_nt" : "mix3r2w", mib, 4.0 * mib, us,
                        5.0 * m4 * 16.0 / (us * 1e-6) / 1e9);
            std::fflush(stdout);
        }
    }
    CK(hipDeviceSynchronize());
    std::printf("done\n");
    return 0;
}
