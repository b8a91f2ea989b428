// diag.hip -- conservation and health diagnostics of a mesh's state, reduced
// on the device: the quantities the reference's alignment checker computes
// on the host (scripts/check_correctness.py:620-655 _conservation_metrics:
// mass, momenta, kinetic and potential energy, KLAS 10 and KLAS 1 inflow as
// sums of FLUX0 * SIDE; :980-990: non-finite and infinite depths, depth
// range).
//
// Pass 1 (k_diag_partial): workgroup w takes cells [w*kCellsPerBlock, ...)
// and their slots; each thread accumulates its cells and slots, the wave
// reduces with DPP (wave_reduce.hpp), the workgroup's waves combine in LDS
// in wave order, and one record per workgroup goes to `part`.
// Pass 2 (k_diag_final): one workgroup reduces the records in a fixed order
// and writes the rounded results.  Both passes are deterministic.
//
// Every term is formed in double from the stored values, in the checker's
// order of operations (H * U * a = (H * U) * a, ...), so the result equals
// the correctly rounded sum of the checker's terms (double-double sums).
#include <mutex>
#include <vector>

#include "diag.hpp"
#include "wave_reduce.hpp"

namespace hydro {

// Fixed partial-sum buffer of the allocation-free entry point: one record
// per 1024 cells for meshes of up to kDiagScratchRecords * 1024 cells.  Each
// GPU holds its own copy until the process ends.  Host code never passes
// the array's name to a kernel (on the host it is only a stand-in for the
// symbol); it passes the device address diag_scratch() looked up.
static __device__ __align__(256) double scratch[kDiagScratchRecords * kDiagRecord];

namespace {

constexpr int kDiagBlock = 256;
constexpr int kWaves = kDiagBlock / 64;

struct Acc {
    DD s[kDiagSums];
    uint32_t nonfinite, inf;
    double hmin, hmax;
};

__device__ __forceinline__ void acc_init(Acc &a) {
#pragma unroll
    for (int q = 0; q < kDiagSums; q++) a.s[q] = DD{0.0, 0.0};
    a.nonfinite = a.inf = 0;
    a.hmin = __builtin_huge_val();
    a.hmax = -__builtin_huge_val();
}

// Wave reduction, then the waves of the workgroup in order through LDS;
// thread 0 returns the workgroup's totals.
__device__ __forceinline__ void block_reduce(Acc &a, double *lds) {
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int q = 0; q < kDiagSums; q++) a.s[q] = wave_sum(a.s[q]);
    a.nonfinite = wave_sum(a.nonfinite);
    a.inf = wave_sum(a.inf);
    a.hmin = wave_min(a.hmin);
    a.hmax = wave_max(a.hmax);
    if (lane == 0) {
        double *r = lds + w * kDiagRecord;
#pragma unroll
        for (int q = 0; q < kDiagSums; q++) {
            r[2 * q] = a.s[q].hi;
            r[2 * q + 1] = a.s[q].lo;
        }
        r[2 * kDiagSums] = (double)a.nonfinite;
        r[2 * kDiagSums + 1] = (double)a.inf;
        r[2 * kDiagSums + 2] = a.hmin;
        r[2 * kDiagSums + 3] = a.hmax;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int v = 1; v < kWaves; v++) {
            const double *r = lds + v * kDiagRecord;
            for (int q = 0; q < kDiagSums; q++) a.s[q] = dd_add(a.s[q], DD{r[2 * q], r[2 * q + 1]});
            a.nonfinite += (uint32_t)r[2 * kDiagSums];
            a.inf += (uint32_t)r[2 * kDiagSums + 1];
            a.hmin = fmin(a.hmin, r[2 * kDiagSums + 2]);
            a.hmax = fmax(a.hmax, r[2 * kDiagSums + 3]);
        }
    }
}

template <class R> __global__ void __launch_bounds__(kDiagBlock) k_diag_partial(ExtView<R> m, double *part) {
    __shared__ double lds[kWaves * kDiagRecord];
    Acc a;
    acc_init(a);
    const int c0 = blockIdx.x * kDiagCellsPerBlock;
    const int c1 = min(m.CELL, c0 + kDiagCellsPerBlock);
    for (int c = c0 + (int)threadIdx.x; c < c1; c += kDiagBlock) {
        const double H = (double)m.H[c], U = (double)m.U[c], V = (double)m.V[c], A = (double)m.AREA[c];
        a.s[0] = dd_add(a.s[0], H * A);
        a.s[1] = dd_add(a.s[1], H * U * A);
        a.s[2] = dd_add(a.s[2], H * V * A);
        a.s[3] = dd_add(a.s[3], 0.5 * H * (U * U + V * V) * A);
        a.s[4] = dd_add(a.s[4], 0.5 * 9.81 * H * H * A);
        const bool fin = __builtin_isfinite(H);
        a.nonfinite += fin ? 0u : 1u;
        a.inf += __builtin_isinf(H) ? 1u : 0u;
        if (fin) {
            a.hmin = fmin(a.hmin, H);
            a.hmax = fmax(a.hmax, H);
        }
    }
    for (int s = 4 * c0 + (int)threadIdx.x; s < 4 * c1; s += kDiagBlock) {
        const int k = m.KLAS[s];
        if (k == 10 || k == 1) {
            const double t = (double)m.F0[s] * (double)m.SIDE[s];
            if (k == 10) a.s[5] = dd_add(a.s[5], t);
            else a.s[6] = dd_add(a.s[6], t);
        }
    }
    block_reduce(a, lds);
    if (threadIdx.x == 0) {
        double *r = part + (size_t)blockIdx.x * kDiagRecord;
        for (int q = 0; q < kDiagSums; q++) {
            r[2 * q] = a.s[q].hi;
            r[2 * q + 1] = a.s[q].lo;
        }
        r[2 * kDiagSums] = (double)a.nonfinite;
        r[2 * kDiagSums + 1] = (double)a.inf;
        r[2 * kDiagSums + 2] = a.hmin;
        r[2 * kDiagSums + 3] = a.hmax;
    }
}

__global__ void __launch_bounds__(kDiagBlock) k_diag_final(const double *part, int nrec, double *out) {
    __shared__ double lds[kWaves * kDiagRecord];
    Acc a;
    acc_init(a);
    for (int i = threadIdx.x; i < nrec; i += kDiagBlock) {
        const double *r = part + (size_t)i * kDiagRecord;
        for (int q = 0; q < kDiagSums; q++) a.s[q] = dd_add(a.s[q], DD{r[2 * q], r[2 * q + 1]});
        a.nonfinite += (uint32_t)r[2 * kDiagSums];
        a.inf += (uint32_t)r[2 * kDiagSums + 1];
        a.hmin = fmin(a.hmin, r[2 * kDiagSums + 2]);
        a.hmax = fmax(a.hmax, r[2 * kDiagSums + 3]);
    }
    block_reduce(a, lds);
    if (threadIdx.x == 0) {
        for (int q = 0; q < kDiagSums; q++) out[q] = a.s[q].hi + a.s[q].lo;
        out[kDiagSums] = (double)a.nonfinite;
        out[kDiagSums + 1] = (double)a.inf;
        // no finite depth at all: the range is reported as NaN
        out[kDiagSums + 2] = a.hmin <= a.hmax ? a.hmin : __builtin_nan("");
        out[kDiagSums + 3] = a.hmin <= a.hmax ? a.hmax : __builtin_nan("");
    }
}

}  // namespace

int diag_records(int C) { return C <= 0 ? 0 : (C + kDiagCellsPerBlock - 1) / kDiagCellsPerBlock; }

template <class R> hipError_t launch_diagnostics(const ExtView<R> &m, double *part, double *out, hipStream_t s) {
    const int nrec = diag_records(m.CELL);
    if (nrec > 0) {
        hipLaunchKernelGGL((k_diag_partial<R>), dim3(nrec), dim3(kDiagBlock), 0, s, m, part);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(kDiagBlock), 0, s, part, nrec, out);
    return hipGetLastError();
}

hipError_t diag_scratch(double **out) {
    static std::mutex mu;
    static std::vector<double *> addr;   // per device, looked up once with that device current
    int dev = 0, ndev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lock(mu);
    if (addr.empty()) {
        if ((e = hipGetDeviceCount(&ndev)) != hipSuccess) return e;
        addr.assign(ndev > 0 ? ndev : 1, nullptr);
    }
    if (dev < 0 || dev >= (int)addr.size()) return hipErrorInvalidDevice;
    if (!addr[dev]) {
        void *p = nullptr;
        if ((e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch))) != hipSuccess) return e;
        addr[dev] = static_cast<double *>(p);
    }
    *out = addr[dev];
    return hipSuccess;
}

template hipError_t launch_diagnostics<double>(const ExtView<double> &, double *, double *, hipStream_t);
template hipError_t launch_diagnostics<float>(const ExtView<float> &, double *, double *, hipStream_t);

}  // namespace hydro
