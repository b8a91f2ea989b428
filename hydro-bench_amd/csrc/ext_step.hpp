// ext_step.hpp -- one slot of CalculateFluxKernel and one cell of
// UpdateCellKernel on the caller's (reference) layout, callable from the
// device kernels and from host code (cuda_native_impl/hydro-cal-src/src/
// functors.cu F:47-202 and F:717-784).
#pragma once

#include "kernels.hpp"
#include "physics.hpp"

namespace hydro {

template <class R>
__host__ __device__ __forceinline__ BcCtx<R> make_bc(const ExtView<R> &m, int cell, int slot, int kt, int jt) {
    BcCtx<R> bc;
    bc.QT = m.QT; bc.DQT = m.DQT; bc.ZT = m.ZT; bc.DZT = m.DZT;
    bc.ZW = m.ZW; bc.QW = m.QW; bc.BF = m.BF; bc.NHQ1 = m.NHQ1;
    bc.CELL = m.CELL; bc.NHQ = m.NHQ; bc.cell = cell; bc.kt = kt; bc.jt = jt;
    bc.SIDE = m.SIDE[slot]; bc.COSF = m.COSF[slot]; bc.SINF = m.SINF[slot];
    return bc;
}

template <class R> __host__ __device__ __forceinline__ BcCtx<R> empty_bc() {
    BcCtx<R> bc;
    bc.QT = bc.DQT = bc.ZT = bc.DZT = bc.ZW = bc.QW = bc.BF = nullptr;
    bc.NHQ1 = nullptr;
    bc.CELL = bc.NHQ = bc.cell = bc.kt = bc.jt = 0;
    bc.SIDE = bc.COSF = bc.SINF = 0;
    return bc;
}

// Flux of slot idx = 4*pos + j into FLUX0..3 (and KLAS 7 -> 0 when a breach opens).
template <class R, int VAR>
__host__ __device__ inline void ext_flux_slot(const ExtView<R> &m, int idx, int kt, int jt) {
    const int pos = idx >> 2;
    const R H1 = m.H[pos], U1 = m.U[pos], V1 = m.V[pos], Z1 = m.Z[pos];
    const R ZB1 = m.ZB1[pos], ZBp = m.ZBC[pos];
    const int NC = m.NAC[idx] - 1;
    const int klas = m.KLAS[idx];
    const int kind = kind_of_klas(klas, pos < NC);
    const bool has = (NC != -1);
    R Hn = 0, Un = 0, Vn = 0, Zn = 0, ZBn = 0;
    if (has) { Hn = m.H[NC]; Un = m.U[NC]; Vn = m.V[NC]; Zn = m.Z[NC]; ZBn = m.ZBC[NC]; }
    Flux<R> prev = {0, 0, 0, 0};
    BcCtx<R> bc = empty_bc<R>();
    if (kind >= K_B10) {
        bc = make_bc<R>(m, pos, idx, kt, jt);
        if (kind == K_BUNK) prev = {m.F0[idx], m.F1[idx], m.F2[idx], m.F3[idx]};
    }
    bool fired = false;
    const Flux<R> f = slot_flux<R, VAR>(kind, m.S, bc, H1, U1, V1, Z1, ZB1, ZBp, has, Hn, Un, Vn,
                                        Zn, ZBn, m.COSF[idx], m.SINF[idx], prev, &fired);
    if (fired) {
        m.KLAS[idx] = 0;
        if (m.KLASr) m.KLASr[idx] = 0;
    }
    m.F0[idx] = f.f0;
    m.F1[idx] = f.f1;
    m.F2[idx] = f.f2;
    m.F3[idx] = f.f3;
}

// Update of cell pos from its four slot fluxes.
template <class R, int VAR>
__host__ __device__ inline void ext_update_cell(const ExtView<R> &m, int pos) {
    R WH = 0.0, WU = 0.0, WV = 0.0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int idx = 4 * pos + j;
        const Flux<R> f = {m.F0[idx], m.F1[idx], m.F2[idx], m.F3[idx]};
        accumulate<R>(WH, WU, WV, f, m.SIDE[idx], m.SLCOS[idx], m.SLSIN[idx]);
    }
    const int b = 4 * pos;
    const CellOut<R> o = update_math<R, VAR>(m.S, m.H[pos], m.U[pos], m.V[pos], m.ZBC[pos], m.AREA[pos],
                                             m.FNC[pos], WH, WU, WV, m.SIDE[b], m.SIDE[b + 1],
                                             m.SIDE[b + 2], m.SIDE[b + 3]);
    m.H[pos] = o.H;
    m.U[pos] = o.U;
    m.V[pos] = o.V;
    m.Z[pos] = o.Z;
    m.W[pos] = o.W;
}

}  // namespace hydro
