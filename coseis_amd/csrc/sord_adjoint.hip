// Adjoint of the SORD time step (free-surface meshes), for the gradient of a
// waveform misfit with respect to the material.
//
// The forward step (sord/time_integration.f90, sord/stress.f90,
// sord/acceleration.f90 on a rectilinear mesh, no PML) is linear in the
// state (u, v):
//     w = u + g v,  F = h1 u + dth2 v,  f = Ke w + Kh F,  v' = v + dt m f,
//     u' = u + dt v'
// with m = 1/mass, Ke = -V D^T C D (D: cell gradient of node fields, C: Hooke
// with lambda, mu; the divergence weights of the force are -V D^T) and
// Kh = -sum_q P_q^T yy P_q (hourglass modes).  Both are symmetric, so with
// (ub, vb) the adjoint of (u', v'):
//     vb' = vb + dt ub,  r = dt m vb',
//     ub <- ub + Ke r + h1 Kh r,  vb <- vb' + g Ke r + dth2 Kh r,
// and the step's material derivatives, with G = D w and Gr = D r:
//     dJ/dlambda_c += -V tr(G) tr(Gr),  dJ/dmu_c += -V (G + G^T) : Gr,
//     dJ/dyy_c += -sum_q,i (P_q r_i)(P_q F_i),  dJ/dm_n += vb'_n . (v'_n - v_n) / m_n.
// One fused kernel per adjoint step, on the z-march tiles of the forward
// interior kernel (sord_fused.hip): a block of W waves marches the planes l
// of its tile; per plane every lane computes the cell whose lower corner is
// its node from the plane sums of r, w and F (the row above through LDS, the
// next column through DPP lane shifts), and the cell's contributions to its
// corners are gathered as polynomials in the corner signs, carried from plane
// to plane.  Ke r and
// Kh r are gathered separately (they enter ub and vb with different weights).
// Cells outside the mesh contribute nothing (free surfaces); the material
// sums of a cell are accumulated by the block that owns its lower corner.
#include "sord_dev.h"
#include "sord_launch.h"

namespace sord {

// DPP lane shifts (wave_shl:1 / wave_shr:1; the lane without a source reads 0)
__device__ __forceinline__ int adj_dpp_next(int x) { return __builtin_amdgcn_mov_dpp(x, 0x130, 0xF, 0xF, true); }
__device__ __forceinline__ int adj_dpp_prev(int x) { return __builtin_amdgcn_mov_dpp(x, 0x138, 0xF, 0xF, true); }
__device__ __forceinline__ float adj_next(float x) { return __int_as_float(adj_dpp_next(__float_as_int(x))); }
__device__ __forceinline__ float adj_prev(float x) { return __int_as_float(adj_dpp_prev(__float_as_int(x))); }
__device__ __forceinline__ double adj_next(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = adj_dpp_next((int)b), hi = adj_dpp_next((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double adj_prev(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = adj_dpp_prev((int)b), hi = adj_dpp_prev((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Sums over the four nodes (j, k), (j+1, k), (j, k+1), (j+1, k+1) of one node
// plane (f_q, q = dj + 2 dk), shared by the cell below and the cell above the
// plane:  J = (f1 - f0) + (f3 - f2),  K = (f2 - f0) + (f3 - f1),
// S = f0 + f1 + f2 + f3,  X = f0 - f1 - f2 + f3.  A cell between planes lo
// and hi has the differences along j, k, l J_lo + J_hi, K_lo + K_hi,
// S_hi - S_lo and the hourglass modes (s_k s_l, s_j s_l, s_j s_k,
// s_j s_k s_l; s = +1 at the lower corner) K_hi - K_lo, J_hi - J_lo,
// X_lo + X_hi, X_lo - X_hi.
template <class T>
struct AdjSums {
    T r[3][4];  // r: J, K, S, X
    T w[3][3];  // w: J, K, S
    T F[3][3];  // F: J, K, X
};

// An array element as the array's base (scalar registers) plus a 32-bit byte
// offset, one per node and plane shared by every array: the load or store
// adds the two itself, where a 64-bit element index costs a 64-bit add per
// access (76 of the plane loop's 818 VALU instructions).  One array of the
// padded mesh stays below 4 GB (adj_setup).
template <class T>
__device__ __forceinline__ T ldo(const T *base, unsigned ob) {
    using G = const __attribute__((address_space(1))) char *;
    return *(const __attribute__((address_space(1))) T *)((G)(const char *)base + ob);
}
template <class T>
__device__ __forceinline__ void sto(T *base, unsigned ob, T v) {
    using G = __attribute__((address_space(1))) char *;
    *(__attribute__((address_space(1))) T *)((G)(char *)base + ob) = v;
}

// raw node loads of one plane: ub 0-2, vb 3-5, u 6-8, v 9-11, 1/m 12, gamma
// 13, the node's gradient sum 14.  The gradient
// sums (node and cell) are loaded with the plane's other data, one plane
// ahead, and stored with the step's term added: a load consumed at once would
// make the wave wait for every load issued before it, the next plane's
// prefetch included.
#define ADJ_NR 15

template <class T, int W>
__global__ __launch_bounds__(64 * W) void k_adj_fused(AdjP<T> A, const Work *__restrict__ work) {
    const Work wk = work[blockIdx.x];
    if (wk.lz0 > wk.lz1) return;  // uniform over the block
    __shared__ T sN[9][W][64];
    __shared__ T sC[12][W][64];
    const int lane = threadIdx.x, w = threadIdx.y;
    const int j = wk.jb - 1 + lane, k = wk.kb - 1 + w;
    const bool own = lane >= 1 && lane <= 62 && w >= 1 && w <= W - 2 && j <= wk.je && k <= wk.ke;
    const bool colreal = lane <= 62 && w <= W - 2 && j <= wk.je && k <= wk.ke && j >= A.lo[0] && j <= A.hi[0] &&
                         k >= A.lo[1] && k <= A.hi[1];
    const Grid<T> &g = A.g;
    // load coordinates: lanes past the column after the tile's last and waves
    // past the row after its last feed nothing, and load that column or row
    // again (lines already fetched) instead of the next tile's
    const int jc = min(max(j, 1), min(g.nm[0], wk.je + 1)), kc = min(max(k, 1), min(g.nm[1], wk.ke + 1));
    const int lmax = g.nm[2];
    const T dt = A.dt, h1 = A.h1, dth2 = A.dth2, V = A.vol;
    const T ax = V * A.rd[0], ay = V * A.rd[1], az = V * A.rd[2];

    constexpr unsigned TB = sizeof(T);
    auto fetch_node = [&](int l, T (&R)[ADJ_NR]) {
        const unsigned ob = (unsigned)g.at(jc, kc, min(l, lmax)) * TB;
#pragma unroll
        for (int i = 0; i < 3; i++)
            R[i] = ldo(A.uh[i], ob), R[3 + i] = ldo(A.vh[i], ob), R[6 + i] = ldo(A.u[i], ob), R[9 + i] = ldo(A.v[i], ob);
        R[12] = ldo(A.mr, ob), R[13] = ldo(A.gam, ob), R[14] = ldo((const T *)A.gm, ob);
    };
    // lambda, mu, yy and the three gradient sums of cell l
    auto fetch_cell = [&](int l, T (&M)[6]) {
        const unsigned ob = (unsigned)g.at(jc, kc, min(l, lmax)) * TB;
        M[0] = ldo(A.lam, ob), M[1] = ldo(A.mu, ob), M[2] = ldo(A.yy, ob);
        M[3] = ldo((const T *)A.gl, ob), M[4] = ldo((const T *)A.gmu, ob), M[5] = ldo((const T *)A.gy, ob);
    };
    // exchanged: r = dt m (vb + dt ub) 0-2, w = u + gamma v 3-5, F = h1 u + dth2 v 6-8
    auto node_vals = [&](const T (&R)[ADJ_NR], T (&o9)[9]) {
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const T vbp = R[3 + i] + dt * R[i];
            o9[i] = dt * R[12] * vbp;
            o9[3 + i] = R[6 + i] + R[13] * R[9 + i];
            o9[6 + i] = h1 * R[6 + i] + dth2 * R[9 + i];
        }
    };
    // the plane sums of r, w, F: the row above through LDS, the next column
    // through lane shifts (the own value's before the LDS read)
    auto exchange = [&](const T (&o9)[9], AdjSums<T> &N) {
#pragma unroll
        for (int q = 0; q < 9; q++) sN[q][w][lane] = o9[q];
        __syncthreads();
        const int wa = w + 1 < W ? w + 1 : w;
#pragma unroll
        for (int q = 0; q < 9; q++) {
            const T ab = sN[q][wa][lane], on = adj_next(o9[q]), an = adj_next(ab);
            const T e = o9[q] + ab, d = ab - o9[q], ne = on + an, nd = an - on;
            const T J = ne - e, K = d + nd, S = e + ne, X = nd - d;
            const int f = q / 3, c = q % 3;
            if (f == 0) N.r[c][0] = J, N.r[c][1] = K, N.r[c][2] = S, N.r[c][3] = X;
            else if (f == 1) N.w[c][0] = J, N.w[c][1] = K, N.w[c][2] = S;
            else N.F[c][0] = J, N.F[c][1] = K, N.F[c][2] = X;
        }
    };
    // cell (j, k, l): polynomials of Ke r (E: 1, s_j, s_k at s_l = +1; Ce at
    // s_l = -1) and of Kh r (H: s_j, s_k, s_j s_k; Ch), material sums if acc
    auto cell_at = [&](int l, const T (&M)[6], const AdjSums<T> &Nlo, const AdjSums<T> &Nhi, T (&E)[3][3],
                       T (&Ce)[3][3], T (&H)[3][3], T (&Ch)[3][3], bool acc) {
        const bool real = colreal && l >= A.lo[2] && l <= A.hi[2];
        T Gr[3][3], G[3][3];  // [a][c]: derivative along a of component c
#pragma unroll
        for (int c = 0; c < 3; c++) {
            Gr[0][c] = A.rd[0] * (Nlo.r[c][0] + Nhi.r[c][0]);
            Gr[1][c] = A.rd[1] * (Nlo.r[c][1] + Nhi.r[c][1]);
            Gr[2][c] = A.rd[2] * (Nhi.r[c][2] - Nlo.r[c][2]);
            G[0][c] = A.rd[0] * (Nlo.w[c][0] + Nhi.w[c][0]);
            G[1][c] = A.rd[1] * (Nlo.w[c][1] + Nhi.w[c][1]);
            G[2][c] = A.rd[2] * (Nhi.w[c][2] - Nlo.w[c][2]);
        }
        const T lm = real ? M[0] : T(0), m = real ? M[1] : T(0), y = real ? M[2] : T(0);
        const T trr = Gr[0][0] + Gr[1][1] + Gr[2][2];
        const T m2 = m + m;
        T s[6];
        s[0] = m2 * Gr[0][0] + lm * trr;
        s[1] = m2 * Gr[1][1] + lm * trr;
        s[2] = m2 * Gr[2][2] + lm * trr;
        s[3] = m * (Gr[2][1] + Gr[1][2]);
        s[4] = m * (Gr[0][2] + Gr[2][0]);
        s[5] = m * (Gr[1][0] + Gr[0][1]);
        const T Av[3][3] = {{ax * s[0], ay * s[5], az * s[4]},
                            {ax * s[5], ay * s[1], az * s[3]},
                            {ax * s[4], ay * s[3], az * s[2]}};
        T sy = T(0);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            // modes (s_k s_l, s_j s_l, s_j s_k, s_j s_k s_l) of r and of F
            const T hr[4] = {Nhi.r[c][1] - Nlo.r[c][1], Nhi.r[c][0] - Nlo.r[c][0], Nlo.r[c][3] + Nhi.r[c][3],
                             Nlo.r[c][3] - Nhi.r[c][3]};
            const T hf[4] = {Nhi.F[c][1] - Nlo.F[c][1], Nhi.F[c][0] - Nlo.F[c][0], Nlo.F[c][2] + Nhi.F[c][2],
                             Nlo.F[c][2] - Nhi.F[c][2]};
#pragma unroll
            for (int q = 0; q < 4; q++) sy = sy + hr[q] * hf[q];
            T h[4];
#pragma unroll
            for (int q = 0; q < 4; q++) h[q] = y * hr[q];
            E[c][0] = Av[c][2], E[c][1] = Av[c][0], E[c][2] = Av[c][1];
            Ce[c][0] = -Av[c][2], Ce[c][1] = Av[c][0], Ce[c][2] = Av[c][1];
            H[c][0] = -h[1], H[c][1] = -h[0], H[c][2] = -h[2] - h[3];
            Ch[c][0] = h[1], Ch[c][1] = h[0], Ch[c][2] = -h[2] + h[3];
        }
        if (acc && real) {
            const unsigned ob = (unsigned)g.at(j, k, l) * TB;
            const T tr = G[0][0] + G[1][1] + G[2][2];
            T sm = T(0);
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
                for (int c = 0; c < 3; c++) sm = sm + Gr[a][c] * (G[a][c] + G[c][a]);
            sto(A.gl, ob, M[3] + -V * tr * trr);
            sto(A.gmu, ob, M[4] + -V * sm);
            sto(A.gy, ob, M[5] + -sy);
        }
    };

    T o9[9];
    AdjSums<T> NA, NB;  // plane sums of two planes (lower, upper in turn)
    T EA[3][3], HA[3][3];                              // polynomials of the current cell (s_l = +1)
    T CeA[3][3], ChA[3][3], CeB[3][3], ChB[3][3];      // s_l = -1 polynomials (carried, new in turn)
    T RA[ADJ_NR], RB[ADJ_NR], mA[6], mB[6];
    T ow[12];  // own node of the plane being updated: ub, vb, v, 1/m, gamma, gradient sum
    auto keep_own = [&](const T (&R)[ADJ_NR]) {
#pragma unroll
        for (int i = 0; i < 6; i++) ow[i] = R[i];
#pragma unroll
        for (int i = 0; i < 3; i++) ow[6 + i] = R[9 + i];
        ow[9] = R[12], ow[10] = R[13], ow[11] = R[14];
    };
    fetch_node(wk.lz0 - 1, RA);
    node_vals(RA, o9);
    exchange(o9, NA);
    fetch_node(wk.lz0, RB);
    fetch_cell(wk.lz0 - 1, mB);
    node_vals(RB, o9);
    keep_own(RB);
    __syncthreads();  // every wave has read the first plane from LDS
    exchange(o9, NB);
    fetch_node(wk.lz0 + 1, RA);
    fetch_cell(wk.lz0, mA);
    cell_at(wk.lz0 - 1, mB, NA, NB, EA, CeA, HA, ChA, false);
    // every wave has read the second plane from LDS before the loop's first
    // plane writes its own slots again
    __syncthreads();

    auto plane = [&](int l, bool act, const T (&Rc)[ADJ_NR], const T (&mc)[6], T (&Rn)[ADJ_NR], T (&mn)[6],
                     const AdjSums<T> &Nlo, AdjSums<T> &Nhi, const T (&cE)[3][3], const T (&cH)[3][3], T (&nE)[3][3],
                     T (&nH)[3][3]) {
        node_vals(Rc, o9);
        fetch_node(l + 2, Rn);
        fetch_cell(l + 1, mn);
        exchange(o9, Nhi);
        cell_at(l, mc, Nlo, Nhi, EA, nE, HA, nH, own && act);
        T Re[3][2], Rh[3][2];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const T e0 = EA[c][0] + cE[c][0], ej = EA[c][1] + cE[c][1], ek = EA[c][2] + cE[c][2];
            const T hj = HA[c][0] + cH[c][0], hk = HA[c][1] + cH[c][1], hjk = HA[c][2] + cH[c][2];
            Re[c][0] = e0 + ek;
            Re[c][1] = ej;
            Rh[c][0] = hk;
            Rh[c][1] = hj + hjk;
            sC[4 * c][w][lane] = e0 - ek;
            sC[4 * c + 1][w][lane] = ej;
            sC[4 * c + 2][w][lane] = -hk;
            sC[4 * c + 3][w][lane] = hj - hjk;
        }
        __syncthreads();
        const int wb = w > 0 ? w - 1 : 0;
        T fe[3], fh[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const T r0 = Re[c][0] + sC[4 * c][wb][lane], rj = Re[c][1] + sC[4 * c + 1][wb][lane];
            fe[c] = (r0 + rj) + adj_prev(r0 - rj);
            const T q0 = Rh[c][0] + sC[4 * c + 2][wb][lane], qj = Rh[c][1] + sC[4 * c + 3][wb][lane];
            fh[c] = (q0 + qj) + adj_prev(q0 - qj);
        }
        if (own && act) {
            const unsigned ob = (unsigned)g.at(j, k, l) * TB;
            // the residual at receiver nodes of this tile (uniform loop)
            T inj[3] = {T(0), T(0), T(0)};
            for (int e = wk.pad0; e < wk.pad0 + wk.pad1; e++) {
                const int4 r = A.rtile[e];
                if (r.x == j && r.y == k && r.z == l)
#pragma unroll
                    for (int c = 0; c < 3; c++) inj[c] = inj[c] + A.res[3 * r.w + c];
            }
            // the 1/mass sum: (vb + dt ub) . (v_s+1 - v_s) over the steps,
            // summed by parts so that a step needs only its own state's v:
            // the new (vb + dt ub), which the next adjoint step takes as its
            // input, times v_s, less the input's times v_s (none of the new
            // at s = 1; the input of the first step times v_nt comes before
            // the sweep, k_adj_gm_end)
            T gm = T(0);
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const T vbp = ow[3 + c] + dt * ow[c];
                const T un = ow[c] + fe[c] + h1 * fh[c];
                const T vn = (vbp + ow[10] * fe[c] + dth2 * fh[c]) + inj[c];
                sto(A.uo[c], ob, un);
                sto(A.vo[c], ob, vn);
                const T vbo = A.gm_out ? vn + dt * un : T(0);
                gm = gm + (vbo - vbp) * ow[6 + c];
            }
            if (ow[9] != T(0)) sto(A.gm, ob, ow[11] + gm / ow[9]);
        }
        keep_own(Rc);
    };
    for (int l = wk.lz0; l <= wk.lz1; l += 2) {
        plane(l, true, RA, mA, RB, mB, NB, NA, CeA, ChA, CeB, ChB);
        plane(l + 1, l + 1 <= wk.lz1, RB, mB, RA, mA, NA, NB, CeB, ChB, CeA, ChA);
    }
}

// the 1/mass sums' end term: the first adjoint step's input (vb: the
// residual of the last step, ub zero) times v of the last forward state, at
// the distinct receiver nodes (every other node's input is zero)
template <class T>
struct Ptr3 {
    const T *p[3];
};
template <class T>
__global__ void k_adj_gm_end(T *gm, const T *mr, Ptr3<T> vh, Ptr3<T> v, Grid<T> g, const int *node, int n) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const long o = g.at(node[3 * q], node[3 * q + 1], node[3 * q + 2]);
    const T s = (vh.p[0][o] * v.p[0][o] + vh.p[1][o] * v.p[1][o]) + vh.p[2][o] * v.p[2][o];
    if (mr[o] != T(0)) gm[o] = gm[o] + s / mr[o];
}
template <class T>
void launch_adj_gm_end(T *gm, const T *mr, T *const vh[3], const T *const v[3], const Grid<T> &g, const int *node,
                       int n, hipStream_t s) {
    Ptr3<T> a, b;
    for (int i = 0; i < 3; i++) a.p[i] = vh[i], b.p[i] = v[i];
    if (n > 0) hipLaunchKernelGGL((k_adj_gm_end<T>), dim3((n + 63) / 64), dim3(64), 0, s, gm, mr, a, b, g, node, n);
}

// add w * r to the adjoint velocity at receiver nodes
template <class T>
__global__ void k_adj_inject(T *vh0, T *vh1, T *vh2, Grid<T> g, const int *rec, int n, const T *res) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const long o = g.at(rec[3 * q], rec[3 * q + 1], rec[3 * q + 2]);
    vh0[o] += res[3 * q];
    vh1[o] += res[3 * q + 1];
    vh2[o] += res[3 * q + 2];
}

// velocity at receiver nodes
template <class T>
__global__ void k_rec_gather(const T *v0, const T *v1, const T *v2, Grid<T> g, const int *rec, int n, T *out) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const long o = g.at(rec[3 * q], rec[3 * q + 1], rec[3 * q + 2]);
    out[3 * q] = v0[o], out[3 * q + 1] = v1[o], out[3 * q + 2] = v2[o];
}

// velocity at receiver nodes of many states (one thread per state and
// receiver): out[s][q][c] = v_c of states[s] at receiver q
template <class T>
__global__ void k_rec_gather_states(const T *const *states, int nstate, size_t S, Grid<T> g, const int *rec, int n,
                                    T *out) {
    const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (q >= (long)nstate * n) return;
    const int st = (int)(q / n), r = (int)(q % n);
    const long o = g.at(rec[3 * r], rec[3 * r + 1], rec[3 * r + 2]);
    const T *v = states[st] + 3 * S;
#pragma unroll
    for (int c = 0; c < 3; c++) out[3 * q + c] = v[c * S + o];
}

template <class T>
void launch_adjoint(const AdjP<T> &A, const Work *work, int nwork, hipStream_t s) {
    constexpr int W = adj_waves<T>();
    if (nwork > 0) hipLaunchKernelGGL((k_adj_fused<T, W>), dim3(nwork), dim3(64, W), 0, s, A, work);
}
template <class T>
void launch_adj_inject(T *const vh[3], const Grid<T> &g, const int *rec, int n, const T *res, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL((k_adj_inject<T>), dim3((n + 63) / 64), dim3(64), 0, s, vh[0], vh[1], vh[2], g, rec, n, res);
}
template <class T>
void launch_rec_gather_states(const T *const *states, int nstate, size_t S, const Grid<T> &g, const int *rec, int n,
                              T *out, hipStream_t s) {
    const long m = (long)nstate * n;
    if (m > 0)
        hipLaunchKernelGGL((k_rec_gather_states<T>), dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, states, nstate,
                           S, g, rec, n, out);
}
template <class T>
void launch_rec_gather(T *const v[3], const Grid<T> &g, const int *rec, int n, T *out, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL((k_rec_gather<T>), dim3((n + 63) / 64), dim3(64), 0, s, v[0], v[1], v[2], g, rec, n, out);
}

// dJ/drho, dJ/dvp, dJ/dvs of every real cell from the material sums: mu =
// rho vs^2, lambda = rho vp^2 - 2 mu, yy through lambda and mu, 1/m_n = 8 /
// (sum of the 8 cells' rho V) (one thread per cell, the host restatement's
// order of operations)
template <class T>
__global__ void k_adj_chain(AdjChain<T> c) {
    const int n0 = c.hi[0] - c.lo[0] + 1, n1 = c.hi[1] - c.lo[1] + 1, n2 = c.hi[2] - c.lo[2] + 1;
    const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (q >= (long)n0 * n1 * n2) return;
    const int j = c.lo[0] + (int)(q % n0), k = c.lo[1] + (int)((q / n0) % n1), l = c.lo[2] + (int)(q / ((long)n0 * n1));
    const long o = c.g.at(j, k, l);
    const T rho = c.rho[o], vp = c.vp[o], vs = c.vs[o];
    const T L = c.lam[o], M = c.mu[o], D = L + T(2) * M;
    const T dyl = D == 0 ? T(0) : c.sy * M * M / (D * D);
    const T dym = D == 0 ? T(0) : c.sy * (L * L + T(2) * L * M + T(2) * M * M) / (D * D);
    const T jl = c.gl[o] + c.gy[o] * dyl, jm = c.gmu[o] + c.gy[o] * dym;
    T jr = T(0);  // through the node masses of the 8 corners
    for (int kq = 0; kq < 8; kq++) {
        const long on = c.g.at(j + (kq & 1), k + (kq >> 1 & 1), l + (kq >> 2 & 1));
        const T m = c.mr[on];
        jr += c.gm[on] * (-m * m * c.vol / T(8));
    }
    c.grho[q] = jr + jl * (vp * vp - T(2) * vs * vs) + jm * vs * vs;
    c.gvp[q] = jl * T(2) * rho * vp;
    c.gvs[q] = jl * (-T(4) * rho * vs) + jm * T(2) * rho * vs;
}
template <class T>
void launch_adj_chain(const AdjChain<T> &c, hipStream_t s) {
    const long n = (long)(c.hi[0] - c.lo[0] + 1) * (c.hi[1] - c.lo[1] + 1) * (c.hi[2] - c.lo[2] + 1);
    if (n > 0) hipLaunchKernelGGL((k_adj_chain<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, c);
}

#define INSTA(T)                                                                                       \
    template void launch_adj_chain<T>(const AdjChain<T> &, hipStream_t);                              \
    template void launch_adjoint<T>(const AdjP<T> &, const Work *, int, hipStream_t);                 \
    template void launch_adj_inject<T>(T *const *, const Grid<T> &, const int *, int, const T *, hipStream_t); \
    template void launch_adj_gm_end<T>(T *, const T *, T *const *, const T *const *, const Grid<T> &, const int *, int, \
                                       hipStream_t);                                                   \
    template void launch_rec_gather<T>(T *const *, const Grid<T> &, const int *, int, T *, hipStream_t);         \
    template void launch_rec_gather_states<T>(const T *const *, int, size_t, const Grid<T> &, const int *, int, T *, \
                                              hipStream_t);
INSTA(float)
INSTA(double)

}  // namespace sord
