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
// Two kernels per adjoint step: cells (gradients, stress of r, hourglass
// modes of r, material sums) and nodes (gather, update).
#include "sord_dev.h"
#include "sord_launch.h"

namespace sord {


// corner k (bits: dj, dk, dl) of a cell: e_a = +1 at the upper corner along a
__device__ __forceinline__ int esgn(int kq, int a) { return (kq >> a & 1) ? 1 : -1; }
// hourglass patterns (modes s_k s_l, s_j s_l, s_j s_k, s_j s_k s_l; s = +1 at the lower corner)
__device__ __forceinline__ int pq(int q, int kq) {
    const int sj = (kq & 1) ? -1 : 1, sk = (kq & 2) ? -1 : 1, sl = (kq & 4) ? -1 : 1;
    return q == 0 ? sk * sl : q == 1 ? sj * sl : q == 2 ? sj * sk : sj * sk * sl;
}

template <class T>
__global__ __launch_bounds__(256) void k_adj_cell(AdjP<T> A) {
    const int j = A.lo[0] + blockIdx.x * blockDim.x + threadIdx.x;
    const int k = A.lo[1] + blockIdx.y * blockDim.y + threadIdx.y;
    const int l = A.lo[2] + blockIdx.z;
    if (j > A.hi[0] || k > A.hi[1]) return;
    const Grid<T> &g = A.g;
    T r[8][3], w[8][3], F[8][3];
#pragma unroll
    for (int kq = 0; kq < 8; kq++) {
        const long o = g.at(j + (kq & 1), k + (kq >> 1 & 1), l + (kq >> 2 & 1));
        const T m = A.mr[o], ga = A.gam[o];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const T vb = A.vh[i][o] + A.dt * A.uh[i][o];
            r[kq][i] = A.dt * m * vb;
            w[kq][i] = A.u[i][o] + ga * A.v[i][o];
            F[kq][i] = A.h1 * A.u[i][o] + A.dth2 * A.v[i][o];
        }
    }
    T G[3][3], Gr[3][3];  // [a][i]: derivative along a of component i
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int i = 0; i < 3; i++) {
            T x = T(0), y = T(0);
#pragma unroll
            for (int kq = 0; kq < 8; kq++) {
                x += T(esgn(kq, a)) * w[kq][i];
                y += T(esgn(kq, a)) * r[kq][i];
            }
            G[a][i] = A.rd[a] * x;
            Gr[a][i] = A.rd[a] * y;
        }
    const long oc = g.at(j, k, l);
    const T lm = A.lam[oc], mu = A.mu[oc], yy = A.yy[oc];
    const T tr = G[0][0] + G[1][1] + G[2][2], trr = Gr[0][0] + Gr[1][1] + Gr[2][2];
    T sm = T(0);
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int i = 0; i < 3; i++) sm += Gr[a][i] * (G[a][i] + G[i][a]);
    A.gl[oc] += -A.vol * tr * trr;
    A.gmu[oc] += -A.vol * sm;
    // V * Hooke(Gr)
    const T V = A.vol;
    A.s6[0][oc] = V * (T(2) * mu * Gr[0][0] + lm * trr);
    A.s6[1][oc] = V * (T(2) * mu * Gr[1][1] + lm * trr);
    A.s6[2][oc] = V * (T(2) * mu * Gr[2][2] + lm * trr);
    A.s6[3][oc] = V * mu * (Gr[1][2] + Gr[2][1]);
    A.s6[4][oc] = V * mu * (Gr[2][0] + Gr[0][2]);
    A.s6[5][oc] = V * mu * (Gr[0][1] + Gr[1][0]);
    T sy = T(0);
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
        for (int i = 0; i < 3; i++) {
            T hr = T(0), hf = T(0);
#pragma unroll
            for (int kq = 0; kq < 8; kq++) {
                hr += T(pq(q, kq)) * r[kq][i];
                hf += T(pq(q, kq)) * F[kq][i];
            }
            sy += hr * hf;
            A.h12[3 * q + i][oc] = yy * hr;
        }
    A.gy[oc] += -sy;
}

template <class T>
__global__ __launch_bounds__(256) void k_adj_node(AdjP<T> A) {
    const int j = A.nlo[0] + blockIdx.x * blockDim.x + threadIdx.x;
    const int k = A.nlo[1] + blockIdx.y * blockDim.y + threadIdx.y;
    const int l = A.nlo[2] + blockIdx.z;
    if (j > A.nhi[0] || k > A.nhi[1]) return;
    const Grid<T> &g = A.g;
    // stress component index of (i, a): xx yy zz yz zx xy
    const int sidx[3][3] = {{0, 5, 4}, {5, 1, 3}, {4, 3, 2}};
    T q3[3] = {T(0), T(0), T(0)}, h3[3] = {T(0), T(0), T(0)};
#pragma unroll
    for (int kq = 0; kq < 8; kq++) {  // the node is corner kq of cell (j,k,l) - offsets
        const int cj = j - (kq & 1), ck = k - (kq >> 1 & 1), cl = l - (kq >> 2 & 1);
        if (cj < A.lo[0] || cj > A.hi[0] || ck < A.lo[1] || ck > A.hi[1] || cl < A.lo[2] || cl > A.hi[2]) continue;
        const long oc = g.at(cj, ck, cl);
#pragma unroll
        for (int i = 0; i < 3; i++) {
            T x = T(0);
#pragma unroll
            for (int a = 0; a < 3; a++) x += T(esgn(kq, a)) * A.rd[a] * A.s6[sidx[i][a]][oc];
            q3[i] -= x;
            T y = T(0);
#pragma unroll
            for (int q = 0; q < 4; q++) y += T(pq(q, kq)) * A.h12[3 * q + i][oc];
            h3[i] -= y;
        }
    }
    const long o = g.at(j, k, l);
    const T m = A.mr[o], ga = A.gam[o];
    T gm = T(0);
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const T vb = A.vh[i][o] + A.dt * A.uh[i][o];
        A.uo[i][o] = A.uh[i][o] + q3[i] + A.h1 * h3[i];
        A.vo[i][o] = vb + ga * q3[i] + A.dth2 * h3[i];
        gm += vb * (A.vn[i][o] - A.v[i][o]);
    }
    if (m != T(0)) A.gm[o] += gm / m;
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

template <class T>
void launch_adjoint(const AdjP<T> &A, hipStream_t s) {
    const dim3 b(64, 4);
    const dim3 gc((A.hi[0] - A.lo[0] + 64) / 64, (A.hi[1] - A.lo[1] + 4) / 4, A.hi[2] - A.lo[2] + 1);
    const dim3 gn((A.nhi[0] - A.nlo[0] + 64) / 64, (A.nhi[1] - A.nlo[1] + 4) / 4, A.nhi[2] - A.nlo[2] + 1);
    hipLaunchKernelGGL((k_adj_cell<T>), gc, b, 0, s, A);
    hipLaunchKernelGGL((k_adj_node<T>), gn, b, 0, s, A);
}
template <class T>
void launch_adj_inject(T *const vh[3], const Grid<T> &g, const int *rec, int n, const T *res, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL((k_adj_inject<T>), dim3((n + 63) / 64), dim3(64), 0, s, vh[0], vh[1], vh[2], g, rec, n, res);
}
template <class T>
void launch_rec_gather(T *const v[3], const Grid<T> &g, const int *rec, int n, T *out, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL((k_rec_gather<T>), dim3((n + 63) / 64), dim3(64), 0, s, v[0], v[1], v[2], g, rec, n, out);
}

#define INSTA(T)                                                                                       \
    template void launch_adjoint<T>(const AdjP<T> &, hipStream_t);                                     \
    template void launch_adj_inject<T>(T *const *, const Grid<T> &, const int *, int, const T *, hipStream_t); \
    template void launch_rec_gather<T>(T *const *, const Grid<T> &, const int *, int, T *, hipStream_t);
INSTA(float)
INSTA(double)

}  // namespace sord
