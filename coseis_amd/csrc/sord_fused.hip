// Fused single-pass SORD step for interior nodes (CDNA4, gfx950).
//
// One launch advances every node of a box whose 8 surrounding cells are
// elastic (no PML memory, no fault, no boundary condition): strain
// (sord/stress.f90:23-26 with sord/diff_nc_op.f90), Hooke's law and 1/V
// (sord/stress.f90:192-238), divergence (sord/acceleration.f90:284-293 with
// sord/diff_cn_op.f90), hourglass stiffness + viscosity
// (sord/acceleration.f90:390-421 with sord/hourglass_op.f90), Newton
// (sord/acceleration.f90:498-509) and the leapfrog of the next step
// (sord/time_integration.f90:26-41), on a mesh with uniform spacing per axis.
//
// Work decomposition ("z-march"): a block is W wavefronts; wave w owns node
// row k = kb - 1 + w, lane i owns node column j = jb - 1 + i, and the block
// marches the planes l of a chunk.  Lanes 0 and 63 and waves 0 and W-1 are a
// one-node halo: they load and compute cells but update no node.  Per plane,
// every lane computes the one cell whose lower corner is its node: its stress
// and its four hourglass modes.  The cell's contribution to its 8 corner nodes
// is a polynomial in the corner signs (s_j, s_k, s_l) = +-1:
//     A_j s_j + A_k s_k + A_l s_l + B_jk s_j s_k + B_kl s_k s_l + B_jl s_j s_l + B_jkl s_j s_k s_l
// (A: divergence of the stress row, B: minus the hourglass modes), so the
// gather over the 8 cells of a node factorises: the plane l-1 part is carried
// in registers, the row k-1 part comes through LDS, the column j-1 part
// through a lane shuffle.  u, v, gamma, 1/m and lambda, mu, yy are read once
// and u, v written once per node: 68 bytes per node-update in FP32.
#include "sord_dev.h"
#include "sord_launch.h"

namespace sord {

template <class T>
__device__ __forceinline__ void atomic_max_bits(void *slot, T v);
template <>
__device__ __forceinline__ void atomic_max_bits<float>(void *slot, float v) {
    // v >= 0 or NaN: the ordering of the magnitude bits is the ordering of the
    // values, and NaN sorts above +inf
    atomicMax((unsigned int *)slot, __float_as_uint(v) & 0x7fffffffu);
}
template <>
__device__ __forceinline__ void atomic_max_bits<double>(void *slot, double v) {
    unsigned long long b = (unsigned long long)__double_as_longlong(v) & 0x7fffffffffffffffull;
    atomicMax((unsigned long long *)slot, b);
}

template <class T>
__device__ __forceinline__ T wave_max(T v) {
    for (int o = 32; o > 0; o >>= 1) {
        T w = __shfl_xor(v, o);
        v = (w > v || w != w) ? w : v;
    }
    return v;
}

template <class T>
struct FusedP {
    T *u[3], *v[3], *un[3], *vn[3];
    T *gam, *mr, *lam, *mu, *yy;
    Grid<T> g;
    int lo[3], hi[3];   // node box (1-based local, inclusive)
    int nz;             // planes per block
    int nbx, nby, nbz;  // blocks per axis
    T gx, gy, gz;       // 1 / (4 dx_a): gradient from the sum of four edge differences
    T ax, ay, az;       // dx_b dx_c / 4: divergence weights
    T h1, dth2, dt;
    T vol, rvol;        // cell volume and its inverse (source ops)
    void *stat;         // 4 max slots (a, v, u, w squared norms) or null
    int stat_now, stat_next;
};

template <class T, int W>
__global__ __launch_bounds__(64 * W) void k_fused(FusedP<T> Q, PointOps<T> ops) {
    __shared__ T sN[6][W][64];
    __shared__ T sC[6][W][64];
    const int lane = threadIdx.x, w = threadIdx.y;
    // block -> tile: consecutive tiles along k share halo rows, so map tiles
    // with the same hardware XCD (round-robin of the linear block id over 8
    // XCDs) to neighbouring k positions
    const int nblk = Q.nbx * Q.nby * Q.nbz;
    const int bid = blockIdx.x;
    int t = bid;
    if (nblk % 8 == 0) t = (bid % 8) * (nblk / 8) + bid / 8;
    const int by = t % Q.nby, bx = (t / Q.nby) % Q.nbx, bz = t / (Q.nby * Q.nbx);
    const int jb = Q.lo[0] + bx * 62, kb = Q.lo[1] + by * (W - 2);
    const int lz0 = Q.lo[2] + bz * Q.nz;
    if (lz0 > Q.hi[2]) return;  // uniform over the block
    const int lz1 = min(lz0 + Q.nz - 1, Q.hi[2]);
    const int j = jb - 1 + lane, k = kb - 1 + w;
    const bool own = lane >= 1 && lane <= 62 && w >= 1 && w <= W - 2 && j <= Q.hi[0] && k <= Q.hi[1];
    const bool cellv = lane <= 62 && w <= W - 2 && j <= Q.hi[0] && k <= Q.hi[1];
    const Grid<T> &g = Q.g;
    const int jc = min(max(j, 1), g.nm[0]), kc = min(max(k, 1), g.nm[1]);
    const bool stat_now = Q.stat && Q.stat_now, stat_next = Q.stat && Q.stat_next;
    T st_a = T(0), st_v = T(0), st_u = T(0), st_w = T(0);

    // node data: W = u + gamma v (strain), F = h1 u + dt h2 v (hourglass)
    T Nlo[4][6], Nhi[4][6];
    T uo[3], vo[3], un1[3], vn1[3];

    auto load_node = [&](int l, T own6[6], T uu[3], T vv[3]) {
        const long o = g.at(jc, kc, l);
        const T ga = Q.gam[o];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            uu[i] = Q.u[i][o];
            vv[i] = Q.v[i][o];
            own6[i] = uu[i] + ga * vv[i];
            own6[3 + i] = Q.h1 * uu[i] + Q.dth2 * vv[i];
        }
    };
    auto exchange = [&](const T own6[6], T N[4][6]) {
#pragma unroll
        for (int q = 0; q < 6; q++) sN[q][w][lane] = own6[q];
        __syncthreads();
        const int wa = w + 1 < W ? w + 1 : w;
#pragma unroll
        for (int q = 0; q < 6; q++) {
            const T ab = sN[q][wa][lane];
            N[0][q] = own6[q];
            N[2][q] = ab;
            N[1][q] = __shfl_down(own6[q], 1);
            N[3][q] = __shfl_down(ab, 1);
        }
    };
    // stress and hourglass of cell (j, k, l) from corner planes Nlo (l) and Nhi (l+1);
    // E: polynomial evaluated at s_l = +1, C: at s_l = -1
    auto cell = [&](int l, T E[3][4], T C[3][4]) {
        const long oc = g.at(jc, kc, l);
        const T lm = Q.lam[oc], m = Q.mu[oc], y = Q.yy[oc];
        T G[3][3];  // G[a][c]: gradient along a of component c
#pragma unroll
        for (int c = 0; c < 3; c++) {
            G[0][c] = Q.gx * (((Nlo[1][c] - Nlo[0][c]) + (Nlo[3][c] - Nlo[2][c])) +
                              ((Nhi[1][c] - Nhi[0][c]) + (Nhi[3][c] - Nhi[2][c])));
            G[1][c] = Q.gy * (((Nlo[2][c] - Nlo[0][c]) + (Nlo[3][c] - Nlo[1][c])) +
                              ((Nhi[2][c] - Nhi[0][c]) + (Nhi[3][c] - Nhi[1][c])));
            G[2][c] = Q.gz * (((Nhi[0][c] - Nlo[0][c]) + (Nhi[1][c] - Nlo[1][c])) +
                              ((Nhi[2][c] - Nlo[2][c]) + (Nhi[3][c] - Nlo[3][c])));
        }
        T sxx, syy, szz, syz, szx, sxy;
        const long o = g.at(j, k, l);
        const bool inbox = ops.n > 0 && cellv && j >= ops.lo[0] && j <= ops.hi[0] && k >= ops.lo[1] &&
                           k <= ops.hi[1] && l >= ops.lo[2] && l <= ops.hi[2];
        if (!inbox) {
            const T tr = G[0][0] + G[1][1] + G[2][2];
            const T m2 = m + m;
            sxx = m2 * G[0][0] + lm * tr;
            syy = m2 * G[1][1] + lm * tr;
            szz = m2 * G[2][2] + lm * tr;
            syz = m * (G[2][1] + G[1][2]);
            szx = m * (G[0][2] + G[2][0]);
            sxy = m * (G[1][0] + G[0][1]);
        } else {
            // reference units: volume-integrated strain, potency / moment / stress ops
            T w1[3], w2[3];
#pragma unroll
            for (int i = 0; i < 3; i++) w1[i] = G[i][i] * Q.vol;
            w2[0] = T(0.5) * (G[2][1] * Q.vol) + T(0.5) * (G[1][2] * Q.vol);
            w2[1] = T(0.5) * (G[0][2] * Q.vol) + T(0.5) * (G[2][0] * Q.vol);
            w2[2] = T(0.5) * (G[1][0] * Q.vol) + T(0.5) * (G[0][1] * Q.vol);
            auto apply = [&](int stage) {
                for (int q = 0; q < ops.n; q++) {
                    if (ops.stage[q] != stage || ops.cell[q] != o) continue;
                    const int cc = ops.comp[q], opq = ops.op[q];
                    const T x = ops.val[q];
#pragma unroll
                    for (int t2 = 0; t2 < 6; t2++) {
                        if (t2 != cc) continue;
                        T &r = t2 < 3 ? w1[t2] : w2[t2 - 3];
                        r = opq == 0 ? x : opq == 1 ? r + x : r * x;
                    }
                }
            };
            apply(0);
            const T tr = lm * (w1[0] + w1[1] + w1[2]);
#pragma unroll
            for (int i = 0; i < 3; i++) {
                w1[i] = T(2) * m * w1[i] + tr;
                w2[i] = T(2) * m * w2[i];
            }
            apply(1);
#pragma unroll
            for (int i = 0; i < 3; i++) {
                w1[i] = w1[i] * Q.rvol;
                w2[i] = w2[i] * Q.rvol;
            }
            apply(2);
            sxx = w1[0], syy = w1[1], szz = w1[2], syz = w2[0], szx = w2[1], sxy = w2[2];
        }
        if (stat_now && cellv) {
            const T n2 = (sxx * sxx + syy * syy + szz * szz) + T(2) * (syz * syz + szx * szx + sxy * sxy);
            st_w = (n2 > st_w || n2 != n2) ? n2 : st_w;
        }
        const T A[3][3] = {{Q.ax * sxx, Q.ay * sxy, Q.az * szx},
                           {Q.ax * sxy, Q.ay * syy, Q.az * syz},
                           {Q.ax * szx, Q.ay * syz, Q.az * szz}};
#pragma unroll
        for (int c = 0; c < 3; c++) {
            // hourglass modes over the corners (d_j, d_k, d_l), sign +1 for the lower corner
            const T a00 = Nlo[0][3 + c] - Nlo[1][3 + c], a10 = Nlo[2][3 + c] - Nlo[3][3 + c];
            const T a01 = Nhi[0][3 + c] - Nhi[1][3 + c], a11 = Nhi[2][3 + c] - Nhi[3][3 + c];
            const T b00 = Nlo[0][3 + c] + Nlo[1][3 + c], b10 = Nlo[2][3 + c] + Nlo[3][3 + c];
            const T b01 = Nhi[0][3 + c] + Nhi[1][3 + c], b11 = Nhi[2][3 + c] + Nhi[3][3 + c];
            const T h1 = y * ((b00 - b10) - (b01 - b11));  // s_k s_l
            const T h2 = y * ((a00 + a10) - (a01 + a11));  // s_j s_l
            const T d0 = a00 - a10, d1 = a01 - a11;
            const T h3 = y * (d0 + d1);                    // s_j s_k
            const T h4 = y * (d0 - d1);                    // s_j s_k s_l
            const T Aj = A[c][0], Ak = A[c][1], Al = A[c][2];
            E[c][0] = Al;
            E[c][1] = Aj - h2;
            E[c][2] = Ak - h1;
            E[c][3] = -h3 - h4;
            C[c][0] = -Al;
            C[c][1] = Aj + h2;
            C[c][2] = Ak + h1;
            C[c][3] = -h3 + h4;
        }
    };

    T own6[6], carry[3][4], E[3][4], Cn[3][4];
    load_node(lz0 - 1, own6, uo, vo);
    exchange(own6, Nlo);
    load_node(lz0, own6, uo, vo);
    exchange(own6, Nhi);
    cell(lz0 - 1, E, carry);
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
        for (int c = 0; c < 6; c++) Nlo[q][c] = Nhi[q][c];

    for (int l = lz0; l <= lz1; l++) {
        const long on = g.at(jc, kc, l);
        const T mrn = Q.mr[on];
        load_node(l + 1, own6, un1, vn1);
        exchange(own6, Nhi);
        cell(l, E, Cn);
        T R[3][2];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const T q0 = E[c][0] + carry[c][0], qj = E[c][1] + carry[c][1];
            const T qk = E[c][2] + carry[c][2], qjk = E[c][3] + carry[c][3];
            R[c][0] = q0 + qk;
            R[c][1] = qj + qjk;
            sC[2 * c][w][lane] = q0 - qk;
            sC[2 * c + 1][w][lane] = qj - qjk;
#pragma unroll
            for (int q = 0; q < 4; q++) carry[c][q] = Cn[c][q];
        }
        __syncthreads();
        const int wb = w > 0 ? w - 1 : 0;
        T f[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const T r0 = R[c][0] + sC[2 * c][wb][lane];
            const T rj = R[c][1] + sC[2 * c + 1][wb][lane];
            const T send = r0 - rj;
            f[c] = (r0 + rj) + __shfl_up(send, 1);
        }
        if (own) {
            const long o = g.at(j, k, l);
            T a2 = T(0), v2 = T(0), u2 = T(0);
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const T a = f[c] * mrn;
                const T vv = vo[c] + Q.dt * a;
                const T uu = uo[c] + Q.dt * vv;
                Q.vn[c][o] = vv;
                Q.un[c][o] = uu;
                a2 = a2 + a * a;
                v2 = v2 + vv * vv;
                u2 = u2 + uu * uu;
            }
            if (stat_now) st_a = (a2 > st_a || a2 != a2) ? a2 : st_a;
            if (stat_next) {
                st_v = (v2 > st_v || v2 != v2) ? v2 : st_v;
                st_u = (u2 > st_u || u2 != u2) ? u2 : st_u;
            }
        }
#pragma unroll
        for (int c = 0; c < 3; c++) uo[c] = un1[c], vo[c] = vn1[c];
#pragma unroll
        for (int q = 0; q < 4; q++)
#pragma unroll
            for (int c = 0; c < 6; c++) Nlo[q][c] = Nhi[q][c];
    }
    if (Q.stat) {
        T *slot = (T *)Q.stat;
        if (stat_now) {
            st_a = wave_max(st_a);
            st_w = wave_max(st_w);
            if (lane == 0) {
                atomic_max_bits<T>(slot + 0, st_a);
                atomic_max_bits<T>(slot + 3, st_w);
            }
        }
        if (stat_next) {
            st_v = wave_max(st_v);
            st_u = wave_max(st_u);
            if (lane == 0) {
                atomic_max_bits<T>(slot + 1, st_v);
                atomic_max_bits<T>(slot + 2, st_u);
            }
        }
    }
}

template <class T>
void launch_fused(const Params<T> &P, const PointOps<T> &ops, const int lo[3], const int hi[3], const T sp[3],
                  int waves, int nz, void *stat, int stat_now, int stat_next, hipStream_t s) {
    FusedP<T> Q;
    for (int i = 0; i < 3; i++) {
        Q.u[i] = P.u[i], Q.v[i] = P.v[i], Q.un[i] = P.un[i], Q.vn[i] = P.vn[i];
        Q.lo[i] = lo[i], Q.hi[i] = hi[i];
    }
    if (lo[0] > hi[0] || lo[1] > hi[1] || lo[2] > hi[2]) return;
    Q.gam = P.gam, Q.mr = P.mr, Q.lam = P.lam, Q.mu = P.mu, Q.yy = P.yy;
    Q.g = P.g;
    Q.nz = nz;
    Q.gx = T(1) / (T(4) * sp[0]);
    Q.gy = T(1) / (T(4) * sp[1]);
    Q.gz = T(1) / (T(4) * sp[2]);
    Q.ax = sp[1] * sp[2] * T(0.25);
    Q.ay = sp[0] * sp[2] * T(0.25);
    Q.az = sp[0] * sp[1] * T(0.25);
    Q.h1 = P.hg_on ? P.h1 : T(0);
    Q.dth2 = P.hg_on ? P.dth2 : T(0);
    Q.dt = P.dt;
    Q.vol = sp[0] * sp[1] * sp[2];
    Q.rvol = T(1) / Q.vol;
    Q.stat = stat;
    Q.stat_now = stat_now;
    Q.stat_next = stat_next;
    Q.nbx = (hi[0] - lo[0] + 62) / 62;
    Q.nby = (hi[1] - lo[1] + waves - 2) / (waves - 2);
    Q.nbz = (hi[2] - lo[2] + nz) / nz;
    const int nblk = Q.nbx * Q.nby * Q.nbz;
    switch (waves) {
    case 4: hipLaunchKernelGGL((k_fused<T, 4>), dim3(nblk), dim3(64, 4), 0, s, Q, ops); break;
    case 8: hipLaunchKernelGGL((k_fused<T, 8>), dim3(nblk), dim3(64, 8), 0, s, Q, ops); break;
    case 12: hipLaunchKernelGGL((k_fused<T, 12>), dim3(nblk), dim3(64, 12), 0, s, Q, ops); break;
    default: hipLaunchKernelGGL((k_fused<T, 16>), dim3(nblk), dim3(64, 16), 0, s, Q, ops); break;
    }
}

template void launch_fused<float>(const Params<float> &, const PointOps<float> &, const int *, const int *,
                                  const float *, int, int, void *, int, int, hipStream_t);
template void launch_fused<double>(const Params<double> &, const PointOps<double> &, const int *, const int *,
                                   const double *, int, int, void *, int, int, hipStream_t);

}  // namespace sord
