// HIP kernels of the SORD solver step for CDNA4 (gfx950).
//
// Stage kernels (general path): each follows one stage of the reference step
// and reproduces its region bookkeeping exactly; the fast interior kernel lives
// in sord_fused.hip.
//
//   k_stress  : sord/stress.f90:16-238      strain (diff_nc + PML g), Hooke, x 1/V
//   k_accel   : sord/acceleration.f90:277-211 force (diff_cn + PML p), hourglass,
//               optional Newton + leapfrog (sord/time_integration.f90:26-41)
//   k_time    : sord/time_integration.f90:26-41
//   k_vbc     : sord/boundary_cond.f90:30-115 on boundary planes
//   k_rupture : sord/dynamic_rupture.f90:243-335 (fault plane)
//   k_fault_time : sord/time_integration.f90:15-23, 67-93 (fault plane)
//
// The mesh is rectilinear: the support operators reduce to
//   diff_nc(f, a)(cell) = (w_b w_c / 4) * sum over the 4 a-edges of the cell of (f+ - f-)
//   diff_cn(s, a)(node) = sum over the 8 cells of (w_b w_c / 4) * (+-s)
// with w_b, w_c the cell widths across a (the exact value of the reference's
// general hexahedral formula, sord/diff_nc_op.f90:22-67, for a box).
#include <mutex>

#include "sord_dev.h"
#include "sord_launch.h"

namespace sord {

// corner c = dj + 2 dk + 4 dl of a cell; the reference's summation order and the
// signs of the four hourglass modes (sord/hourglass_op.f90:7-146)
__constant__ int kHgOrd[8] = {0, 7, 6, 1, 5, 2, 3, 4};
__constant__ signed char kHgSign[4][8] = {
    {+1, +1, -1, -1, -1, -1, +1, +1},
    {+1, -1, +1, -1, -1, +1, -1, +1},
    {+1, -1, -1, +1, +1, -1, -1, +1},
    {+1, -1, -1, +1, -1, +1, +1, -1},
};

template <class T>
__device__ __forceinline__ void atomic_max_sq(T *slot, T v);
template <>
__device__ __forceinline__ void atomic_max_sq<float>(float *slot, float v) {
    // v >= 0 or NaN: magnitude bits order like the values, NaN above +inf
    atomicMax((unsigned int *)slot, __float_as_uint(v) & 0x7fffffffu);
}
template <>
__device__ __forceinline__ void atomic_max_sq<double>(double *slot, double v) {
    atomicMax((unsigned long long *)slot, (unsigned long long)__double_as_longlong(v) & 0x7fffffffffffffffull);
}
// maximum over the wavefront, then one atomic per wavefront
template <class T>
__device__ __forceinline__ void stat_max(T *slot, T v) {
    for (int o = 32; o > 0; o >>= 1) {
        T w = __shfl_xor(v, o);
        v = (w > v || w != w) ? w : v;
    }
    if ((threadIdx.x + blockDim.x * threadIdx.y) % 64 == 0) atomic_max_sq<T>(slot, v);
}

template <class T>
__device__ __forceinline__ long slab_at(const int *n, int i, int j, int k, int comp) {
    return (long)(i - 1) + (long)n[0] * ((long)(j - 1) + (long)n[1] * (long)(k - 1)) +
           (long)comp * n[0] * n[1] * n[2];
}

// ---------------------------------------------------------------------------
// stress: sord/stress.f90:16-238
// mode 0: stress (updates PML memory), 1: volume strain (potency), 2: moment,
// 3: stress without advancing the PML memory (the fused kernels advance it)
// ---------------------------------------------------------------------------
template <class T, int MODE>
__device__ __forceinline__ T stress_cell(Params<T> &P, const PointOps<T> &ops, int j, int k, int l);

template <class T, int MODE>
__global__ __launch_bounds__(256) void k_stress(Params<T> P, PointOps<T> ops, int3 lo, int3 hi) {
    const int j = lo.x + blockIdx.x * blockDim.x + threadIdx.x;
    const int k = lo.y + blockIdx.y * blockDim.y + threadIdx.y;
    const int l = lo.z + blockIdx.z;
    const bool in = j <= hi.x && k <= hi.y;
    T n2 = T(0);
    if (in) n2 = stress_cell<T, MODE>(P, ops, j, k, l);
    if ((MODE == 0 || MODE == 3) && P.stat_now) stat_max<T>(P.stat_cur + 3, n2);
}

// one cell of the stress stage; returns the squared Frobenius norm of the stress
template <class T, int MODE>
__device__ __forceinline__ T stress_cell(Params<T> &P, const PointOps<T> &ops, int j, int k, int l) {
    const Grid<T> &g = P.g;
    const int c[3] = {j, k, l};

    // region flags (cells): lower slab c <= i1pml, upper slab c >= i2pml - 1
    bool inLo[3], inHi[3], inS[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        inLo[a] = c[a] <= P.i1pml[a];
        inHi[a] = c[a] >= P.i2pml[a] - 1;
        inS[a] = inLo[a] || inHi[a];
    }
    // corner values u, v, w = u + gam v (sord/time_integration.f90:36)
    T U[3][8], V[3][8], W[3][8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const long o = g.at(j + (q & 1), k + ((q >> 1) & 1), l + ((q >> 2) & 1));
        const T ga = P.gam[o];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            U[i][q] = P.u[i][o];
            V[i][q] = P.v[i][o];
            W[i][q] = U[i][q] + ga * V[i][q];
        }
    }
    const T wx = P.hx[j], wy = P.hy[k], wz = P.hz[l];
    const T coef[3] = {wy * wz * T(0.25), wx * wz * T(0.25), wx * wy * T(0.25)};
    const long oc = g.at(j, k, l);
    T e[3][3];
#pragma unroll
    for (int ic = 0; ic < 3; ic++) {
#pragma unroll
        for (int id = 0; id < 3; id++) {
            bool other = false;
#pragma unroll
            for (int a = 0; a < 3; a++)
                if (a != id && inS[a]) other = true;
            const bool useV = inS[id], useU = !inS[id] && other;
            // edge differences along index axis m, over the 4 edges (other two corner bits)
            auto edges = [&](int m) {
                const int bit = 1 << m;
                T s = T(0);
                int first = 1;
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    if (q & bit) continue;
                    const T fp = useV ? V[ic][q | bit] : useU ? U[ic][q | bit] : W[ic][q | bit];
                    const T fm = useV ? V[ic][q] : useU ? U[ic][q] : W[ic][q];
                    const T d = fp - fm;
                    s = first ? d : s + d;
                    first = 0;
                }
                return s;
            };
            T D;
            if (!P.gen) {
                D = coef[id] * edges(id);
            } else {  // general cell: its corner weights along id
                D = T(0);
#pragma unroll
                for (int q = 0; q < 8; q++)
                    D = D + P.cw[8 * id + q][oc] * (useV ? V[ic][q] : useU ? U[ic][q] : W[ic][q]);
            }
            if (inS[id]) {
                // damped direction: g' recursion (sord/stress.f90:42-131)
                const int a = id;
                int si, pidx, sl;
                if (inLo[a]) {
                    si = c[a] - P.i1cell[a] + 1;
                    pidx = c[a] + g.nnoff[a];
                    sl = a;
                } else {
                    si = P.i2cell[a] - c[a] + 1;
                    pidx = g.nn[a] - c[a] - g.nnoff[a];
                    sl = a + 3;
                }
                int sj = j, sk = k, sll = l;
                if (a == 0) sj = si;
                if (a == 1) sk = si;
                if (a == 2) sll = si;
                const long go = slab_at<T>(P.gn[sl], sj, sk, sll, ic);
                T *gp = P.gm[sl];
                D = P.dc2[pidx - 1] * D + P.dc1[pidx - 1] * gp[go];
                if (MODE == 0) gp[go] = D;
            }
            e[ic][id] = D;
        }
    }
    // symmetric strain (sord/stress.f90:133-165): diag, then 0.5*(e_ij) + 0.5*(e_ji)
    T w1[3], w2[3];
    w1[0] = e[0][0];
    w1[1] = e[1][1];
    w1[2] = e[2][2];
    w2[0] = T(0.5) * e[1][2] + T(0.5) * e[2][1];  // yz
    w2[1] = T(0.5) * e[0][2] + T(0.5) * e[2][0];  // zx
    w2[2] = T(0.5) * e[0][1] + T(0.5) * e[1][0];  // xy

    const long o = g.at(j, k, l);
    const unsigned boxes = ops.n > 0 ? ops_boxes(ops, (1u << ops.nbox) - 1, j, k, l) : 0u;
    auto apply = [&](int stage) {
        for (unsigned mb = boxes; mb; mb &= mb - 1) {
            const int b = __builtin_ctz(mb);
            for (int q = ops.bop[b]; q < ops.bop[b + 1]; q++) {
                if (ops.stage[q] != stage || ops.cell[q] != o) continue;
                const int cc = ops.comp[q], opq = ops.op[q];
                const T x = ops.val[q];
#pragma unroll
                for (int t = 0; t < 6; t++) {
                    if (t != cc) continue;
                    T &r = t < 3 ? w1[t] : w2[t - 3];
                    r = opq == 0 ? x : opq == 1 ? r + x : r * x;
                }
            }
        }
    };
    // finite source values of this step (sord/kinematic_source.f90:67-118),
    // added before the field-I/O ops of their stage, as the reference calls it
    auto add_fs = [&](int kind) {
        if (P.fs_kind != kind) return;
        const int dj = j - P.fs_lo[0], dk = k - P.fs_lo[1], dl = l - P.fs_lo[2];
        if (dj < 0 || dj >= P.fs_n[0] || dk < 0 || dk >= P.fs_n[1] || dl < 0 || dl >= P.fs_n[2]) return;
        const long q = dj + (long)P.fs_n[0] * (dk + (long)P.fs_n[1] * dl);
#pragma unroll
        for (int i = 0; i < 3; i++) w1[i] = w1[i] + P.fs[i][q], w2[i] = w2[i] + P.fs[3 + i][q];
    };
    add_fs(1);
    apply(0);
    if (MODE == 1) {
#pragma unroll
        for (int i = 0; i < 3; i++) {
            P.sig[i][o] = w1[i];
            P.sig[i + 3][o] = w2[i];
        }
        return T(0);
    }
    // Hooke's law (sord/stress.f90:192-213)
    const T lm = P.lam[o], m2 = T(2) * P.mu[o];
    const T tr = lm * (w1[0] + w1[1] + w1[2]);
#pragma unroll
    for (int i = 0; i < 3; i++) {
        w1[i] = m2 * w1[i] + tr;
        w2[i] = m2 * w2[i];
    }
    add_fs(2);
    apply(1);
    if (MODE == 2) {
#pragma unroll
        for (int i = 0; i < 3; i++) {
            P.sig[i][o] = w1[i];
            P.sig[i + 3][o] = w2[i];
        }
        return T(0);
    }
    const T rv = P.vc[o];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        w1[i] = w1[i] * rv;
        w2[i] = w2[i] * rv;
    }
    apply(2);
#pragma unroll
    for (int i = 0; i < 3; i++) {
        P.sig[i][o] = w1[i];
        P.sig[i + 3][o] = w2[i];
    }
    return (w1[0] * w1[0] + w1[1] * w1[1] + w1[2] * w1[2]) + T(2) * (w2[0] * w2[0] + w2[1] * w2[1] + w2[2] * w2[2]);
}

// ---------------------------------------------------------------------------
// acceleration: sord/acceleration.f90:277-211
// MODE 0: write the force to P.a
// MODE 1: local boundary conditions, a = f * mr into P.a
// MODE 2: as 1, then leapfrog into (un, vn); nodes on the two fault planes
//         store the force in P.a and are advanced by the rupture kernel
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ void local_bc(const Params<T> &P, const int n[3], T f[3]) {
    // sord/boundary_cond.f90:44-66: -1 anti-mirror, 1 mirror, 3 rigid (plane-local)
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const int b = (a + 1) % 3, c = (a + 2) % 3;
        if (n[a] == P.i1bc[a]) {
            const int t = P.bc1[a] % 10;
            if (t == -1) f[b] = f[c] = T(0);
            if (t == 1) f[a] = T(0);
            if (t == 3) f[0] = f[1] = f[2] = T(0);
        }
        if (n[a] == P.i2bc[a]) {
            const int t = P.bc2[a] % 10;
            if (t == -1) f[b] = f[c] = T(0);
            if (t == 1) f[a] = T(0);
            if (t == 3) f[0] = f[1] = f[2] = T(0);
        }
    }
}

template <class T, int MODE>
__device__ __forceinline__ void accel_node(Params<T> &P, int j, int k, int l, T st[3]);

template <class T, int MODE>
__global__ __launch_bounds__(256) void k_accel(Params<T> P, int3 lo, int3 hi) {
    const int j = lo.x + blockIdx.x * blockDim.x + threadIdx.x;
    const int k = lo.y + blockIdx.y * blockDim.y + threadIdx.y;
    const int l = lo.z + blockIdx.z;
    T st[3] = {T(0), T(0), T(0)};  // |a|^2, |v|^2, |u|^2 of the updated node
    if (j <= hi.x && k <= hi.y) accel_node<T, MODE>(P, j, k, l, st);
    if (MODE == 2 && P.stat_now) stat_max<T>(P.stat_cur + 0, st[0]);
    if (MODE == 2 && P.stat_next) {
        stat_max<T>(P.stat_nxt + 1, st[1]);
        stat_max<T>(P.stat_nxt + 2, st[2]);
    }
}

template <class T, int MODE>
__device__ __forceinline__ void accel_node(Params<T> &P, int j, int k, int l, T st[3]) {
    const Grid<T> &g = P.g;
    const int n[3] = {j, k, l};
    const long o = g.at(j, k, l);

    // cell widths around the node: index 0 -> cell n-1, 1 -> cell n
    const T wxm = P.hx[j - 1], wxp = P.hx[j];
    const T wym = P.hy[k - 1], wyp = P.hy[k];
    const T wzm = P.hz[l - 1], wzp = P.hz[l];
    const T WX[2] = {wxm, wxp}, WY[2] = {wym, wyp}, WZ[2] = {wzm, wzp};

    T f[3];
#pragma unroll
    for (int ic = 0; ic < 3; ic++) {
        T acc = T(0);
#pragma unroll
        for (int iid = 0; iid < 3; iid++) {
            const int id = (ic + iid) % 3;
            const int sc = ic == id ? ic : 3 + (3 - ic - id);
            const T *S = P.sig[sc];
            // diff_cn along id: sum over the 4 cell pairs (minus side, plus side)
            T s = T(0);
            int first = 1;
            if (P.gen) {
                // general cells: minus the transpose of the cell difference, the
                // node being corner 7 - q of the cell at offset q
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    const long oq = g.at(j - 1 + (q & 1), k - 1 + ((q >> 1) & 1), l - 1 + ((q >> 2) & 1));
                    const T t = -(P.cw[8 * id + 7 - q][oq] * S[oq]);
                    s = first ? t : s + t;
                    first = 0;
                }
            }
#pragma unroll
            for (int q = 0; q < 4 && !P.gen; q++) {
                int d[3];
                const int b = (id + 1) % 3, cc = (id + 2) % 3;
                d[b] = q & 1;
                d[cc] = (q >> 1) & 1;
                d[id] = 0;
                const T wb = b == 0 ? WX[d[0]] : b == 1 ? WY[d[1]] : WZ[d[2]];
                const T wc = cc == 0 ? WX[d[0]] : cc == 1 ? WY[d[1]] : WZ[d[2]];
                const int cj = j - 1 + (id == 0 ? 0 : d[0]);
                const int ck = k - 1 + (id == 1 ? 0 : d[1]);
                const int cl = l - 1 + (id == 2 ? 0 : d[2]);
                const long om = g.at(cj, ck, cl);
                const long op = g.at(cj + (id == 0), ck + (id == 1), cl + (id == 2));
                const T t = (wb * wc * T(0.25)) * (S[op] - S[om]);
                s = first ? t : s + t;
                first = 0;
            }
            // PML memory (sord/acceleration.f90:295-362)
            const int a = id;
            if (n[a] <= P.i1pml[a]) {
                const int si = n[a] - P.i1node[a] + 1, pidx = n[a] + g.nnoff[a];
                int sj = j, sk = k, sl = l;
                if (a == 0) sj = si;
                if (a == 1) sk = si;
                if (a == 2) sl = si;
                T *pp = P.p[a];
                const long po = slab_at<T>(P.pn[a], sj, sk, sl, ic);
                s = P.dn2[pidx - 1] * s + P.dn1[pidx - 1] * pp[po];
                pp[po] = pp[po] + P.dt * s;
            }
            if (n[a] >= P.i2pml[a]) {
                const int si = P.i2node[a] - n[a] + 1, pidx = g.nn[a] - n[a] - g.nnoff[a] + 1;
                int sj = j, sk = k, sl = l;
                if (a == 0) sj = si;
                if (a == 1) sk = si;
                if (a == 2) sl = si;
                T *pp = P.p[a + 3];
                const long po = slab_at<T>(P.pn[a + 3], sj, sk, sl, ic);
                s = P.dn2[pidx - 1] * s + P.dn1[pidx - 1] * pp[po];
                pp[po] = pp[po] + P.dt * s;
            }
            acc = iid == 0 ? s : acc + s;
        }
        f[ic] = acc;
    }

    // hourglass control, only viscous in PML (sord/acceleration.f90:390-469)
    if (P.hg_on) {
        bool interior = true;
#pragma unroll
        for (int a = 0; a < 3; a++)
            if (n[a] < P.i1pml[a] + 1 || n[a] > P.i2pml[a] - 1) interior = false;
        if (interior || P.hg_pml) {
            // 27 node values of the filtered field, per component
            T F[3][27];
#pragma unroll
            for (int q = 0; q < 27; q++) {
                const int dj = q % 3 - 1, dk = (q / 3) % 3 - 1, dl = q / 9 - 1;
                const long oq = g.at(j + dj, k + dk, l + dl);
#pragma unroll
                for (int i = 0; i < 3; i++)
                    F[i][q] = interior ? P.h1 * P.u[i][oq] + P.dth2 * P.v[i][oq] : P.v[i][oq];
            }
            // cell weights: yy (interior) or dt*h2*yy (PML)
            T Y[8];
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int dj = q & 1, dk = (q >> 1) & 1, dl = (q >> 2) & 1;
                const T yv = P.yy[g.at(j - dj, k - dk, l - dl)];
                Y[q] = interior ? yv : P.dth2 * yv;
            }
#pragma unroll
            for (int iq = 0; iq < 4; iq++) {
#pragma unroll
                for (int ic = 0; ic < 3; ic++) {
                    T s2 = T(0);
#pragma unroll
                    for (int qc = 0; qc < 8; qc++) {
                        // cell at offset -(dj,dk,dl) from the node, in cn order
                        const int cq = kHgOrd[qc];
                        const int cj = -(cq & 1), ck = -((cq >> 1) & 1), cl = -((cq >> 2) & 1);
                        // hourglass_nc over the cell's 8 corners, nc order
                        T h = T(0);
#pragma unroll
                        for (int qn = 0; qn < 8; qn++) {
                            const int nq = kHgOrd[qn];
                            const int dj = cj + (nq & 1), dk = ck + ((nq >> 1) & 1),
                                      dl = cl + ((nq >> 2) & 1);
                            const T x = F[ic][(dj + 1) + 3 * (dk + 1) + 9 * (dl + 1)];
                            const int sg = kHgSign[iq][nq];
                            h = qn == 0 ? (sg > 0 ? x : -x) : (sg > 0 ? h + x : h - x);
                        }
                        const T sv = interior ? h * Y[cq] : Y[cq] * h;
                        const int sg = kHgSign[iq][cq];
                        s2 = qc == 0 ? (sg > 0 ? sv : -sv) : (sg > 0 ? s2 + sv : s2 - sv);
                    }
                    f[ic] = f[ic] - s2;
                }
            }
        }
    }

    if (MODE == 0) {
#pragma unroll
        for (int i = 0; i < 3; i++) P.a[i][o] = f[i];
        return;
    }
    local_bc(P, n, f);
    if (MODE == 2 && P.ifn && (n[P.ifn - 1] == P.irup || n[P.ifn - 1] == P.irup + 1)) {
#pragma unroll
        for (int i = 0; i < 3; i++) P.a[i][o] = f[i];
        return;
    }
    const T m = P.mr[o];
#pragma unroll
    for (int i = 0; i < 3; i++) f[i] = f[i] * m;
    if (MODE == 1 || P.store_a) {
#pragma unroll
        for (int i = 0; i < 3; i++) P.a[i][o] = f[i];
        if (MODE == 1) return;
    }
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const T vv = P.v[i][o] + P.dt * f[i];
        const T uu = P.u[i][o] + P.dt * vv;
        P.vn[i][o] = vv;
        P.un[i][o] = uu;
        st[0] = st[0] + f[i] * f[i];
        st[1] = st[1] + vv * vv;
        st[2] = st[2] + uu * uu;
    }
}

// ---------------------------------------------------------------------------
// leapfrog: sord/time_integration.f90:26-41 (the viscous w1 = u + gam v is
// formed where it is used, in k_stress)
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(256) void k_time(Params<T> P, int3 lo, int3 hi) {
    const int j = lo.x + blockIdx.x * blockDim.x + threadIdx.x;
    const int k = lo.y + blockIdx.y * blockDim.y + threadIdx.y;
    const int l = lo.z + blockIdx.z;
    if (j > hi.x || k > hi.y) return;
    const long o = P.g.at(j, k, l);
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const T vv = P.v[i][o] + P.dt * P.a[i][o];
        P.v[i][o] = vv;
        P.u[i][o] = P.u[i][o] + P.dt * vv;
    }
}

// a = f * mr over a node box (sord/acceleration.f90:498-509)
template <class T>
__global__ __launch_bounds__(256) void k_newton(Params<T> P, int3 lo, int3 hi) {
    const int j = lo.x + blockIdx.x * blockDim.x + threadIdx.x;
    const int k = lo.y + blockIdx.y * blockDim.y + threadIdx.y;
    const int l = lo.z + blockIdx.z;
    if (j > hi.x || k > hi.y) return;
    const long o = P.g.at(j, k, l);
    const T m = P.mr[o];
#pragma unroll
    for (int i = 0; i < 3; i++) P.a[i][o] = P.a[i][o] * m;
}

// ---------------------------------------------------------------------------
// generic strided box operations on one 3-D array (field i/o)
// ---------------------------------------------------------------------------
// out[q] = f(box point q), q in Fortran order of the box
template <class T>
__global__ void k_gather(const T *f, View g, Box b, T *out) {
    const long N = (long)b.n[0] * b.n[1] * b.n[2];
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < N; q += (long)gridDim.x * blockDim.x) {
        const int i = q % b.n[0], jj = (q / b.n[0]) % b.n[1], kk = q / ((long)b.n[0] * b.n[1]);
        out[q] = f[g.at(b.lo[0] + i * b.st[0], b.lo[1] + jj * b.st[1], b.lo[2] + kk * b.st[2])];
    }
}

// f(box) op= src (src == nullptr: the scalar val); op 0 '=', 1 '+', 2 '*'
template <class T>
__global__ void k_boxop(T *f, View g, Box b, int op, T val, const T *src) {
    const long N = (long)b.n[0] * b.n[1] * b.n[2];
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < N; q += (long)gridDim.x * blockDim.x) {
        const int i = q % b.n[0], jj = (q / b.n[0]) % b.n[1], kk = q / ((long)b.n[0] * b.n[1]);
        const long o = g.at(b.lo[0] + i * b.st[0], b.lo[1] + jj * b.st[1], b.lo[2] + kk * b.st[2]);
        const T x = src ? src[q] : val;
        if (op == 0) f[o] = x;
        else if (op == 1) f[o] = f[o] + x;
        else f[o] = f[o] * x;
    }
}

// f(plane dst) = f(plane src) along axis ax over the full extents n (scalar_bc,
// sord/boundary_cond.f90:7-27)
template <class T>
__global__ void k_plane_copy(T *f, View g, int n0, int n1, int n2, int ax, int dst, int src) {
    const int n[3] = {n0, n1, n2};
    const int b = ax == 0 ? 1 : 0, c = ax == 2 ? 1 : 2;  // in-plane axes, the lower (faster in memory) first
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < (long)n[b] * n[c]; q += (long)gridDim.x * blockDim.x) {
        int idx[3];
        idx[b] = 1 + q % n[b];
        idx[c] = 1 + q / n[b];
        idx[ax] = src;
        const T x = f[g.at(idx[0], idx[1], idx[2])];
        idx[ax] = dst;
        f[g.at(idx[0], idx[1], idx[2])] = x;
    }
}

// sub-cell samples: out[r] = sum over the 8 weights of f (sord/input_output.f90:470-479)
template <class T>
__global__ void k_sample(const T *f, View g, Samples<T> s, T *out) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= s.n) return;
    T acc = T(0);
    for (int l = s.lo[r][2]; l <= s.hi[r][2]; l++)
        for (int k = s.lo[r][1]; k <= s.hi[r][1]; k++)
            for (int j = s.lo[r][0]; j <= s.hi[r][0]; j++) {
                const T w = (T(1) - fabs(s.xi[r][0] - T(j))) * (T(1) - fabs(s.xi[r][1] - T(k))) *
                            (T(1) - fabs(s.xi[r][2] - T(l)));
                acc = acc + f[g.at(j, k, l)] * w;
            }
    out[s.slot[r]] = acc;
}

// f(j,k,l) += val * w over the 8 nodes of a sub-cell point (sord/input_output.f90:183-191)
template <class T>
__global__ void k_point_add(T *f, View g, int3 lo, int3 hi, T x0, T x1, T x2, T val) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (int l = lo.z; l <= hi.z; l++)
        for (int k = lo.y; k <= hi.y; k++)
            for (int j = lo.x; j <= hi.x; j++) {
                const T r = (T(1) - fabs(x0 - T(j))) * (T(1) - fabs(x1 - T(k))) * (T(1) - fabs(x2 - T(l)));
                const long o = g.at(j, k, l);
                f[o] = f[o] + val * r;
            }
}

// squared norms (sord/utilities.f90:120-166): kind 0 vector of 3, 1 tensor (w1, w2)
template <class T>
__global__ void k_norm(T *out, const T *a0, const T *a1, const T *a2, const T *b0, const T *b1,
                       const T *b2, Grid<T> g, Box b) {
    const long N = (long)b.n[0] * b.n[1] * b.n[2];
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < N; q += (long)gridDim.x * blockDim.x) {
        const int i = q % b.n[0], jj = (q / b.n[0]) % b.n[1], kk = q / ((long)b.n[0] * b.n[1]);
        const long o = g.at(b.lo[0] + i * b.st[0], b.lo[1] + jj * b.st[1], b.lo[2] + kk * b.st[2]);
        T s = a0[o] * a0[o] + a1[o] * a1[o] + a2[o] * a2[o];
        if (b0) s = s + T(2) * (b0[o] * b0[o] + b1[o] * b1[o] + b2[o] * b2[o]);
        out[o] = s;
    }
}

// max and first location (Fortran order over the whole padded array) of a
// squared-norm field over the core box; other points count as -1
// (sord/time_integration.f90:47-52).  Two-pass: per-block partials, then host.
template <class T>
__global__ void k_maxnorm(const T *a0, const T *a1, const T *a2, const T *b0, const T *b1, const T *b2,
                          Grid<T> g, Box b, T *pval, long *ploc) {
    __shared__ T sv[256];
    __shared__ long sl[256];
    const long N = (long)b.n[0] * b.n[1] * b.n[2];
    T best = T(-1);
    long bl = 0x7fffffffffffffffL;
    bool nan = false;
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < N; q += (long)gridDim.x * blockDim.x) {
        const int i = q % b.n[0], jj = (q / b.n[0]) % b.n[1], kk = q / ((long)b.n[0] * b.n[1]);
        const int j = b.lo[0] + i, k = b.lo[1] + jj, l = b.lo[2] + kk;
        const long o = g.at(j, k, l);
        T s = a0[o] * a0[o] + a1[o] * a1[o] + a2[o] * a2[o];
        if (b0) s = s + T(2) * (b0[o] * b0[o] + b1[o] * b1[o] + b2[o] * b2[o]);
        const long lin = (long)(j - 1) + (long)g.nm[0] * ((long)(k - 1) + (long)g.nm[1] * (l - 1));
        if (s != s) nan = true;
        else if (s > best || (s == best && lin < bl)) {
            best = s;
            bl = lin;
        }
    }
    if (nan) best = T(NAN);
    sv[threadIdx.x] = best;
    sl[threadIdx.x] = bl;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            const T x = sv[threadIdx.x + w], cur = sv[threadIdx.x];
            const long y = sl[threadIdx.x + w];
            if (x != x) sv[threadIdx.x] = x;
            else if (cur == cur && (x > cur || (x == cur && y < sl[threadIdx.x]))) {
                sv[threadIdx.x] = x;
                sl[threadIdx.x] = y;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        pval[blockIdx.x] = sv[0];
        ploc[blockIdx.x] = sl[0];
    }
}

// ---------------------------------------------------------------------------
// vector boundary conditions on boundary planes (sord/boundary_cond.f90:30-115)
// one launch per rule and face: kind 0 zero components mask, 1 copy with signs
// ---------------------------------------------------------------------------
template <class T>
__global__ void k_vbc_plane(Params<T> P, int ax, int dst, int src, int mask, T s0, T s1, T s2, int kind) {
    const Grid<T> &g = P.g;
    const int b = ax == 0 ? 1 : 0, c = ax == 2 ? 1 : 2;  // in-plane axes, the lower (faster in memory) first
    const int nb = g.nm[b], nc = g.nm[c];
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < (long)nb * nc; q += (long)gridDim.x * blockDim.x) {
        int idx[3];
        idx[ax] = dst;
        idx[b] = 1 + q % nb;
        idx[c] = 1 + q / nb;
        const long o = g.at(idx[0], idx[1], idx[2]);
        if (kind == 0) {
            for (int i = 0; i < 3; i++)
                if (mask & (1 << i)) P.a[i][o] = T(0);
        } else {
            idx[ax] = src;
            const long os = g.at(idx[0], idx[1], idx[2]);
            const T sg[3] = {s0, s1, s2};
            for (int i = 0; i < 3; i++) P.a[i][o] = sg[i] * P.a[i][os];
        }
    }
}

// ---------------------------------------------------------------------------
// rupture (sord/dynamic_rupture.f90:243-335) on the fault plane.
// Fault arrays are 2-D planes (the local mesh with extent 1 along the normal),
// Fortran order.  Writes the traction outputs and, in fused mode, the Newton
// and leapfrog update of both sides of the fault.
// ---------------------------------------------------------------------------
template <class T>
__global__ void k_rupture(Params<T> P, Fault<T> F, int it, T tm, int fused) {
    const Grid<T> &g = P.g;
    const int a = P.ifn - 1;
    const int b = a == 0 ? 1 : 0, c = a == 2 ? 1 : 2;  // in-plane axes, the lower (faster in memory) first
    const int nb = g.nm[b], nc = g.nm[c];
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < (long)nb * nc; q += (long)gridDim.x * blockDim.x) {
        int idx[3];
        idx[a] = P.irup;
        idx[b] = 1 + q % nb;
        idx[c] = 1 + q / nb;
        // fault-array linear index (Fortran order with extent 1 along the normal)
        int fi[3] = {idx[0], idx[1], idx[2]};
        fi[a] = 1;
        const long fq = (long)(fi[0] - 1) + (long)F.fm[0] * ((long)(fi[1] - 1) + (long)F.fm[1] * (fi[2] - 1));
        const long om = g.at(idx[0], idx[1], idx[2]);
        idx[a] = P.irup + 1;
        const long op = g.at(idx[0], idx[1], idx[2]);
        const T mrm = P.mr[om], mrp = P.mr[op];
        const T ar = F.area[fq];
        T f1 = P.dt * P.dt * ar * (mrm + mrp);
        if (f1 != T(0)) f1 = T(1) / f1;
        const T nh[3] = {F.nhat[0][fq], F.nhat[1][fq], F.nhat[2][fq]};
        T t1[3], t2[3], t3[3];
        for (int i = 0; i < 3; i++) {
            t1[i] = F.t0[i][fq] + f1 * P.dt *
                                      (P.v[i][op] - P.v[i][om] + P.a[i][op] * mrp * P.dt - P.a[i][om] * mrm * P.dt);
            t2[i] = t1[i] + f1 * (P.u[i][op] - P.u[i][om]);
        }
        T tn = t1[0] * nh[0] + t1[1] * nh[1] + t1[2] * nh[2];
        for (int i = 0; i < 3; i++) t3[i] = t1[i] - tn * nh[i];
        T ts = sqrt(t3[0] * t3[0] + t3[1] * t3[1] + t3[2] * t3[2]);
        T fr = f1;
        if (it > 1) {
            tn = t2[0] * nh[0] + t2[1] * nh[1] + t2[2] * nh[2];
            if (F.faultopening == 1) tn = fmin(T(0), tn);
            const T sl = F.sl[fq], dc = F.dc[fq], mus = F.mus[fq], mud = F.mud[fq], co = F.co[fq];
            fr = mud;
            if (sl < dc) fr = fr + (T(1) - sl / dc) * (mus - mud);
            fr = -(fmin(T(0), tn) * fr) + co;
            if (F.nucleate) {
                T f2 = T(1);
                if (F.trelax > T(0)) f2 = fmin((tm - F.rhypo[fq] / F.vrup) / F.trelax, T(1));
                f2 = (T(1) - f2) * ts + f2 * (-(tn * mud) + co);
                if (F.rhypo[fq] < fmin(F.rcrit, tm * F.vrup) && f2 < fr) fr = f2;
            }
            T f2 = T(1);
            if (ts > fr) f2 = fr / ts;
            for (int i = 0; i < 3; i++) t3[i] = f2 * t3[i];
            ts = fmin(ts, fr);
            for (int i = 0; i < 3; i++) t1[i] = t3[i] + tn * nh[i];
        }
        // friction work, strain energy and moment density (sord/dynamic_rupture.f90:353-368)
        {
            T dv[3], du[3];
            for (int i = 0; i < 3; i++) {
                dv[i] = P.v[i][op] - P.v[i][om];
                du[i] = P.u[i][op] - P.u[i][om];
            }
            const T ef = (t1[0] * dv[0] + t1[1] * dv[1] + t1[2] * dv[2]) * ar;
            F.efacc[fq] = F.efacc[fq] + P.dt * ef;
            F.es[fq] = ((F.t0[0][fq] + t1[0]) * du[0] + (F.t0[1][fq] + t1[1]) * du[1] +
                        (F.t0[2][fq] + t1[2]) * du[2]) *
                       ar;
            F.mo[fq] = F.muf[fq] * ar * sqrt(du[0] * du[0] + du[1] * du[1] + du[2] * du[2]);
        }
        T fm_[3], fp_[3];
        for (int i = 0; i < 3; i++) {
            const T d = ar * (t1[i] - F.t0[i][fq]);
            fm_[i] = P.a[i][om] + d;
            fp_[i] = P.a[i][op] - d;
        }
        // outputs of this step
        for (int i = 0; i < 3; i++) {
            F.t1[i][fq] = t1[i];
            F.t3[i][fq] = t3[i];
        }
        F.ts[fq] = ts;
        F.tn[fq] = tn;
        F.fr[fq] = fr;
        // boundary conditions on both sides (local rules; others by the host)
        int nmn[3] = {idx[0], idx[1], idx[2]}, npn[3] = {idx[0], idx[1], idx[2]};
        nmn[a] = P.irup;
        npn[a] = P.irup + 1;
        if (fused) {
            local_bc(P, nmn, fm_);
            local_bc(P, npn, fp_);
        }
        if (!fused) {
            for (int i = 0; i < 3; i++) {
                P.a[i][om] = fm_[i];
                P.a[i][op] = fp_[i];
            }
            continue;
        }
        // slip acceleration output from the forces (sord/dynamic_rupture.f90:370-376)
        T sa2 = T(0);
        for (int i = 0; i < 3; i++) {
            F.sa[i][fq] = fp_[i] * mrp - fm_[i] * mrm;
            sa2 = sa2 + F.sa[i][fq] * F.sa[i][fq];
            P.a[i][om] = fm_[i] * mrm;
            P.a[i][op] = fp_[i] * mrp;
        }
        F.sam[fq] = sqrt(sa2);
        // Newton and leapfrog for both sides; previous slip rate for the next step
        T sv0 = T(0);
        for (int i = 0; i < 3; i++) {
            const T d = P.v[i][op] - P.v[i][om];
            sv0 = sv0 + d * d;
        }
        F.svold[fq] = sqrt(sv0);
        T sa_m = 0, sa_p = 0, sv_m = 0, sv_p = 0, su_m = 0, su_p = 0;
        for (int i = 0; i < 3; i++) {
            const T am = fm_[i] * mrm, ap = fp_[i] * mrp;
            const T vm = P.v[i][om] + P.dt * am, vp = P.v[i][op] + P.dt * ap;
            const T um = P.u[i][om] + P.dt * vm, up = P.u[i][op] + P.dt * vp;
            P.vn[i][om] = vm;
            P.vn[i][op] = vp;
            P.un[i][om] = um;
            P.un[i][op] = up;
            sa_m += am * am, sa_p += ap * ap, sv_m += vm * vm, sv_p += vp * vp, su_m += um * um, su_p += up * up;
        }
        // statistics over the fault nodes (maxima; one atomic per node pair, the plane is small)
        if (P.stat_now) atomic_max_sq<T>(P.stat_cur + 0, fmax(sa_m, sa_p));
        if (P.stat_next) {
            atomic_max_sq<T>(P.stat_nxt + 1, fmax(sv_m, sv_p));
            atomic_max_sq<T>(P.stat_nxt + 2, fmax(su_m, su_p));
        }
    }
}

// fault part of step_time after the leapfrog (sord/time_integration.f90:67-93)
template <class T>
__global__ void k_fault_time(Params<T> P, Fault<T> F, T tm, T svtol) {
    const Grid<T> &g = P.g;
    const int a = P.ifn - 1;
    const int b = a == 0 ? 1 : 0, c = a == 2 ? 1 : 2;  // in-plane axes, the lower (faster in memory) first
    const int nb = g.nm[b], nc = g.nm[c];
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < (long)nb * nc; q += (long)gridDim.x * blockDim.x) {
        int idx[3];
        idx[a] = P.irup;
        idx[b] = 1 + q % nb;
        idx[c] = 1 + q / nb;
        int fi[3] = {idx[0], idx[1], idx[2]};
        fi[a] = 1;
        const long fq = (long)(fi[0] - 1) + (long)F.fm[0] * ((long)(fi[1] - 1) + (long)F.fm[1] * (fi[2] - 1));
        const long om = g.at(idx[0], idx[1], idx[2]);
        idx[a] = P.irup + 1;
        const long op = g.at(idx[0], idx[1], idx[2]);
        T t1[3], t2[3];
        for (int i = 0; i < 3; i++) {
            t1[i] = P.v[i][op] - P.v[i][om];
            t2[i] = P.u[i][op] - P.u[i][om];
        }
        const T f1 = sqrt(t1[0] * t1[0] + t1[1] * t1[1] + t1[2] * t1[2]);
        const T f2 = F.svold[fq];
        F.sl[fq] = F.sl[fq] + P.dt * f1;
        F.psv[fq] = fmax(F.psv[fq], f1);
        if (svtol > T(0)) {
            if (f1 >= svtol && F.trup[fq] > T(1e8)) F.trup[fq] = tm - P.dt * (T(0.5) + (svtol - f1) / (f2 - f1));
            if (f1 >= svtol) F.tarr[fq] = T(1e9);
            if (f1 < svtol && f2 >= svtol) F.tarr[fq] = tm - P.dt * (T(0.5) + (svtol - f1) / (f2 - f1));
        }
        for (int i = 0; i < 3; i++) {
            F.sv[i][fq] = t1[i];
            F.su[i][fq] = t2[i];
        }
        F.svm[fq] = f1;
        F.sum[fq] = sqrt(t2[0] * t2[0] + t2[1] * t2[1] + t2[2] * t2[2]);
    }
}

// previous slip rate before the leapfrog (sord/time_integration.f90:15-23)
template <class T>
__global__ void k_fault_svold(Params<T> P, Fault<T> F) {
    const Grid<T> &g = P.g;
    const int a = P.ifn - 1;
    const int b = a == 0 ? 1 : 0, c = a == 2 ? 1 : 2;  // in-plane axes, the lower (faster in memory) first
    const int nb = g.nm[b], nc = g.nm[c];
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < (long)nb * nc; q += (long)gridDim.x * blockDim.x) {
        int idx[3];
        idx[a] = P.irup;
        idx[b] = 1 + q % nb;
        idx[c] = 1 + q / nb;
        int fi[3] = {idx[0], idx[1], idx[2]};
        fi[a] = 1;
        const long fq = (long)(fi[0] - 1) + (long)F.fm[0] * ((long)(fi[1] - 1) + (long)F.fm[1] * (fi[2] - 1));
        const long om = g.at(idx[0], idx[1], idx[2]);
        idx[a] = P.irup + 1;
        const long op = g.at(idx[0], idx[1], idx[2]);
        T s = T(0);
        for (int i = 0; i < 3; i++) {
            const T d = P.v[i][op] - P.v[i][om];
            s = s + d * d;
        }
        F.svold[fq] = sqrt(s);
    }
}

// slip acceleration after Newton in split mode (sord/dynamic_rupture.f90:370-376)
template <class T>
__global__ void k_fault_sa(Params<T> P, Fault<T> F) {
    const Grid<T> &g = P.g;
    const int a = P.ifn - 1;
    const int b = a == 0 ? 1 : 0, c = a == 2 ? 1 : 2;  // in-plane axes, the lower (faster in memory) first
    const int nb = g.nm[b], nc = g.nm[c];
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < (long)nb * nc; q += (long)gridDim.x * blockDim.x) {
        int idx[3];
        idx[a] = P.irup;
        idx[b] = 1 + q % nb;
        idx[c] = 1 + q / nb;
        int fi[3] = {idx[0], idx[1], idx[2]};
        fi[a] = 1;
        const long fq = (long)(fi[0] - 1) + (long)F.fm[0] * ((long)(fi[1] - 1) + (long)F.fm[1] * (fi[2] - 1));
        const long om = g.at(idx[0], idx[1], idx[2]);
        idx[a] = P.irup + 1;
        const long op = g.at(idx[0], idx[1], idx[2]);
        T sa2 = T(0);
        for (int i = 0; i < 3; i++) {
            F.sa[i][fq] = P.a[i][op] * P.mr[op] - P.a[i][om] * P.mr[om];
            sa2 = sa2 + F.sa[i][fq] * F.sa[i][fq];
        }
        F.sam[fq] = sqrt(sa2);
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline dim3 box_grid(int3 lo, int3 hi, dim3 blk) {
    return dim3((hi.x - lo.x + blk.x) / blk.x, (hi.y - lo.y + blk.y) / blk.y, hi.z - lo.z + 1);
}

// Finite kinematic source (sord/kinematic_source.f90:67-118), one thread per
// subfault: the history value at tm (linear between samples), spread over the
// 2 x 2 x 2 cells or nodes around the subfault with trilinear weights, added
// to the step's source box (potency, moment: minus the tensor; force with
// box set) or to the force a.  Subfaults that share a cell add atomically.
template <class T>
__global__ __launch_bounds__(256) void k_finite_source(FSrc<T> S, T tm, Params<T> P, int box) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= S.n) return;
    const T t0 = S.t0[q], dt = S.dt[q];
    int i = (int)floor((tm - t0) / dt) + 1;
    T xi[3];
    int i1[3], i2[3];
    bool ok = i >= 0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        xi[a] = S.xi[a][q];
        const int fl = (int)floor(xi[a]);
        i1[a] = max(S.lo[a], fl);
        i2[a] = min(S.hi[a], fl + 1);
        if (i2[a] < i1[a]) ok = false;
    }
    if (!ok) return;
    i = min(i, S.nt[q] - 1);
    const T t = t0 + dt * (T)(i - 1);
    T h = min(T(1), (tm - t) / dt);
    const T *hs = S.hist + S.off[q];
    h = i == 0 ? h * hs[0] : (T(1) - h) * hs[i - 1] + h * hs[i];
    for (int l = i1[2]; l <= i2[2]; l++)
        for (int k = i1[1]; k <= i2[1]; k++)
            for (int j = i1[0]; j <= i2[0]; j++) {
                const T w = h * ((T(1) - fabs(xi[0] - (T)j)) * (T(1) - fabs(xi[1] - (T)k)) *
                                 (T(1) - fabs(xi[2] - (T)l)));
                const long o = (j - P.fs_lo[0]) + (long)P.fs_n[0] * ((k - P.fs_lo[1]) + (long)P.fs_n[1] * (l - P.fs_lo[2]));
                if (S.kind == 3 && !box) {
                    const long on = P.g.at(j, k, l);
                    for (int c = 0; c < 3; c++) atomicAdd(&P.a[c][on], w * S.w[c][q]);
                } else if (S.kind == 3) {
                    for (int c = 0; c < 3; c++) atomicAdd(&P.fs[c][o], w * S.w[c][q]);
                } else {
                    for (int c = 0; c < 6; c++) atomicAdd(&P.fs[c][o], -(w * S.w[c][q]));
                }
            }
}

template <class T>
void launch_finite_source(const FSrc<T> &S, T tm, const Params<T> &P, bool box, hipStream_t s) {
    if (S.n > 0) hipLaunchKernelGGL((k_finite_source<T>), dim3((S.n + 255) / 256), dim3(256), 0, s, S, tm, P, (int)box);
}

template <class T>
void launch_stress(const Params<T> &P, const PointOps<T> &ops, int3 lo, int3 hi, int mode, hipStream_t s) {
    if (lo.x > hi.x || lo.y > hi.y || lo.z > hi.z) return;
    dim3 blk(64, 4, 1);
    dim3 grd = box_grid(lo, hi, blk);
    if (mode == 0) hipLaunchKernelGGL((k_stress<T, 0>), grd, blk, 0, s, P, ops, lo, hi);
    if (mode == 1) hipLaunchKernelGGL((k_stress<T, 1>), grd, blk, 0, s, P, ops, lo, hi);
    if (mode == 2) hipLaunchKernelGGL((k_stress<T, 2>), grd, blk, 0, s, P, ops, lo, hi);
    if (mode == 3) hipLaunchKernelGGL((k_stress<T, 3>), grd, blk, 0, s, P, ops, lo, hi);
}

template <class T>
void launch_accel(const Params<T> &P, int3 lo, int3 hi, int mode, hipStream_t s) {
    if (lo.x > hi.x || lo.y > hi.y || lo.z > hi.z) return;
    dim3 blk(64, 4, 1);
    dim3 grd = box_grid(lo, hi, blk);
    if (mode == 0) hipLaunchKernelGGL((k_accel<T, 0>), grd, blk, 0, s, P, lo, hi);
    if (mode == 1) hipLaunchKernelGGL((k_accel<T, 1>), grd, blk, 0, s, P, lo, hi);
    if (mode == 2) hipLaunchKernelGGL((k_accel<T, 2>), grd, blk, 0, s, P, lo, hi);
}

template <class T>
void launch_time(const Params<T> &P, int3 lo, int3 hi, hipStream_t s) {
    dim3 blk(64, 4, 1);
    hipLaunchKernelGGL((k_time<T>), box_grid(lo, hi, blk), blk, 0, s, P, lo, hi);
}
template <class T>
void launch_newton(const Params<T> &P, int3 lo, int3 hi, hipStream_t s) {
    dim3 blk(64, 4, 1);
    hipLaunchKernelGGL((k_newton<T>), box_grid(lo, hi, blk), blk, 0, s, P, lo, hi);
}

static inline int nblocks(long n) {
    long b = (n + 255) / 256;
    if (b > 4096) b = 4096;
    return (int)(b < 1 ? 1 : b);
}

template <class T>
void launch_gather(const T *f, View g, const int lo[3], const int hi[3], const int st[3], const int n[3], T *out,
                   hipStream_t s) {
    Box b;
    for (int i = 0; i < 3; i++) b.lo[i] = lo[i], b.hi[i] = hi[i], b.st[i] = st[i], b.n[i] = n[i];
    const long N = (long)n[0] * n[1] * n[2];
    hipLaunchKernelGGL((k_gather<T>), dim3(nblocks(N)), dim3(256), 0, s, f, g, b, out);
}
template <class T>
void launch_boxop(T *f, View g, const int lo[3], const int hi[3], const int st[3], const int n[3], int op, T val,
                  const T *src, hipStream_t s) {
    Box b;
    for (int i = 0; i < 3; i++) b.lo[i] = lo[i], b.hi[i] = hi[i], b.st[i] = st[i], b.n[i] = n[i];
    const long N = (long)n[0] * n[1] * n[2];
    hipLaunchKernelGGL((k_boxop<T>), dim3(nblocks(N)), dim3(256), 0, s, f, g, b, op, val, src);
}
template <class T>
void launch_plane_copy(T *f, View g, const int n[3], int ax, int dst, int src, hipStream_t s) {
    const int b = (ax + 1) % 3, c = (ax + 2) % 3;
    hipLaunchKernelGGL((k_plane_copy<T>), dim3(nblocks((long)n[b] * n[c])), dim3(256), 0, s, f, g, n[0], n[1], n[2],
                       ax, dst, src);
}
template <class T>
void launch_sample(const T *f, View g, const Samples<T> &smp, T *out, hipStream_t s) {
    if (smp.n <= 0) return;
    hipLaunchKernelGGL((k_sample<T>), dim3(1), dim3(64), 0, s, f, g, smp, out);
}
template <class T>
void launch_point_add(T *f, View g, const int lo[3], const int hi[3], const T xi[3], T val, hipStream_t s) {
    hipLaunchKernelGGL((k_point_add<T>), dim3(1), dim3(64), 0, s, f, g, make_int3(lo[0], lo[1], lo[2]),
                       make_int3(hi[0], hi[1], hi[2]), xi[0], xi[1], xi[2], val);
}
template <class T>
void launch_norm(T *out, T *const a[3], T *const b[3], const Grid<T> &g, const int lo[3], const int hi[3],
                 const int st[3], const int n[3], hipStream_t s) {
    Box bx;
    for (int i = 0; i < 3; i++) bx.lo[i] = lo[i], bx.hi[i] = hi[i], bx.st[i] = st[i], bx.n[i] = n[i];
    const long N = (long)n[0] * n[1] * n[2];
    hipLaunchKernelGGL((k_norm<T>), dim3(nblocks(N)), dim3(256), 0, s, out, a[0], a[1], a[2],
                       b ? b[0] : nullptr, b ? b[1] : nullptr, b ? b[2] : nullptr, g, bx);
}
template <class T>
int launch_maxnorm(T *const a[3], T *const b[3], const Grid<T> &g, const int lo[3], const int hi[3], T *pval,
                   long *ploc, hipStream_t s) {
    Box bx;
    for (int i = 0; i < 3; i++) bx.lo[i] = lo[i], bx.hi[i] = hi[i], bx.st[i] = 1, bx.n[i] = hi[i] - lo[i] + 1;
    const long N = (long)bx.n[0] * bx.n[1] * bx.n[2];
    int nb = nblocks(N);
    if (nb > MAXNORM_BLOCKS) nb = MAXNORM_BLOCKS;
    hipLaunchKernelGGL((k_maxnorm<T>), dim3(nb), dim3(256), 0, s, a[0], a[1], a[2], b ? b[0] : nullptr,
                       b ? b[1] : nullptr, b ? b[2] : nullptr, g, bx, pval, ploc);
    return nb;
}
template <class T>
void launch_vbc_plane(const Params<T> &P, int ax, int dst, int src, int mask, const T sg[3], int kind,
                      hipStream_t s) {
    const int b = (ax + 1) % 3, c = (ax + 2) % 3;
    const long N = (long)P.g.nm[b] * P.g.nm[c];
    hipLaunchKernelGGL((k_vbc_plane<T>), dim3(nblocks(N)), dim3(256), 0, s, P, ax, dst, src, mask, sg[0], sg[1],
                       sg[2], kind);
}
template <class T>
void launch_rupture(const Params<T> &P, const Fault<T> &F, int it, T tm, int fused, hipStream_t s) {
    const int a = P.ifn - 1, b = (a + 1) % 3, c = (a + 2) % 3;
    const long N = (long)P.g.nm[b] * P.g.nm[c];
    hipLaunchKernelGGL((k_rupture<T>), dim3(nblocks(N)), dim3(256), 0, s, P, F, it, tm, fused);
}
template <class T>
void launch_fault_time(const Params<T> &P, const Fault<T> &F, T tm, T svtol, hipStream_t s) {
    const int a = P.ifn - 1, b = (a + 1) % 3, c = (a + 2) % 3;
    const long N = (long)P.g.nm[b] * P.g.nm[c];
    hipLaunchKernelGGL((k_fault_time<T>), dim3(nblocks(N)), dim3(256), 0, s, P, F, tm, svtol);
}
template <class T>
void launch_fault_svold(const Params<T> &P, const Fault<T> &F, hipStream_t s) {
    const int a = P.ifn - 1, b = (a + 1) % 3, c = (a + 2) % 3;
    const long N = (long)P.g.nm[b] * P.g.nm[c];
    hipLaunchKernelGGL((k_fault_svold<T>), dim3(nblocks(N)), dim3(256), 0, s, P, F);
}
template <class T>
void launch_fault_sa(const Params<T> &P, const Fault<T> &F, hipStream_t s) {
    const int a = P.ifn - 1, b = (a + 1) % 3, c = (a + 2) % 3;
    const long N = (long)P.g.nm[b] * P.g.nm[c];
    hipLaunchKernelGGL((k_fault_sa<T>), dim3(nblocks(N)), dim3(256), 0, s, P, F);
}

#define INST(T)                                                                                            \
    template void launch_stress<T>(const Params<T> &, const PointOps<T> &, int3, int3, int, hipStream_t);   \
    template void launch_finite_source<T>(const FSrc<T> &, T, const Params<T> &, bool, hipStream_t);       \
    template void launch_accel<T>(const Params<T> &, int3, int3, int, hipStream_t);                         \
    template void launch_time<T>(const Params<T> &, int3, int3, hipStream_t);                               \
    template void launch_newton<T>(const Params<T> &, int3, int3, hipStream_t);                             \
    template void launch_gather<T>(const T *, View, const int *, const int *, const int *, const int *, T *, \
                                   hipStream_t);                                                            \
    template void launch_boxop<T>(T *, View, const int *, const int *, const int *, const int *, int, T,     \
                                  const T *, hipStream_t);                                                  \
    template void launch_plane_copy<T>(T *, View, const int *, int, int, int, hipStream_t);                  \
    template void launch_sample<T>(const T *, View, const Samples<T> &, T *, hipStream_t);                   \
    template void launch_point_add<T>(T *, View, const int *, const int *, const T *, T, hipStream_t);       \
    template void launch_norm<T>(T *, T *const *, T *const *, const Grid<T> &, const int *, const int *,     \
                                 const int *, const int *, hipStream_t);                                    \
    template int launch_maxnorm<T>(T *const *, T *const *, const Grid<T> &, const int *, const int *, T *,   \
                                   long *, hipStream_t);                                                    \
    template void launch_vbc_plane<T>(const Params<T> &, int, int, int, int, const T *, int, hipStream_t);  \
    template void launch_rupture<T>(const Params<T> &, const Fault<T> &, int, T, int, hipStream_t);          \
    template void launch_fault_time<T>(const Params<T> &, const Fault<T> &, T, T, hipStream_t);              \
    template void launch_fault_svold<T>(const Params<T> &, const Fault<T> &, hipStream_t);                   \
    template void launch_fault_sa<T>(const Params<T> &, const Fault<T> &, hipStream_t);

INST(float)
INST(double)

// ---------------------------------------------------------------------------
// calibration copy: one 4-byte word per lane per access, the access width of
// the solver kernels (the HBM counters count 128-byte requests differently)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_copy_words(const unsigned *src, unsigned *dst, long n) {
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x)
        dst[q] = src[q];
}
void launch_copy_words(const void *src, void *dst, long nwords, hipStream_t s) {
    if (nwords <= 0) return;
    const long blocks = std::min<long>((nwords + 255) / 256, 65536);
    hipLaunchKernelGGL(k_copy_words, dim3((unsigned)blocks), dim3(256), 0, s, (const unsigned *)src, (unsigned *)dst,
                       nwords);
}

// ---------------------------------------------------------------------------
// Scratch of the allocation-free step (sord_step_static): the statistics
// slots and the partial maxima of the max-norm reduction, in a file-scope
// device array of SCRATCH_BYTES that every device holds from the library's
// load until the process ends.  Host code never passes the array's name to a
// kernel or a copy (on the host it is a stand-in, not a device address): it
// uses the address hipGetSymbolAddress gives, taken once per device with that
// device current, at the handle's initialisation.
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

void *static_scratch(int device) {
    static std::mutex mu;
    static void *addr[64] = {};
    if (device < 0 || device >= 64) return nullptr;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[device]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
            (void)hipGetLastError();
            return nullptr;
        }
        addr[device] = p;
    }
    return addr[device];
}

}  // namespace sord
