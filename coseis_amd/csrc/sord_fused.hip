// Fused single-pass SORD step (CDNA4, gfx950).
//
// One launch advances a set of tiles by one time step: strain
// (sord/stress.f90:16-165 with sord/diff_nc_op.f90), Hooke's law and 1/V
// (sord/stress.f90:192-238), divergence (sord/acceleration.f90:280-388 with
// sord/diff_cn_op.f90), hourglass stiffness and viscosity
// (sord/acceleration.f90:390-469 with sord/hourglass_op.f90), plane-local
// boundary conditions (sord/boundary_cond.f90:44-66), Newton
// (sord/acceleration.f90:498-509) and the leapfrog of the next step
// (sord/time_integration.f90:26-41), on a mesh with uniform spacing per axis.
//
// Two instantiations share the decomposition:
//   k_fused   interior tiles: every cell elastic, every node interior
//   k_fusedp  the rest: PML slabs (split-field memory g on cells, p on
//             nodes, viscous-only hourglass at PML nodes), boundary planes
//             (zeroed halo cells, boundary conditions) and the two fault
//             planes (force stored for the rupture kernel, no update)
//
// Decomposition ("z-march"): a block is W wavefronts; wave w owns node row
// k = kb - 1 + w, lane i owns node column j = jb - 1 + i, and the block marches
// the planes l of its work item.  Lanes 0 and 63 and waves 0 and W-1 are a
// one-node halo: they load and compute cells but update no node.  Per plane
// every lane computes the one cell whose lower corner is its node.  A cell's
// contribution to its 8 corner nodes is a polynomial in the corner signs
// (s_j, s_k, s_l) = +-1 (+1 for the lower corner):
//     A_j s_j + A_k s_k + A_l s_l + B_jk s_j s_k + B_kl s_k s_l + B_jl s_j s_l + B_jkl s_j s_k s_l
// with A the divergence weights times the stress row and B minus the
// hourglass modes, so the gather over the 8 cells of a node factorises: the
// plane l-1 part is carried in registers, the row k-1 part comes through LDS,
// the column j-1 part through a lane shuffle.  u, v, gamma, 1/m and
// lambda, mu, yy are read once and u, v written once per node: 68 bytes per
// node-update in FP32 (plus the PML memory in the slabs).
#include <hip/hip_ext.h>

#include "sord_dev.h"
#include "sord_launch.h"

namespace sord {

template <class T>
__device__ __forceinline__ void atomic_max_bits(T *slot, T v);
template <>
__device__ __forceinline__ void atomic_max_bits<float>(float *slot, float v) {
    // v >= 0 or NaN: the magnitude bits order like the values, NaN above +inf
    atomicMax((unsigned int *)slot, __float_as_uint(v) & 0x7fffffffu);
}
template <>
__device__ __forceinline__ void atomic_max_bits<double>(double *slot, double v) {
    atomicMax((unsigned long long *)slot, (unsigned long long)__double_as_longlong(v) & 0x7fffffffffffffffull);
}

// value of lane + 1 / lane - 1 of the wavefront (DPP wave_shl:1 / wave_shr:1;
// the lane without a source reads 0); a VALU operand modifier instead of a
// trip through the LDS crossbar
__device__ __forceinline__ int dpp_next(int x) { return __builtin_amdgcn_mov_dpp(x, 0x130, 0xF, 0xF, true); }
__device__ __forceinline__ int dpp_prev(int x) { return __builtin_amdgcn_mov_dpp(x, 0x138, 0xF, 0xF, true); }
__device__ __forceinline__ float lane_next(float x) { return __int_as_float(dpp_next(__float_as_int(x))); }
__device__ __forceinline__ float lane_prev(float x) { return __int_as_float(dpp_prev(__float_as_int(x))); }
__device__ __forceinline__ double lane_next(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = dpp_next((int)b), hi = dpp_next((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double lane_prev(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = dpp_prev((int)b), hi = dpp_prev((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// a global-memory pointer held in scalar registers: selecting between two
// such pointers per lane stays a register select (a select between two
// argument fields would become a per-lane load of the pointer, consumed at
// once, which waits for every load in flight); typed as global memory so that
// its accesses are global (not flat) loads and stores
template <class T>
using gptr = __attribute__((address_space(1))) T *;
// the same for an integer argument field (selected per lane between two)
__device__ __forceinline__ int uniform_int(int x) { return __builtin_amdgcn_readfirstlane(x); }
template <class T>
__device__ __forceinline__ gptr<T> uniform_ptr(T *p) {
    const unsigned long long v = (unsigned long long)p;
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    return (gptr<T>)(T *)(((unsigned long long)hi << 32) | lo);
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
__device__ __forceinline__ T tmax(T a, T b) {
    return (b > a || b != b) ? b : a;
}

// the statistics maxima of a block (slots: 0 acceleration, 1 velocity, 2
// displacement, 3 strain; 0 and 3 on a step that takes them now, 1 and 2 on
// the step before): each wave's maxima through LDS to the first wave, then
// one atomic per slot and block (the interior kernel).  Atomics of every
// wave on the same four words queue in one L2 channel (16-wave blocks on
// 200^3 nodes: the step with statistics 46 % slower than the step without).
// Every wave of the block calls it.
template <class T, int W>
__device__ __forceinline__ void block_stat_max(const T (&m)[4], bool now, bool next, T *cur, T *nxt, int w,
                                               int lane) {
    __shared__ T sM[4][W];
    if (!now && !next) return;  // uniform over the block
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const bool use = (q == 0 || q == 3) ? now : next;
        if (use) {
            const T x = wave_max(m[q]);
            if (lane == 0) sM[q][w] = x;
        }
    }
    __syncthreads();
    if (w == 0) {
        T r[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const bool use = (q == 0 || q == 3) ? now : next;
            r[q] = use ? wave_max(lane < W ? sM[q][lane] : T(0)) : T(0);
        }
        if (lane == 0) {
            if (now) atomic_max_bits<T>(cur + 0, r[0]), atomic_max_bits<T>(cur + 3, r[3]);
            if (next) atomic_max_bits<T>(nxt + 1, r[1]), atomic_max_bits<T>(nxt + 2, r[2]);
        }
    }
}


// ---------------------------------------------------------------------------
// shared pieces
// ---------------------------------------------------------------------------

// source ops on one cell in the reference's units: volume-integrated strain
// (potency), moment, stress; G is the true gradient, returns the stress.  fv:
// the finite source's values of this cell, added to the potency (fk 1) or the
// moment (fk 2) before that stage's ops, as sord/stress.f90:170-173,215-218
// call it.
template <class T>
__device__ __forceinline__ void cell_ops(const PointOps<T> &ops, unsigned boxes, long o, const T G[3][3], T lm,
                                         T m, T vol, T rvol, int fk, const T fv[6], T s[6]) {
    T w1[3], w2[3];
#pragma unroll
    for (int i = 0; i < 3; i++) w1[i] = G[i][i] * vol;
    w2[0] = T(0.5) * (G[2][1] * vol) + T(0.5) * (G[1][2] * vol);
    w2[1] = T(0.5) * (G[0][2] * vol) + T(0.5) * (G[2][0] * vol);
    w2[2] = T(0.5) * (G[1][0] * vol) + T(0.5) * (G[0][1] * vol);
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
    auto add_fs = [&](int kind) {
        if (fk != kind) return;
#pragma unroll
        for (int i = 0; i < 3; i++) w1[i] = w1[i] + fv[i], w2[i] = w2[i] + fv[3 + i];
    };
    add_fs(1);
    apply(0);
    const T tr = lm * (w1[0] + w1[1] + w1[2]);
#pragma unroll
    for (int i = 0; i < 3; i++) {
        w1[i] = T(2) * m * w1[i] + tr;
        w2[i] = T(2) * m * w2[i];
    }
    add_fs(2);
    apply(1);
#pragma unroll
    for (int i = 0; i < 3; i++) {
        w1[i] = w1[i] * rvol;
        w2[i] = w2[i] * rvol;
    }
    apply(2);
    s[0] = w1[0], s[1] = w1[1], s[2] = w1[2], s[3] = w2[0], s[4] = w2[1], s[5] = w2[2];
}

// Finite source box (cells of the potency or moment, nodes of the force):
// whether a tile's cell rows [j0, j1] x [k0, k1] (its node rows for the force)
// meet it, uniform over the block, and the box index of one cell or node
// (-1 outside).
struct FsTile {
    bool cells, nodes;
};
template <class T>
__device__ __forceinline__ FsTile fs_tile(const Params<T> &P, int j0, int j1, int k0, int k1) {
    FsTile t{false, false};
    if (!P.fs_kind || P.fs_n[0] <= 0) return t;
    const bool fn = P.fs_kind == 3;
    const int jn0 = fn ? j0 + 1 : j0, kn0 = fn ? k0 + 1 : k0;  // owned nodes start one row in
    const bool meet = jn0 < P.fs_lo[0] + P.fs_n[0] && j1 >= P.fs_lo[0] && kn0 < P.fs_lo[1] + P.fs_n[1] &&
                      k1 >= P.fs_lo[1];
    t.cells = meet && !fn;
    t.nodes = meet && fn;
    return t;
}
template <class T>
__device__ __forceinline__ long fs_index(const Params<T> &P, int j, int k, int l) {
    const unsigned dj = j - P.fs_lo[0], dk = k - P.fs_lo[1], dl = l - P.fs_lo[2];
    if (dj >= (unsigned)P.fs_n[0] || dk >= (unsigned)P.fs_n[1] || dl >= (unsigned)P.fs_n[2]) return -1;
    return (long)dj + (long)P.fs_n[0] * ((long)dk + (long)P.fs_n[1] * dl);
}
// the finite force on node (j, k, l), added before the boundary conditions as
// sord/acceleration.f90:213-216 adds it
template <class T>
__device__ __forceinline__ void fs_force(const Params<T> &P, int j, int k, int l, T f[3]) {
    const long q = fs_index(P, j, k, l);
    if (q < 0) return;
#pragma unroll
    for (int c = 0; c < 3; c++) f[c] = f[c] + P.fs[c][q];
}

// Sums over the four nodes (j, k), (j+1, k), (j, k+1), (j+1, k+1) of one node
// plane, from which a cell's differences and hourglass modes follow: each
// plane's sums serve the cell below it and the cell above it.  With f_q the
// node values (q = dj + 2 dk):
//   J = (f1 - f0) + (f3 - f2)   K = (f2 - f0) + (f3 - f1)
//   S = f0 + f1 + f2 + f3       X = f0 - f1 - f2 + f3
// A cell between planes lo and hi has the differences along j, k, l
// J_lo + J_hi, K_lo + K_hi, S_hi - S_lo, and the hourglass modes
// (s_k s_l, s_j s_l, s_j s_k, s_j s_k s_l) K_hi - K_lo, J_hi - J_lo,
// X_lo + X_hi, X_lo - X_hi (s = +1 at the lower corner).
template <class T>
struct PlaneSums {
    T w[3][3];  // u + gamma v: J, K, S (strain)
    T F[3][3];  // h1 u + dt h2 v: J, K, X (full hourglass)
    T v[3][4];  // v: J, K, S (strain in a damped direction), X (viscous hourglass)
    T u[3][3];  // u: J, K, S (strain in the other directions of a PML slab)
};

// ---------------------------------------------------------------------------
// interior tiles
// ---------------------------------------------------------------------------
// DEPTH: planes of node data in flight ahead of the plane being computed (1:
// two register sets alternate; 2: three rotate, for 8-wave blocks that run
// one block per CU, whose waves have the register file of two)
template <class T, int W, bool FS, bool ST, int DEPTH = 1>
__device__ __forceinline__ void interior_item(const Params<T> &P, const FusedC<T> &C, const PointOps<T> &ops,
                                              const Work &wk) {
    __shared__ T sN[6][W][64];
    __shared__ T sC[6][W][64];
    const int lane = threadIdx.x, w = threadIdx.y;
    const int j = wk.jb - 1 + lane, k = wk.kb - 1 + w;
    const bool own = lane >= 1 && lane <= 62 && w >= 1 && w <= W - 2 && j <= wk.je && k <= wk.ke;
    const bool cellv = lane <= 62 && w <= W - 2 && j <= wk.je && k <= wk.ke;
    const Grid<T> &g = P.g;
    const int jc = min(max(j, 1), g.nm[0]), kc = min(max(k, 1), g.nm[1]);
    // ST: a statistics step (its maxima add registers and work; the other
    // steps run an instantiation without them)
    const bool stat_now = ST && P.stat_now, stat_next = ST && P.stat_next;
    const unsigned tile_boxes = ops.n ? ops_boxes_jk(ops, wk.jb - 1, wk.je, wk.kb - 1, wk.ke) : 0u;
    const FsTile fst = FS ? fs_tile(P, wk.jb - 1, wk.je, wk.kb - 1, wk.ke) : FsTile{false, false};
    // statistics maxima of each lane (|a|^2, |v|^2, |u|^2, |w|^2)
    T stm[4] = {T(0), T(0), T(0), T(0)};
    auto st_max = [&](int q, T x) {
        if (ST) stm[q] = tmax(stm[q], x);
    };

    // node data: W = u + gamma v (strain), F = h1 u + dt h2 v (hourglass)
    T uo[3], vo[3];

    // raw loads of one node plane (u, v, gamma), issued one iteration before use
    const int lmax = g.nm[2];
    auto fetch_node = [&](int l, T uu[3], T vv[3], T &ga) {
        const long o = g.at(jc, kc, min(l, lmax));
#pragma unroll
        for (int i = 0; i < 3; i++) uu[i] = P.u[i][o], vv[i] = P.v[i][o];
        ga = P.gam[o];
    };
    auto node_vals = [&](const T uu[3], const T vv[3], T ga, T own6[6]) {
#pragma unroll
        for (int i = 0; i < 3; i++) {
            own6[i] = uu[i] + ga * vv[i];
            own6[3 + i] = C.h1 * uu[i] + C.dth2 * vv[i];
        }
    };
    auto fetch_cell = [&](int l, T mat[4]) {  // lambda, mu, yy of cell l and 1/m of node l
        const long o = g.at(jc, kc, min(l, lmax));
        mat[0] = P.lam[o], mat[1] = P.mu[o], mat[2] = P.yy[o], mat[3] = P.mr[o];
    };
    // the plane sums of W (J, K, S) and F (J, K, X): the row above through
    // LDS, the next column through a lane shift of the row pair's sum and
    // difference
    auto exchange = [&](const T own6[6], PlaneSums<T> &S) {
#pragma unroll
        for (int q = 0; q < 6; q++) sN[q][w][lane] = own6[q];
        __syncthreads();
        const int wa = w + 1 < W ? w + 1 : w;
#pragma unroll
        for (int q = 0; q < 6; q++) {
            // the own value's shift does not wait for the LDS read
            const T ab = sN[q][wa][lane], on = lane_next(own6[q]), an = lane_next(ab);
            const T e = own6[q] + ab, d = ab - own6[q], ne = on + an, nd = an - on;
            if (q < 3) S.w[q][0] = ne - e, S.w[q][1] = d + nd, S.w[q][2] = e + ne;
            else S.F[q - 3][0] = ne - e, S.F[q - 3][1] = d + nd, S.F[q - 3][2] = nd - d;
        }
    };
    // stress and hourglass of cell (j, k, l) from the plane sums of planes l
    // (Slo) and l+1 (Shi); E: polynomial at s_l = +1, Cc: at s_l = -1
    // st: the cell counts in the statistics (not the spare plane past the item's
    // last, which may lie where another kernel's damping applies)
    auto cell_at = [&](int l, const T mat[4], const PlaneSums<T> &Slo, const PlaneSums<T> &Shi, T E[3][4],
                       T Cc[3][4], bool st) {
        const T lm = mat[0], m = mat[1], y = mat[2];
        T G[3][3];  // G[a][c]: gradient along a of component c
#pragma unroll
        for (int c = 0; c < 3; c++) {
            G[0][c] = C.gx * (Slo.w[c][0] + Shi.w[c][0]);
            G[1][c] = C.gy * (Slo.w[c][1] + Shi.w[c][1]);
            G[2][c] = C.gz * (Shi.w[c][2] - Slo.w[c][2]);
        }
        T s[6];
        // source ops: the boxes met by this tile, then those of plane l (both
        // uniform over the block), then those holding this cell
        unsigned bx = 0;
        if (tile_boxes) {
            const unsigned cl = ops_boxes_l(ops, tile_boxes, l);
            if (cl && cellv) bx = ops_boxes(ops, cl, j, k, l);
        }
        // finite source (potency, moment) on this cell
        const long fq = fst.cells && cellv ? fs_index(P, j, k, l) : -1;
        if (!bx && fq < 0) {
            const T tr = G[0][0] + G[1][1] + G[2][2];
            const T m2 = m + m;
            s[0] = m2 * G[0][0] + lm * tr;
            s[1] = m2 * G[1][1] + lm * tr;
            s[2] = m2 * G[2][2] + lm * tr;
            s[3] = m * (G[2][1] + G[1][2]);
            s[4] = m * (G[0][2] + G[2][0]);
            s[5] = m * (G[1][0] + G[0][1]);
        } else {
            T fv[6];
#pragma unroll
            for (int c = 0; c < 6; c++) fv[c] = fq >= 0 ? P.fs[c][fq] : T(0);
            cell_ops<T>(ops, bx, g.at(j, k, l), G, lm, m, C.vol, C.rvol, fq >= 0 ? P.fs_kind : 0, fv, s);
        }
        if (stat_now && cellv && st)
            st_max(3, (s[0] * s[0] + s[1] * s[1] + s[2] * s[2]) + T(2) * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]));
        const T A[3][3] = {{C.ax * s[0], C.ay * s[5], C.az * s[4]},
                           {C.ax * s[5], C.ay * s[1], C.az * s[3]},
                           {C.ax * s[4], C.ay * s[3], C.az * s[2]}};
#pragma unroll
        for (int c = 0; c < 3; c++) {
            // modes (s_k s_l, s_j s_l, s_j s_k, s_j s_k s_l) of F
            T h[4];
            h[0] = Shi.F[c][1] - Slo.F[c][1];
            h[1] = Shi.F[c][0] - Slo.F[c][0];
            h[2] = Slo.F[c][2] + Shi.F[c][2];
            h[3] = Slo.F[c][2] - Shi.F[c][2];
#pragma unroll
            for (int q = 0; q < 4; q++) h[q] = y * h[q];
            E[c][0] = A[c][2];
            E[c][1] = A[c][0] - h[1];
            E[c][2] = A[c][1] - h[0];
            E[c][3] = -h[2] - h[3];
            Cc[c][0] = -A[c][2];
            Cc[c][1] = A[c][0] + h[1];
            Cc[c][2] = A[c][1] + h[0];
            Cc[c][3] = -h[2] + h[3];
        }
    };

    // The plane loop is unrolled with rotating register sets: the loads for
    // plane l+1+DEPTH go to the set that no plane before them reads, so no
    // register holding a load in flight is ever copied (a copy would make the
    // compiler wait for the load at the end of every iteration).
    T own6[6], E[3][4];
    PlaneSums<T> NA, NB;  // plane sums of two planes (lower, upper in turn)
    T CA[3][4], CB[3][4];  // s_l = -1 polynomials of two cells (carried, new in turn)
    T uA[3], vA[3], gA, mA[4], uB[3], vB[3], gB, mB[4];  // loads: plane l+1 raw, cell l + 1/m
    T uC[3], vC[3], gC, mC[4];                            // the third set (DEPTH 2)
    fetch_node(wk.lz0 - 1, uA, vA, gA);
    node_vals(uA, vA, gA, own6);
    exchange(own6, NA);
    fetch_node(wk.lz0, uo, vo, gB);
    fetch_cell(wk.lz0 - 1, mB);
    if constexpr (DEPTH == 2) {
        fetch_node(wk.lz0 + 1, uC, vC, gC);
        fetch_cell(wk.lz0, mC);
        fetch_node(wk.lz0 + 2, uA, vA, gA);
        fetch_cell(wk.lz0 + 1, mA);
    }
    node_vals(uo, vo, gB, own6);
    __syncthreads();  // every wave has read the first plane from LDS
    exchange(own6, NB);
    if constexpr (DEPTH == 1) {
        fetch_node(wk.lz0 + 1, uA, vA, gA);
        fetch_cell(wk.lz0, mA);
    }
    cell_at(wk.lz0 - 1, mB, NA, NB, E, CA, true);
    // every wave has read the second plane from LDS before the first plane of
    // the loop writes its own slots again (a wave that ran ahead would
    // overwrite the row its neighbour below has not read yet)
    __syncthreads();

    // one plane: cu..cm hold plane l+1 and cell l (arrived), nu..nm receive plane
    // l+2 and cell l+1; Nlo holds plane l, Nhi receives plane l+1; carry holds
    // cell l-1's polynomial at s_l = -1, Cn receives cell l's
    auto plane = [&](int l, bool act, T (&cu)[3], T (&cv)[3], T &cg, T (&cm)[4], T (&nu)[3], T (&nv)[3], T &ng,
                     T (&nm)[4], const PlaneSums<T> &Nlo, PlaneSums<T> &Nhi, T (&carry)[3][4], T (&Cn)[3][4]) {
        node_vals(cu, cv, cg, own6);
        fetch_node(l + 1 + DEPTH, nu, nv, ng);
        fetch_cell(l + DEPTH, nm);
        exchange(own6, Nhi);
        cell_at(l, cm, Nlo, Nhi, E, Cn, act);
        T R[3][2];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const T q0 = E[c][0] + carry[c][0], qj = E[c][1] + carry[c][1];
            const T qk = E[c][2] + carry[c][2], qjk = E[c][3] + carry[c][3];
            R[c][0] = q0 + qk;
            R[c][1] = qj + qjk;
            sC[2 * c][w][lane] = q0 - qk;
            sC[2 * c + 1][w][lane] = qj - qjk;
        }
        __syncthreads();
        const int wb = w > 0 ? w - 1 : 0;
        T f[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const T r0 = R[c][0] + sC[2 * c][wb][lane];
            const T rj = R[c][1] + sC[2 * c + 1][wb][lane];
            f[c] = (r0 + rj) + lane_prev(r0 - rj);
        }
        if (own && act) {
            const long o = g.at(j, k, l);
            if (fst.nodes) fs_force(P, j, k, l, f);
            const T mrn = cm[3];
            T a2 = T(0), v2 = T(0), u2 = T(0);
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const T a = f[c] * mrn;
                const T vv = vo[c] + P.dt * a;
                const T uu = uo[c] + P.dt * vv;
                P.vn[c][o] = vv;
                P.un[c][o] = uu;
                a2 = a2 + a * a;
                v2 = v2 + vv * vv;
                u2 = u2 + uu * uu;
            }
            if (stat_now) st_max(0, a2);
            if (stat_next) st_max(1, v2), st_max(2, u2);
        }
#pragma unroll
        for (int c = 0; c < 3; c++) uo[c] = cu[c], vo[c] = cv[c];
    };
    // all planes of a group always run (one loop back-edge keeps the register
    // sets in place); past the last plane of the item they store nothing
    if constexpr (DEPTH == 1) {
        for (int l = wk.lz0; l <= wk.lz1; l += 2) {
            plane(l, true, uA, vA, gA, mA, uB, vB, gB, mB, NB, NA, CA, CB);
            plane(l + 1, l + 1 <= wk.lz1, uB, vB, gB, mB, uA, vA, gA, mA, NA, NB, CB, CA);
        }
    } else {
        // three load sets (C, A, B hold planes l+1, l+2, l+3 in turn) and the
        // two corner and polynomial sets: six planes per back-edge
        for (int l = wk.lz0; l <= wk.lz1; l += 6) {
            plane(l, true, uC, vC, gC, mC, uB, vB, gB, mB, NB, NA, CA, CB);
            plane(l + 1, l + 1 <= wk.lz1, uA, vA, gA, mA, uC, vC, gC, mC, NA, NB, CB, CA);
            plane(l + 2, l + 2 <= wk.lz1, uB, vB, gB, mB, uA, vA, gA, mA, NB, NA, CA, CB);
            plane(l + 3, l + 3 <= wk.lz1, uC, vC, gC, mC, uB, vB, gB, mB, NA, NB, CB, CA);
            plane(l + 4, l + 4 <= wk.lz1, uA, vA, gA, mA, uC, vC, gC, mC, NB, NA, CA, CB);
            plane(l + 5, l + 5 <= wk.lz1, uB, vB, gB, mB, uA, vA, gA, mA, NA, NB, CB, CA);
        }
    }
    if (ST) block_stat_max<T, W>(stm, stat_now, stat_next, P.stat_cur, P.stat_nxt, w, lane);
}

// one work item per block; the plain kernel gets one scenario in its
// arguments, the batch kernel looks its item's scenario up (Work.pad1).  The
// steps without statistics (9 in 10 by default) run instantiations without
// the maxima: their four accumulators pushed the FP32 kernel past its 128
// VGPRs, and the spill reloads in the plane loop waited for every load in
// flight, the next plane's prefetch included (C2: 0.482 against 0.427 ms per
// interior launch).
#ifndef SORD_INTERIOR_DEPTH
#define SORD_INTERIOR_DEPTH 1
#endif
// prefetch depth and waves per SIMD of the plain interior kernels: FP32
// 8-wave blocks (one per CU) may take the deeper prefetch and the register
// file of two waves per SIMD; 16-wave blocks need four
template <class T, int W>
constexpr int interior_depth() { return sizeof(T) == 4 && W == 8 ? SORD_INTERIOR_DEPTH : 1; }
template <class T, int W>
constexpr int interior_waves_per_eu() { return sizeof(T) == 4 ? (interior_depth<T, W>() == 2 ? 2 : 4) : 1; }

template <class T, int W>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(interior_waves_per_eu<T, W>())))
void k_fused(Params<T> P, FusedC<T> C, PointOps<T> ops) {
    const Work wk = C.work[blockIdx.x];
    if (wk.lz0 > wk.lz1) return;  // uniform over the block
    interior_item<T, W, false, false, interior_depth<T, W>()>(P, C, ops, wk);
}
// the same on a statistics step
template <class T, int W>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(interior_waves_per_eu<T, W>())))
void k_fused_stat(Params<T> P, FusedC<T> C, PointOps<T> ops) {
    const Work wk = C.work[blockIdx.x];
    if (wk.lz0 > wk.lz1) return;
    interior_item<T, W, false, true, interior_depth<T, W>()>(P, C, ops, wk);
}
// with a finite source: its box enters the stress or the force
template <class T, int W>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : 1)))
void k_fused_fs(Params<T> P, FusedC<T> C, PointOps<T> ops) {
    const Work wk = C.work[blockIdx.x];
    if (wk.lz0 > wk.lz1) return;
    interior_item<T, W, true, true>(P, C, ops, wk);
}
template <class T, int W>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : 1)))
void k_fused_batch(const Scen<T> *__restrict__ sc, const Work *__restrict__ work) {
    const Work wk = work[blockIdx.x];
    if (wk.lz0 > wk.lz1) return;
    const Scen<T> &S = sc[wk.pad1];
    if (!S.active) return;  // uniform over the block
    interior_item<T, W, false, false>(S.P, S.C, S.ops, wk);
}
// a batch step where some merged scenario takes statistics
template <class T, int W>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : 1)))
void k_fused_batch_stat(const Scen<T> *__restrict__ sc, const Work *__restrict__ work) {
    const Work wk = work[blockIdx.x];
    if (wk.lz0 > wk.lz1) return;
    const Scen<T> &S = sc[wk.pad1];
    if (!S.active) return;  // uniform over the block
    interior_item<T, W, false, true>(S.P, S.C, S.ops, wk);
}

// ---------------------------------------------------------------------------
// PML, boundary and fault tiles
// ---------------------------------------------------------------------------

// MASK bit a: the item has real cells of a PML slab of axis a (else no cell or
// node of the item is damped along a, and that axis's PML code folds away).
// Rows: an item of width je-jb+1 runs LW = width+2 lanes per node row (Work.pad0)
// and R = 64/LW rows per wavefront, so narrow items (the x slabs) fill the
// wavefront; node rows of the block are w*R + lane/LW, with halo rows 0 and W*R-1.
// LDS of the boundary kernel: 12 node values and 21 cell coefficients per slot,
// W * 64 + 64 slots (lanes past R*LW use the last 64)
template <int W>
constexpr int boundary_lds_slots() { return 33 * (W * 64 + 64); }


template <class T, int W, int MASK, bool FS>
__device__ __forceinline__ void boundary_item(const Params<T> &P, const FusedC<T> &C, const PointOps<T> &ops,
                                              const Work &wk, T *lds) {
    constexpr int NS = W * 64 + 64;
    // node values exchanged per plane: u + gamma v and h1 u + dt h2 v, and with
    // damping also v and u
    constexpr int NF = MASK ? 12 : 6;
    T (*sN)[NS] = reinterpret_cast<T (*)[NS]>(lds);
    T (*sC)[NS] = reinterpret_cast<T (*)[NS]>(lds + NF * NS);
    const int lane = threadIdx.x, w = threadIdx.y;
    const int LW = wk.pad0, R = 64 / LW, NR = W * R;
    const int jj = lane % LW, rr = lane / LW;
    const bool valid = rr < R;
    const int row = w * R + rr;
    const int t = valid ? w * R * LW + lane : W * 64 + lane;  // LDS slot
    const int tab = valid && row + 1 < NR ? t + LW : t;     // row above
    const int tbe = valid && row > 0 ? t - LW : t;           // row below
    // one damped axis PA: its PML memory comes with the plane's other loads
    constexpr bool PF = MASK == 1 || MASK == 2 || MASK == 4;
    constexpr int PA = MASK == 2 ? 1 : MASK == 4 ? 2 : 0;
    // per plane: lambda, mu, yy, 1/m; g and p of the damped axis (PF); their
    // coefficients when the damped axis is l (along j and k they are constant
    // over the lane's planes)
    constexpr int NMAT = !PF ? 4 : PA == 2 ? 14 : 10;
    constexpr int ZC = NMAT > 10 ? 10 : 0;  // the coefficients along l (unused otherwise)
    const int j = wk.jb - 1 + jj, k = wk.kb - 1 + row;
    const bool ownjk = valid && jj >= 1 && jj <= LW - 2 && row >= 1 && row <= NR - 2 && j <= wk.je && k <= wk.ke;
    const bool cellv = valid && jj <= LW - 2 && row <= NR - 2 && j <= wk.je && k <= wk.ke;
    const Grid<T> &g = P.g;
    const unsigned tile_boxes = ops.n ? ops_boxes_jk(ops, wk.jb - 1, wk.je, wk.kb - 1, wk.ke) : 0u;
    const FsTile fst = FS ? fs_tile(P, wk.jb - 1, wk.je, wk.kb - 1, wk.ke) : FsTile{false, false};
    const int jc = min(max(j, 1), g.nm[0]), kc = min(max(k, 1), g.nm[1]);
    const int lmax = g.nm[2];
    const bool stat_now = P.stat_now, stat_next = P.stat_next;
    T st_a = T(0), st_v = T(0), st_u = T(0), st_w = T(0);
    const int fa = P.ifn - 1;  // fault normal axis (-1: none)
    T uo[3], vo[3];

    // ---- the lane's geometry along j and k, constant over its planes ----
    // cell (j, k, .) inside the mesh and not the fault's zero-width cell; node
    // (j, k, .) inside; node on a fault plane
    const bool cjk = j >= P.i1cell[0] && j <= P.i2cell[0] && k >= P.i1cell[1] && k <= P.i2cell[1] &&
                     !(fa == 0 && j == P.irup) && !(fa == 1 && k == P.irup);
    const bool njk = ownjk && j >= P.i1node[0] && j <= P.i2node[0] && k >= P.i1node[1] && k <= P.i2node[1];
    const bool fjk = (fa == 0 && (j == P.irup || j == P.irup + 1)) || (fa == 1 && (k == P.irup || k == P.irup + 1));
    // force components that the boundary conditions of a node plane zero
    // (sord/boundary_cond.f90:44-66; -1: the two tangential, 1: the normal,
    // 3: all)
    auto bc_zero = [&](int a, int n) -> int {
        int bt = 0;
        if (n == P.i1bc[a]) bt = P.bc1[a] % 10;
        if (n == P.i2bc[a]) bt = P.bc2[a] % 10;
        return bt == -1 ? (7 & ~(1 << a)) : bt == 1 ? 1 << a : bt == 3 ? 7 : 0;
    };
    const int zjk = bc_zero(0, j) | bc_zero(1, k);
    // slab membership along the damped axes j, k: cell (g memory, source
    // fields), node (p memory), node inside the undamped range (full hourglass)
    bool cin[3] = {false, false, false}, nin[3] = {false, false, false};
    bool hint_jk = true;
    // PML memory of the lane's cell and node along j and k: pointers to plane 1,
    // per-plane steps and component strides, coefficients
    gptr<T> lg[2] = {nullptr, nullptr}, lgn[2] = {nullptr, nullptr}, lp[2] = {nullptr, nullptr};
    int lgl[2] = {0, 0}, lgs[2] = {0, 0}, lpl[2] = {0, 0}, lps[2] = {0, 0};
    T lc2[2] = {T(0), T(0)}, lc1[2] = {T(0), T(0)}, ln2[2] = {T(0), T(0)}, ln1[2] = {T(0), T(0)};
#pragma unroll
    for (int a = 0; a < 2; a++) {
        if (!(MASK >> a & 1)) continue;
        const int c = a ? k : j;
        const bool clo = c <= P.i1pml[a];
        cin[a] = clo || c >= P.i2pml[a] - 1;
        const bool nlo = c <= P.i1pml[a];
        nin[a] = nlo || c >= P.i2pml[a];
        if (c < P.i1pml[a] + 1 || c > P.i2pml[a] - 1) hint_jk = false;
        // memory layout (sord/arrays.f90): the slab index along a fastest (a
        // = 0) or the middle index (a = 1)
        if (cin[a] && cjk) {
            const int si = clo ? c - P.i1cell[a] + 1 : P.i2cell[a] - c + 1;
            const int pidx = clo ? c + g.nnoff[a] : g.nn[a] - c - g.nnoff[a];
            const int dep = clo ? uniform_int(P.gn[a][a]) : uniform_int(P.gn[a + 3][a]);
            long e;
            if (a == 0) e = (long)(si - 1) + (long)dep * (k - 1), lgl[a] = dep * g.nm[1];
            else e = (long)(j - 1) + (long)g.nm[0] * (si - 1), lgl[a] = g.nm[0] * dep;
            lgs[a] = lgl[a] * g.nm[2];
            lg[a] = (clo ? uniform_ptr(P.gm[a]) : uniform_ptr(P.gm[a + 3])) + e;
            lgn[a] = (clo ? uniform_ptr(P.gmn[a]) : uniform_ptr(P.gmn[a + 3])) + e;
            lc2[a] = P.dc2[pidx - 1], lc1[a] = P.dc1[pidx - 1];
        }
        if (nin[a] && njk) {
            const int si = nlo ? c - P.i1node[a] + 1 : P.i2node[a] - c + 1;
            const int pidx = nlo ? c + g.nnoff[a] : g.nn[a] - c - g.nnoff[a] + 1;
            const int dep = nlo ? uniform_int(P.pn[a][a]) : uniform_int(P.pn[a + 3][a]);
            long e;
            if (a == 0) e = (long)(si - 1) + (long)dep * (k - 1), lpl[a] = dep * g.nm[1];
            else e = (long)(j - 1) + (long)g.nm[0] * (si - 1), lpl[a] = g.nm[0] * dep;
            lps[a] = lpl[a] * g.nm[2];
            lp[a] = (nlo ? uniform_ptr(P.p[a]) : uniform_ptr(P.p[a + 3])) + e;
            ln2[a] = P.dn2[pidx - 1], ln1[a] = P.dn1[pidx - 1];
        }
    }
    const long ejk = (long)(j - 1) + (long)g.nm[0] * (k - 1);  // (j, k) in a slab of axis l

    // ---- per plane along l (uniform over the block) ----
    // cell l inside the mesh and not the fault cell
    auto creal_l = [&](int l) { return l >= P.i1cell[2] && l <= P.i2cell[2] && !(fa == 2 && l == P.irup); };
    // PML memory of cell l (g) and node l (p) in the slab of axis l: element
    // offset of (j, k), component stride, coefficient index, side
    struct LSlot {
        bool on, lo;
        long e, s;
        int pidx;
    };
    auto gslot_l = [&](int l) -> LSlot {
        LSlot r{false, false, 0, 0, 0};
        if (!(MASK & 4) || !creal_l(l)) return r;
        r.lo = l <= P.i1pml[2];
        if (!r.lo && l < P.i2pml[2] - 1) return r;
        r.on = true;
        const int si = r.lo ? l - P.i1cell[2] + 1 : P.i2cell[2] - l + 1;
        r.pidx = r.lo ? l + g.nnoff[2] : g.nn[2] - l - g.nnoff[2];
        const long dep = r.lo ? P.gn[2][2] : P.gn[5][2];
        r.s = (long)g.nm[0] * g.nm[1] * dep;
        r.e = ejk + (long)g.nm[0] * g.nm[1] * (si - 1);
        return r;
    };
    auto pslot_l = [&](int l) -> LSlot {
        LSlot r{false, false, 0, 0, 0};
        if (!(MASK & 4) || l < P.i1node[2] || l > P.i2node[2]) return r;
        r.lo = l <= P.i1pml[2];
        if (!r.lo && l < P.i2pml[2]) return r;
        r.on = true;
        const int si = r.lo ? l - P.i1node[2] + 1 : P.i2node[2] - l + 1;
        r.pidx = r.lo ? l + g.nnoff[2] : g.nn[2] - l - g.nnoff[2] + 1;
        const long dep = r.lo ? P.pn[2][2] : P.pn[5][2];
        r.s = (long)g.nm[0] * g.nm[1] * dep;
        r.e = ejk + (long)g.nm[0] * g.nm[1] * (si - 1);
        return r;
    };

    // raw loads: node (u, v, gamma) of a plane; cell (lambda, mu, yy) and node 1/m
    auto fetch_node = [&](int l, T (&nd)[7]) {
        const long o = g.at(jc, kc, min(l, lmax));
#pragma unroll
        for (int i = 0; i < 3; i++) nd[i] = P.u[i][o], nd[3 + i] = P.v[i][o];
        nd[6] = P.gam[o];
    };
    // lambda, mu, yy of cell l and 1/m of node l; with one damped axis (PF)
    // also the PML memory of cell l (g, 3 components) and node l (p, 3
    // components), and along l their coefficients, loaded a plane ahead like
    // the rest: a load consumed at once makes the wave wait for every load
    // issued before it, the next plane's prefetch included
    auto fetch_cell = [&](int l, T (&m)[NMAT]) {
        const long o = g.at(jc, kc, min(l, lmax));
        m[0] = P.lam[o], m[1] = P.mu[o], m[2] = P.yy[o], m[3] = P.mr[o];
        if constexpr (PF) {
#pragma unroll
            for (int q = 4; q < NMAT; q++) m[q] = T(0);
            if constexpr (PA < 2) {
                if (cin[PA] && cjk && creal_l(l)) {
                    const long e = (long)lgl[PA] * (l - 1);
#pragma unroll
                    for (int ic = 0; ic < 3; ic++) m[4 + ic] = lg[PA][e + (long)ic * lgs[PA]];
                }
                if (nin[PA] && njk && l >= P.i1node[2] && l <= P.i2node[2]) {
                    const long e = (long)lpl[PA] * (l - 1);
#pragma unroll
                    for (int ic = 0; ic < 3; ic++) m[7 + ic] = lp[PA][e + (long)ic * lps[PA]];
                }
            } else {
                const LSlot gs = gslot_l(l);
                if (gs.on && cjk) {
                    const auto gp = gs.lo ? uniform_ptr(P.gm[2]) : uniform_ptr(P.gm[5]);
#pragma unroll
                    for (int ic = 0; ic < 3; ic++) m[4 + ic] = gp[gs.e + ic * gs.s];
                    m[ZC] = P.dc2[gs.pidx - 1], m[ZC + 1] = P.dc1[gs.pidx - 1];
                }
                const LSlot ps = pslot_l(l);
                if (ps.on && njk) {
                    const auto pp = ps.lo ? uniform_ptr(P.p[2]) : uniform_ptr(P.p[5]);
#pragma unroll
                    for (int ic = 0; ic < 3; ic++) m[7 + ic] = pp[ps.e + ic * ps.s];
                    m[ZC + 2] = P.dn2[ps.pidx - 1], m[ZC + 3] = P.dn1[ps.pidx - 1];
                }
            }
        }
    };
    // the plane sums of a plane from the raw node loads (u 0..2, v 3..5,
    // gamma 6): each node's values once, the row above through LDS, the next
    // column through a lane shift of the row pair's sum and difference
    auto exchange = [&](const T (&nd)[7], PlaneSums<T> &S) {
        T own[NF];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            own[i] = nd[i] + nd[6] * nd[3 + i];
            own[3 + i] = C.h1 * nd[i] + C.dth2 * nd[3 + i];
            if constexpr (MASK != 0) own[6 + i] = nd[3 + i], own[9 + i] = nd[i];
        }
#pragma unroll
        for (int q = 0; q < NF; q++) sN[q][t] = own[q];
        __syncthreads();
#pragma unroll
        for (int q = 0; q < NF; q++) {
            const T ab = sN[q][tab];
            const T e = own[q] + ab, d = ab - own[q];  // rows k + (k+1), (k+1) - k
            const T ne = lane_next(e), nd1 = lane_next(d);
            const T J = ne - e, K = d + nd1, Sm = e + ne, X = nd1 - d;
            const int f = q / 3, c = q % 3;
            if (f == 0) S.w[c][0] = J, S.w[c][1] = K, S.w[c][2] = Sm;
            else if (f == 1) S.F[c][0] = J, S.F[c][1] = K, S.F[c][2] = X;
            else if (f == 2) S.v[c][0] = J, S.v[c][1] = K, S.v[c][2] = Sm, S.v[c][3] = X;
            else S.u[c][0] = J, S.u[c][1] = K, S.u[c][2] = Sm;
        }
    };

    // cell (j, k, l) from the plane sums of planes l (Slo) and l+1 (Shi):
    // coefficients per component at s_l = +1 (E) and -1 (Cc): [0..2] divergence
    // along x, y, z (with the sign of s_l on the z term), [3..5] full hourglass
    // (s_j, s_k, s_j s_k), [6..8] viscous hourglass.  wr: this block owns the
    // cell and writes its PML memory: a cell on a tile edge is also computed by
    // the neighbouring tile, so every block reads the memory of the previous
    // step from gm and only the owner writes the new value to gmn (the host
    // swaps the two after the launch).  st: the cell counts in the statistics
    // (not the spare plane past the item's last, whose damping this kernel's
    // mask may lack).
    auto cell_at = [&](int l, const T (&mat)[NMAT], const PlaneSums<T> &Slo, const PlaneSums<T> &Shi, T (&E)[3][9],
                       T (&Cc)[3][9], bool wr, bool st) {
        const bool real = cjk && creal_l(l);
        const T lm = mat[0], m = mat[1], y = mat[2];
        bool inS[3] = {cin[0], cin[1], false};
        if (MASK & 4) inS[2] = l <= P.i1pml[2] || l >= P.i2pml[2] - 1;
        const T gsp[3] = {C.gx, C.gy, C.gz};
        T G[3][3];  // G[id][ic]
#pragma unroll
        for (int id = 0; id < 3; id++) {
            bool other = false;
#pragma unroll
            for (int a = 0; a < 3; a++)
                if (a != id && (MASK >> a & 1) && inS[a]) other = true;
            const bool useV = (MASK >> id & 1) && inS[id], useU = !useV && other;
#pragma unroll
            for (int ic = 0; ic < 3; ic++) {
                // difference of the source field: v in the damped direction, u
                // in the other directions of a PML slab, u + gamma v elsewhere
                auto diff = [&](const T (&lo)[3], const T (&hi)[3]) { return id == 2 ? hi[2] - lo[2] : lo[id] + hi[id]; };
                const T dw = diff(Slo.w[ic], Shi.w[ic]);
                T d = dw;
                if constexpr (MASK != 0) {
                    const T dv = id == 2 ? Shi.v[ic][2] - Slo.v[ic][2] : Slo.v[ic][id] + Shi.v[ic][id];
                    const T du = diff(Slo.u[ic], Shi.u[ic]);
                    d = useV ? dv : useU ? du : dw;
                }
                G[id][ic] = gsp[id] * d;
            }
            if ((MASK >> id & 1) && inS[id] && real) {
                // split-field memory of the damped direction in the reference's
                // units (volume-integrated strain), sord/stress.f90:42-131
                T c2, c1, old[3];
                gptr<T> gq;
                long e, s;
                if (id < 2) {
                    e = (long)lgl[id] * (l - 1), s = lgs[id];
                    gq = lgn[id] + e;
                    c2 = lc2[id], c1 = lc1[id];
                    if constexpr (PF) {
#pragma unroll
                        for (int ic = 0; ic < 3; ic++) old[ic] = mat[4 + ic];
                    } else {
#pragma unroll
                        for (int ic = 0; ic < 3; ic++) old[ic] = lg[id][e + ic * s];
                    }
                    e = 0;
                } else {
                    const LSlot gs = gslot_l(l);
                    e = gs.e, s = gs.s;
                    gq = gs.lo ? uniform_ptr(P.gmn[2]) : uniform_ptr(P.gmn[5]);
                    if constexpr (PF) {
                        c2 = mat[ZC], c1 = mat[ZC + 1];
#pragma unroll
                        for (int ic = 0; ic < 3; ic++) old[ic] = mat[4 + ic];
                    } else {
                        const auto gp = gs.lo ? uniform_ptr(P.gm[2]) : uniform_ptr(P.gm[5]);
                        c2 = P.dc2[gs.pidx - 1], c1 = P.dc1[gs.pidx - 1];
#pragma unroll
                        for (int ic = 0; ic < 3; ic++) old[ic] = gp[e + ic * s];
                    }
                }
#pragma unroll
                for (int ic = 0; ic < 3; ic++) {
                    const T dv = c2 * (G[id][ic] * C.vol) + c1 * old[ic];
                    if (wr) gq[e + ic * s] = dv;
                    G[id][ic] = dv * C.rvol;
                }
            }
        }
        T s[6];
        // source ops: the boxes met by this tile, then those of plane l (both
        // uniform over the block), then those holding this cell
        unsigned bx = 0;
        if (tile_boxes) {
            const unsigned cl = ops_boxes_l(ops, tile_boxes, l);
            if (cl && cellv) bx = ops_boxes(ops, cl, j, k, l);
        }
        // finite source (potency, moment) on this cell
        const long fq = fst.cells && cellv ? fs_index(P, j, k, l) : -1;
        if (!bx && fq < 0) {
            const T tr = G[0][0] + G[1][1] + G[2][2];
            const T m2 = m + m;
            s[0] = m2 * G[0][0] + lm * tr;
            s[1] = m2 * G[1][1] + lm * tr;
            s[2] = m2 * G[2][2] + lm * tr;
            s[3] = m * (G[2][1] + G[1][2]);
            s[4] = m * (G[0][2] + G[2][0]);
            s[5] = m * (G[1][0] + G[0][1]);
        } else {
            T fv[6];
#pragma unroll
            for (int c = 0; c < 6; c++) fv[c] = fq >= 0 ? P.fs[c][fq] : T(0);
            cell_ops<T>(ops, bx, g.at(j, k, l), G, lm, m, C.vol, C.rvol, fq >= 0 ? P.fs_kind : 0, fv, s);
        }
        if (!real)
#pragma unroll
            for (int q = 0; q < 6; q++) s[q] = T(0);
        if (stat_now && cellv && st)
            st_w = tmax(st_w, (s[0] * s[0] + s[1] * s[1] + s[2] * s[2]) + T(2) * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]));
        const T A[3][3] = {{C.ax * s[0], C.ay * s[5], C.az * s[4]},
                           {C.ax * s[5], C.ay * s[1], C.az * s[3]},
                           {C.ax * s[4], C.ay * s[3], C.az * s[2]}};
        const T yf = real ? y : T(0), yv = real ? C.dth2 * y : T(0);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            // modes (s_k s_l, s_j s_l, s_j s_k, s_j s_k s_l) of h1 u + dt h2 v
            // and, for the viscous hourglass of PML nodes, of v
            T h[4], hv[4];
            h[0] = Shi.F[c][1] - Slo.F[c][1];
            h[1] = Shi.F[c][0] - Slo.F[c][0];
            h[2] = Slo.F[c][2] + Shi.F[c][2];
            h[3] = Slo.F[c][2] - Shi.F[c][2];
            if constexpr (MASK != 0) {
                hv[0] = Shi.v[c][1] - Slo.v[c][1];
                hv[1] = Shi.v[c][0] - Slo.v[c][0];
                hv[2] = Slo.v[c][3] + Shi.v[c][3];
                hv[3] = Slo.v[c][3] - Shi.v[c][3];
            } else {
#pragma unroll
                for (int q = 0; q < 4; q++) hv[q] = T(0);
            }
#pragma unroll
            for (int q = 0; q < 4; q++) h[q] = yf * h[q], hv[q] = yv * hv[q];
            E[c][0] = A[c][0];
            E[c][1] = A[c][1];
            E[c][2] = A[c][2];
            Cc[c][0] = A[c][0];
            Cc[c][1] = A[c][1];
            Cc[c][2] = -A[c][2];
            E[c][3] = -h[1], E[c][4] = -h[0], E[c][5] = -h[2] - h[3];
            Cc[c][3] = h[1], Cc[c][4] = h[0], Cc[c][5] = -h[2] + h[3];
            E[c][6] = -hv[1], E[c][7] = -hv[0], E[c][8] = -hv[2] - hv[3];
            Cc[c][6] = hv[1], Cc[c][7] = hv[0], Cc[c][8] = -hv[2] + hv[3];
        }
    };

    // The plane loop is unrolled by two with alternating register sets, as in
    // k_fused: no register holding a load in flight or a carried value is copied.
    T own[7], E[3][9];
    PlaneSums<T> SA, SB;  // plane sums of two planes (lower, upper in turn)
    T CA[3][9], CB[3][9];
    T fA[7], mA[NMAT], fB[7], mB[NMAT];
    fetch_node(wk.lz0 - 1, fA);
    exchange(fA, SA);
    fetch_node(wk.lz0, fB);
    fetch_cell(wk.lz0 - 1, mB);
#pragma unroll
    for (int i = 0; i < 3; i++) uo[i] = fB[i], vo[i] = fB[3 + i];
    __syncthreads();  // every wave has read the first plane from LDS
    exchange(fB, SB);
    fetch_node(wk.lz0 + 1, fA);
    fetch_cell(wk.lz0, mA);
    cell_at(wk.lz0 - 1, mB, SA, SB, E, CA, false, true);
    // every wave has read the second plane from LDS before the loop's first
    // plane writes its own slots again
    __syncthreads();

    auto plane = [&](int l, bool act, const T (&cf)[7], const T (&cm)[NMAT], T (&nf)[7], T (&nm_)[NMAT],
                     const PlaneSums<T> &Slo, PlaneSums<T> &Shi, const T (&carry)[3][9], T (&Cn)[3][9]) {
#pragma unroll
        for (int q = 0; q < 7; q++) own[q] = cf[q];
        fetch_node(l + 2, nf);
        fetch_cell(l + 1, nm_);
        exchange(own, Shi);
        cell_at(l, cm, Slo, Shi, E, Cn, ownjk && act, act);
        // plane combination, then the k stage: own part (s_k = +1) and the part
        // sent to the row above (s_k = -1): divergence x, y, z, full (c0, cj), viscous (c0, cj)
        T R[3][7];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            T q[9];
#pragma unroll
            for (int z = 0; z < 9; z++) q[z] = E[c][z] + carry[c][z];
            R[c][0] = q[0];
            R[c][1] = q[1];
            R[c][2] = q[2];
            R[c][3] = q[4];         // full hourglass: constant = s_k coefficient
            R[c][4] = q[3] + q[5];  // s_j coefficient
            R[c][5] = q[7];
            R[c][6] = q[6] + q[8];
            sC[7 * c + 0][t] = q[0];
            sC[7 * c + 1][t] = -q[1];
            sC[7 * c + 2][t] = q[2];
            sC[7 * c + 3][t] = -q[4];
            sC[7 * c + 4][t] = q[3] - q[5];
            sC[7 * c + 5][t] = -q[7];
            sC[7 * c + 6][t] = q[6] - q[8];
        }
        __syncthreads();
        // j stage: divergences per direction and the two hourglass variants
        T D[3][3], Hf[3], Hv[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            T r[7];
#pragma unroll
            for (int q = 0; q < 7; q++) r[q] = R[c][q] + sC[7 * c + q][tbe];
            D[c][0] = r[0] + lane_prev(-r[0]);
            D[c][1] = r[1] + lane_prev(r[1]);
            D[c][2] = r[2] + lane_prev(r[2]);
            Hf[c] = (r[3] + r[4]) + lane_prev(r[3] - r[4]);
            Hv[c] = (r[5] + r[6]) + lane_prev(r[5] - r[6]);
        }
        const bool nl_in = l >= P.i1node[2] && l <= P.i2node[2];
        if (njk && act && nl_in) {
            const long o = g.at(j, k, l);
            // PML momentum memory of each damped direction (sord/acceleration.f90:37-104)
#pragma unroll
            for (int id = 0; id < 3; id++) {
                if (!(MASK >> id & 1)) continue;
                T c2, c1, old[3];
                gptr<T> pp;
                long e, s;
                if (id < 2) {
                    if (!nin[id]) continue;  // not a PML node of axis id
                    e = (long)lpl[id] * (l - 1), s = lps[id];
                    pp = lp[id];
                    c2 = ln2[id], c1 = ln1[id];
                    if constexpr (PF) {
#pragma unroll
                        for (int ic = 0; ic < 3; ic++) old[ic] = cm[7 + ic];
                    } else {
#pragma unroll
                        for (int ic = 0; ic < 3; ic++) old[ic] = pp[e + ic * s];
                    }
                } else {
                    const LSlot ps = pslot_l(l);
                    if (!ps.on) continue;
                    e = ps.e, s = ps.s;
                    pp = ps.lo ? uniform_ptr(P.p[2]) : uniform_ptr(P.p[5]);
                    if constexpr (PF) {
                        c2 = cm[ZC + 2], c1 = cm[ZC + 3];
#pragma unroll
                        for (int ic = 0; ic < 3; ic++) old[ic] = cm[7 + ic];
                    } else {
                        c2 = P.dn2[ps.pidx - 1], c1 = P.dn1[ps.pidx - 1];
#pragma unroll
                        for (int ic = 0; ic < 3; ic++) old[ic] = pp[e + ic * s];
                    }
                }
#pragma unroll
                for (int ic = 0; ic < 3; ic++) {
                    const T sv = c2 * D[ic][id] + c1 * old[ic];
                    pp[e + ic * s] = old[ic] + P.dt * sv;
                    D[ic][id] = sv;
                }
            }
            T f[3];
#pragma unroll
            for (int ic = 0; ic < 3; ic++)  // reference order: id = ic, ic+1, ic+2
                f[ic] = (D[ic][ic] + D[ic][(ic + 1) % 3]) + D[ic][(ic + 2) % 3];
            bool interior = hint_jk;
            if ((MASK & 4) && (l < P.i1pml[2] + 1 || l > P.i2pml[2] - 1)) interior = false;
#pragma unroll
            for (int c = 0; c < 3; c++) f[c] = interior ? f[c] + Hf[c] : (MASK && P.hg_pml) ? f[c] + Hv[c] : f[c];
            if (fst.nodes) fs_force(P, j, k, l, f);
            // plane-local boundary conditions (sord/boundary_cond.f90:44-66)
            const int z = zjk | bc_zero(2, l);
#pragma unroll
            for (int c = 0; c < 3; c++)
                if (z >> c & 1) f[c] = T(0);
            if (fjk || (fa == 2 && (l == P.irup || l == P.irup + 1))) {
                // fault planes: the rupture kernel finishes these nodes
#pragma unroll
                for (int c = 0; c < 3; c++) P.a[c][o] = f[c];
            } else {
                const T mrn = cm[3];
                T a2 = T(0), v2 = T(0), u2 = T(0);
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const T a = f[c] * mrn;
                    const T vv = vo[c] + P.dt * a;
                    const T uu = uo[c] + P.dt * vv;
                    P.vn[c][o] = vv;
                    P.un[c][o] = uu;
                    a2 = a2 + a * a;
                    v2 = v2 + vv * vv;
                    u2 = u2 + uu * uu;
                }
                if (stat_now) st_a = tmax(st_a, a2);
                if (stat_next) st_v = tmax(st_v, v2), st_u = tmax(st_u, u2);
            }
        }
#pragma unroll
        for (int c = 0; c < 3; c++) uo[c] = cf[c], vo[c] = cf[3 + c];
    };
    // both planes of a pair always run (one loop back-edge keeps the register
    // sets in place); past the last plane of the item the second one stores nothing
    for (int l = wk.lz0; l <= wk.lz1; l += 2) {
        plane(l, true, fA, mA, fB, mB, SB, SA, CA, CB);
        plane(l + 1, l + 1 <= wk.lz1, fB, mB, fA, mA, SA, SB, CB, CA);
    }
    // one atomic per wave and slot (four waves per block).  The block-level
    // reduction of the interior kernel lost maxima here: double precision,
    // the one-axis kernels of the layered model with free surface (tiles 8,
    // 16, 4), 19 in 20 runs, the node with the largest acceleration missing
    // from a step's maximum while its velocity was right; no cause was found
    // in the kernel's source or its LDS layout, and this form lost none in
    // 20 runs.
    if (stat_now) {
        st_a = wave_max(st_a);
        st_w = wave_max(st_w);
        if (lane == 0) atomic_max_bits<T>(P.stat_cur + 0, st_a), atomic_max_bits<T>(P.stat_cur + 3, st_w);
    }
    if (stat_next) {
        st_v = wave_max(st_v);
        st_u = wave_max(st_u);
        if (lane == 0) atomic_max_bits<T>(P.stat_nxt + 1, st_v), atomic_max_bits<T>(P.stat_nxt + 2, st_u);
    }
}

// FP32: at most 256 VGPRs, two waves per SIMD
template <class T, int W, int MASK>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 2 : 1)))
void k_fusedp(Params<T> P, FusedC<T> C, PointOps<T> ops) {
    const Work wk = C.work[blockIdx.x];
    if (wk.lz0 > wk.lz1) return;  // uniform over the block
    __shared__ T lds[boundary_lds_slots<W>()];
    boundary_item<T, W, MASK, false>(P, C, ops, wk, lds);
}
// a finite source: every boundary item on the kernel of all damped axes
template <class T, int W>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 2 : 1)))
void k_fusedp_fs(Params<T> P, FusedC<T> C, PointOps<T> ops) {
    const Work wk = C.work[blockIdx.x];
    if (wk.lz0 > wk.lz1) return;
    __shared__ T lds[boundary_lds_slots<W>()];
    boundary_item<T, W, 7, true>(P, C, ops, wk, lds);
}
template <class T, int W, int MASK>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 2 : 1)))
void k_fusedp_batch(const Scen<T> *__restrict__ sc, const Work *__restrict__ work) {
    const Work wk = work[blockIdx.x];
    if (wk.lz0 > wk.lz1) return;
    const Scen<T> &S = sc[wk.pad1];
    if (!S.active) return;  // uniform over the block
    __shared__ T lds[boundary_lds_slots<W>()];
    boundary_item<T, W, MASK, false>(S.P, S.C, S.ops, wk, lds);
}

// blocks of the general boundary kernel (the most registers of the boundary
// kernels) one CU holds at once
template <class T>
int boundary_blocks_per_cu() {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_fusedp<T, BOUNDARY_WAVES, 7>, 64 * BOUNDARY_WAVES, 0) !=
        hipSuccess) {
        (void)hipGetLastError();
        return 1;
    }
    return n > 0 ? n : 1;
}
template int boundary_blocks_per_cu<float>();
template int boundary_blocks_per_cu<double>();

template <class T>
FusedC<T> fused_consts(const Params<T> &P, const T sp[3], const Work *work) {
    FusedC<T> C;
    C.work = work;
    C.gx = T(1) / (T(4) * sp[0]);
    C.gy = T(1) / (T(4) * sp[1]);
    C.gz = T(1) / (T(4) * sp[2]);
    C.ax = sp[1] * sp[2] * T(0.25);
    C.ay = sp[0] * sp[2] * T(0.25);
    C.az = sp[0] * sp[1] * T(0.25);
    C.h1 = P.hg_on ? P.h1 : T(0);
    C.dth2 = P.hg_on ? P.dth2 : T(0);
    C.vol = sp[0] * sp[1] * sp[2];
    C.rvol = T(1) / C.vol;
    return C;
}

template <class T>
void launch_fused(const Params<T> &P, const PointOps<T> &ops, const Work *work, int nwork, int mask, const T sp[3],
                  int waves, bool fs, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
    if (nwork <= 0) return;
    const FusedC<T> C = fused_consts<T>(P, sp, work);
    using K = void (*)(Params<T>, FusedC<T>, PointOps<T>);
    const bool st = P.stat_now || P.stat_next;
    K k;
    int w = mask >= 0 ? BOUNDARY_WAVES : waves == 4 || waves == 8 ? waves : 16;
    if (fs) {
        k = mask >= 0 ? k_fusedp_fs<T, BOUNDARY_WAVES>
            : w == 4  ? k_fused_fs<T, 4>
            : w == 8  ? k_fused_fs<T, 8>
                      : k_fused_fs<T, 16>;
    } else if (mask < 0) {
        k = st ? (w == 4 ? k_fused_stat<T, 4> : w == 8 ? k_fused_stat<T, 8> : k_fused_stat<T, 16>)
               : (w == 4 ? k_fused<T, 4> : w == 8 ? k_fused<T, 8> : k_fused<T, 16>);
    } else {
        switch (mask) {
        case 0: k = k_fusedp<T, BOUNDARY_WAVES, 0>; break;
        case 1: k = k_fusedp<T, BOUNDARY_WAVES, 1>; break;
        case 2: k = k_fusedp<T, BOUNDARY_WAVES, 2>; break;
        case 3: k = k_fusedp<T, BOUNDARY_WAVES, 3>; break;
        case 4: k = k_fusedp<T, BOUNDARY_WAVES, 4>; break;
        case 5: k = k_fusedp<T, BOUNDARY_WAVES, 5>; break;
        case 6: k = k_fusedp<T, BOUNDARY_WAVES, 6>; break;
        default: k = k_fusedp<T, BOUNDARY_WAVES, 7>; break;
        }
    }
    const dim3 g(nwork), b(64, w);
    // with events: the runtime stamps the kernel's own start and end into them
    if (e0) hipExtLaunchKernelGGL(k, g, b, 0, s, e0, e1, 0, P, C, ops);
    else hipLaunchKernelGGL(k, g, b, 0, s, P, C, ops);
}

template <class T>
void launch_fused_batch(const Scen<T> *sc, const Work *work, int nwork, int mask, bool st, hipStream_t s) {
    if (nwork <= 0) return;
    const dim3 g(nwork), b8(64, 8), bb(64, BOUNDARY_WAVES);
    switch (mask) {
    case -1:
        if (st) hipLaunchKernelGGL((k_fused_batch_stat<T, 8>), g, b8, 0, s, sc, work);
        else hipLaunchKernelGGL((k_fused_batch<T, 8>), g, b8, 0, s, sc, work);
        break;
    case 0: hipLaunchKernelGGL((k_fusedp_batch<T, BOUNDARY_WAVES, 0>), g, bb, 0, s, sc, work); break;
    case 1: hipLaunchKernelGGL((k_fusedp_batch<T, BOUNDARY_WAVES, 1>), g, bb, 0, s, sc, work); break;
    case 2: hipLaunchKernelGGL((k_fusedp_batch<T, BOUNDARY_WAVES, 2>), g, bb, 0, s, sc, work); break;
    case 3: hipLaunchKernelGGL((k_fusedp_batch<T, BOUNDARY_WAVES, 3>), g, bb, 0, s, sc, work); break;
    case 4: hipLaunchKernelGGL((k_fusedp_batch<T, BOUNDARY_WAVES, 4>), g, bb, 0, s, sc, work); break;
    case 5: hipLaunchKernelGGL((k_fusedp_batch<T, BOUNDARY_WAVES, 5>), g, bb, 0, s, sc, work); break;
    case 6: hipLaunchKernelGGL((k_fusedp_batch<T, BOUNDARY_WAVES, 6>), g, bb, 0, s, sc, work); break;
    default: hipLaunchKernelGGL((k_fusedp_batch<T, BOUNDARY_WAVES, 7>), g, bb, 0, s, sc, work); break;
    }
}

// clear the statistics slots the scenarios of a batch fill in the coming step
// (one thread per scenario; replaces per-scenario memsets)
template <class T>
__global__ void k_zero_stats(const StatZero<T> *__restrict__ z, int n) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    if (z[q].now) z[q].cur[0] = T(0), z[q].cur[3] = T(0);
    if (z[q].next) z[q].nxt[1] = T(0), z[q].nxt[2] = T(0);
}
template <class T>
void launch_zero_stats(const StatZero<T> *z, int n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL((k_zero_stats<T>), dim3((n + 63) / 64), dim3(64), 0, s, z, n);
}
template void launch_zero_stats<float>(const StatZero<float> *, int, hipStream_t);
template void launch_zero_stats<double>(const StatZero<double> *, int, hipStream_t);

template FusedC<float> fused_consts<float>(const Params<float> &, const float *, const Work *);
template FusedC<double> fused_consts<double>(const Params<double> &, const double *, const Work *);
template void launch_fused_batch<float>(const Scen<float> *, const Work *, int, int, bool, hipStream_t);
template void launch_fused_batch<double>(const Scen<double> *, const Work *, int, int, bool, hipStream_t);
template void launch_fused<float>(const Params<float> &, const PointOps<float> &, const Work *, int, int,
                                  const float *, int, bool, hipStream_t, hipEvent_t, hipEvent_t);
template void launch_fused<double>(const Params<double> &, const PointOps<double> &, const Work *, int, int,
                                   const double *, int, bool, hipStream_t, hipEvent_t, hipEvent_t);

}  // namespace sord
