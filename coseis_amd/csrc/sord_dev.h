// Device-side state and parameters of the SORD solver step.
//
// Layout in HBM: every volume field is a 3-D array over the padded local
// mesh nm = nn + 2*nhalo (sord/setup.f90:56-57), j fastest, with no row
// padding (row pitch P0 = nm[0]) and a base offset that puts the array's first
// interior node on a 128-byte boundary.  Vectors and
// tensors are structure-of-arrays: one such 3-D array per component.
// Indices used in kernels are the reference's 1-based local indices.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define SORD_MAXPT 64   // sub-cell / single-cell ops applied inside a kernel per step
#define SORD_MAXBOX 8   // boxes grouping the cells of those ops

template <class T>
struct Grid {
    int nm[3];      // padded local shape
    int nn[3];      // global shape
    int nnoff[3];   // local -> global offset
    int P0;         // row pitch (elements)
    long P1;        // plane pitch (elements) = P0 * nm[1]
    long off;       // base offset (elements)
    long S;         // allocation size of one 3-D array (elements)
    __host__ __device__ __forceinline__ long at(int j, int k, int l) const {
        return off + (long)(j - 1) + (long)P0 * (k - 1) + P1 * (long)(l - 1);
    }
};

// limits and constants of the device-side material model
template <class T>
struct MatInit {
    T rho_min, rho_max, vp_min, vp_max, vs_min, vs_max, gam_min, gam_max, vdamp;
    int3 h1, h2;  // the cells: the halo outside is zeroed
    static __host__ __device__ constexpr T big() { return sizeof(T) == 4 ? T(3.40282347e38f) : T(1.7976931348623157e308); }
};
// seven cell arrays reduced together (rho, vp, vs, gam, lam, mu, nu)
template <class T>
struct MatArrays {
    const T *p[7];
};

// A small list of per-cell or per-node operations (value already scaled by the
// time function) applied inside the stress kernel: potency, moment, stress.
template <class T>
struct PointOps {
    int n;
    int stage[SORD_MAXPT];   // 0: strain (potency p**), 1: moment m**, 2: stress w**
    int comp[SORD_MAXPT];    // 0..5: xx yy zz yz zx xy
    int op[SORD_MAXPT];      // 0 '=', 1 '+', 2 '*'
    long cell[SORD_MAXPT];   // linear index
    T val[SORD_MAXPT];
    int nbox;                               // boxes holding the op cells (1-based local cell indices)
    int blo[SORD_MAXBOX][3], bhi[SORD_MAXBOX][3];
    int bop[SORD_MAXBOX + 1];               // the ops of box b are bop[b] .. bop[b+1]-1; all ops of
                                            // one cell are in one box, in their order of application
};

// boxes of `cand` (bit mask) that hold cell (j, k, l)
template <class T>
__host__ __device__ inline unsigned ops_boxes(const PointOps<T> &ops, unsigned cand, int j, int k, int l) {
    unsigned r = 0;
    for (unsigned m = cand; m; m &= m - 1) {
        const int b = __builtin_ctz(m);
        if (j >= ops.blo[b][0] && j <= ops.bhi[b][0] && k >= ops.blo[b][1] && k <= ops.bhi[b][1] &&
            l >= ops.blo[b][2] && l <= ops.bhi[b][2])
            r |= 1u << b;
    }
    return r;
}

// boxes that meet the cell columns [j0, j1] x [k0, k1]
template <class T>
__host__ __device__ inline unsigned ops_boxes_jk(const PointOps<T> &ops, int j0, int j1, int k0, int k1) {
    unsigned r = 0;
    for (int b = 0; b < ops.nbox; b++)
        if (j1 >= ops.blo[b][0] && j0 <= ops.bhi[b][0] && k1 >= ops.blo[b][1] && k0 <= ops.bhi[b][1]) r |= 1u << b;
    return r;
}

// boxes of `cand` whose planes include l
template <class T>
__host__ __device__ inline unsigned ops_boxes_l(const PointOps<T> &ops, unsigned cand, int l) {
    unsigned r = 0;
    for (unsigned m = cand; m; m &= m - 1) {
        const int b = __builtin_ctz(m);
        if (l >= ops.blo[b][2] && l <= ops.bhi[b][2]) r |= 1u << b;
    }
    return r;
}

// Finite kinematic source (sord/kinematic_source.f90): n subfaults; xi are
// local cell (potency, moment) or node (force) coordinates; w[6] the tensor
// (xx yy zz yz zx xy) or force (x y z) weights; history samples of subfault q
// from off[q].
template <class T>
struct FSrc {
    int n, kind;  // kind 1 potency, 2 moment, 3 force
    const T *xi[3], *w[6], *t0, *dt, *hist;
    const int *nt, *off;
    int lo[3], hi[3];  // cell or node range of the mesh
};

template <class T>
struct Params {
    Grid<T> g;
    // finite source values of the current step over the box fs_lo + [0, fs_n):
    // cells of the potency (fs_kind 1) or moment (2), six components, added by
    // the stress computation; nodes of the force (3), three components, added
    // to the force by the fused kernels (the stage path adds it straight to a)
    int fs_kind, fs_lo[3], fs_n[3];
    T *fs[6];
    T dt;
    // node vectors (current state) and next-state buffers
    T *u[3], *v[3], *un[3], *vn[3];
    T *a[3];            // force / acceleration (reference w1 during step_accel)
    T *sig[6];          // cell stress: xx yy zz (reference w1), yz zx xy (reference w2)
    T *gam, *mr;        // node: dt * node-averaged viscosity, inverse nodal mass
    T *lam, *mu, *yy, *vc;  // cell: Lame moduli, hourglass constant, 1/volume
    T *hx, *hy, *hz;    // cell widths along each axis (rectilinear mesh), 1-based cell index
    // non-rectilinear meshes (gen = 1): per cell the support-operator weight
    // of corner q (bits dj + 2 dk + 4 dl) along axis a, cw[8 a + q], so that
    // diff_nc(f, a) = sum_q cw[8 a + q] f(q); diff_cn is minus its transpose
    int gen;
    T *cw[24];
    // region bounds (1-based local)
    int i1node[3], i2node[3], i1cell[3], i2cell[3], i1pml[3], i2pml[3];
    int i1bc[3], i2bc[3];
    int bc1[3], bc2[3];
    // PML
    int npml;
    T *dn1, *dn2, *dc1, *dc2;   // damping tables, index p-1
    T *p[6], *gm[6];            // node / cell memory slabs: lower x,y,z then upper x,y,z
    T *gmn[6];                  // second cell-memory buffer: the fused kernels read gm, write gmn
    int pn[6][3], gn[6][3];     // their shapes (Fortran order, per component)
    // hourglass
    int hg_on, hg_pml;
    T h1, dth2;
    // fault
    int ifn, irup;              // 1-based axis (0 none) and local index of the minus side
    int store_a;                // fused update also stores the acceleration (statistics steps)
    // statistics maxima of squared norms, as bit patterns for atomic max:
    // stat_cur[0] |a|^2 and [3] ||w||^2 of this step, stat_nxt[1] |v|^2 and
    // [2] |u|^2 of the state the update produces (next step's statistics)
    T *stat_cur, *stat_nxt;
    int stat_now, stat_next;
};

// Partial maxima of the max-norm reduction: at most this many blocks.
constexpr int MAXNORM_BLOCKS = 1024;
// Layout of the statically allocated scratch of the allocation-free step
// (sord_step_static): the two sets of four statistics slots, then the
// reduction's partial maxima and their locations (8 bytes each in FP64).
constexpr size_t SCRATCH_STAT = 0, SCRATCH_VAL = 256, SCRATCH_LOC = SCRATCH_VAL + 8 * MAXNORM_BLOCKS,
                 SCRATCH_BYTES = SCRATCH_LOC + 8 * MAXNORM_BLOCKS;

// One block of the fused kernels: node columns [jb, je] x [kb, ke] (at most
// 62 x (W-2)) over planes [lz0, lz1]; lz0 > lz1 marks padding.
// waves per block of the boundary kernel, and the widest column segment of its
// work items (two packed node rows per wavefront with 4 waves)
#ifndef SORD_BOUNDARY_WAVES
#define SORD_BOUNDARY_WAVES 4
#endif
constexpr int BOUNDARY_WAVES = SORD_BOUNDARY_WAVES;
constexpr int BOUNDARY_COLS = BOUNDARY_WAVES == 4 ? 30 : 62;

struct Work {
    int jb, je, kb, ke, lz0, lz1, pad0, pad1;
};

// constants of the fused kernels (uniform spacing per axis)
template <class T>
struct FusedC {
    const Work *work;
    T gx, gy, gz;  // 1 / (4 dx_a): gradient from the sum of four edge differences
    T ax, ay, az;  // dx_b dx_c / 4: divergence weights
    T h1, dth2;    // hourglass(1), dt * hourglass(2)
    T vol, rvol;   // cell volume and its inverse
};

// arguments of the adjoint step kernels
template <class T>
struct AdjP {
    Grid<T> g;
    const T *uh[3], *vh[3];   // adjoint state of step it+1
    T *uo[3], *vo[3];         // adjoint state of step it (out)
    const T *u[3], *v[3];     // forward state S_it
    int gm_out;               // 0 on the last adjoint step (s = 1): no new (vb + dt ub) term
    const T *mr, *gam, *lam, *mu, *yy;
    T *gl, *gmu, *gy, *gm;    // material sums: cells lambda, mu, yy; nodes 1/mass
    T dt, h1, dth2, vol, rd[3];  // rd: 1 / (4 d_a)
    int lo[3], hi[3];          // real cells
    int nlo[3], nhi[3];        // real nodes
    // residual added to the new adjoint velocity at receiver nodes: per work
    // item Work.pad1 entries from Work.pad0 of rtile (j, k, l, receiver)
    const int4 *rtile;
    const T *res;              // this step's residuals (receiver x 3)
};

// waves per block of the adjoint kernel (one node row each; the first and
// last are halo rows): 12 in single precision, which fit its registers (168
// VGPRs, two spilled: three waves per SIMD, a sixth less halo; 200^3 adjoint
// step 305 -> 287 us by rocprofv3), 8 in double precision (its registers
// spill at any count)
template <class T>
constexpr int adj_waves() { return sizeof(T) == 4 ? 12 : 8; }

// the chain rule of the adjoint's material sums to cell rho, vp, vs
template <class T>
struct AdjChain {
    Grid<T> g;
    const T *gl, *gmu, *gy, *gm;     // sums: cells lambda, mu, yy; nodes 1/mass
    const T *rho, *vp, *vs;          // cell rho, vp, vs after the limits
    const T *lam, *mu, *mr;          // cell lambda, mu; node 1/mass
    T *grho, *gvp, *gvs;             // out: real cells, Fortran order
    T vol, sy;                       // cell volume; hourglass scale
    int lo[3], hi[3];                // real cells
};

// statistics slots of one scenario of a batch to clear before a step
template <class T>
struct StatZero {
    T *cur, *nxt;   // slot sets of this step and of the next
    int now, next;  // this step computes statistics / the next one does
};

// one scenario of a batch: what the plain fused kernels get as arguments
template <class T>
struct Scen {
    Params<T> P;
    FusedC<T> C;
    PointOps<T> ops;
    int active;  // 0: the scenario sits this step out (its items return at once)
};

struct Box {
    int lo[3], hi[3], st[3], n[3];  // 1-based local indices, steps, counts
};

// A strided 3-D view: element (j,k,l) (1-based) is at off + (j-1) + s1*(k-1) + s2*(l-1).
struct View {
    long off, s1, s2;
    __host__ __device__ __forceinline__ long at(int j, int k, int l) const {
        return off + (long)(j - 1) + s1 * (long)(k - 1) + s2 * (long)(l - 1);
    }
};

template <class T>
struct Samples {
    int n;
    int lo[SORD_MAXPT][3], hi[SORD_MAXPT][3];
    T xi[SORD_MAXPT][3];
    int slot[SORD_MAXPT];  // output offset
};

// Fault-plane arrays: the local mesh with extent 1 along the normal, Fortran
// order (sord/dynamic_rupture.f90:22-63).
template <class T>
struct Fault {
    int fm[3];
    T *area, *nhat[3], *t0[3];
    T *mus, *mud, *dc, *co, *rhypo;
    T *sl, *psv, *trup, *tarr, *svold;
    T *t1[3], *t3[3], *ts, *tn, *fr, *sa[3], *sam;
    T *sv[3], *su[3], *svm, *sum;
    T *muf, *efacc, *es, *mo;   // energy terms: friction work (accumulated), strain energy, moment
    int faultopening, nucleate;
    T trelax, vrup, rcrit;
};
