// gfx950 kernels for the incompressible SIMPLE + SST k-omega forward solve.
//
// One workgroup owns one draw (one coefficient vector on one mesh) for the
// whole solve: a persistent kernel runs every outer iteration, every Krylov
// iteration and the FP64 convergence decision on the device. The draw's
// fields, matrices and Krylov vectors live in an arena that is placed in LDS
// when it fits (kLDS) and in HBM otherwise; the mesh tables are shared by
// all draws on the mesh and are read through the caches.
//
// All work is cell-centric: a cell gathers the contributions of its faces in
// the order of its face list, which is the face order the reference's face
// loops accumulate in. So no atomics are needed, results do not depend on the
// batch composition, and the FP64 assembly reproduces the reference's
// rounding. Block reductions use 64-lane butterflies, then LDS across waves,
// in a fixed order.
#include <hip/hip_runtime.h>

#include <math.h>

#include <algorithm>
#include <mutex>
#include <type_traits>

#include "rans_device.h"

// The file is compiled twice: part 0 holds the FP64 batch, multi-workgroup
// and 256-thread serving kernels with every launch entry point, part 1 the
// mixed-precision instantiations of the batch kernels and the 512-thread
// serving kernel.
#ifndef RQ_KERNEL_PART
#define RQ_KERNEL_PART 0
#endif

namespace rq {

#define DEV __device__ __forceinline__

DEV double smax(double a, double b) { return (a < b) ? b : a; }
DEV double smin(double a, double b) { return (b < a) ? b : a; }

// ----------------------------------------------------------------------------
// block reductions. A wave reduces with DPP lane moves in a fixed tree (pairs,
// quads, half rows, rows, then row broadcasts), the total is read from lane
// 63 into a scalar register, so every lane holds the same bits. Waves then
// combine through LDS in wave order.
// ----------------------------------------------------------------------------
template <int CTRL, int ROW_MASK = 0xf>
DEV double dpp_move(double v, double old) {
    const int lo = __builtin_amdgcn_update_dpp(__double2loint(old), __double2loint(v), CTRL,
                                               ROW_MASK, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(__double2hiint(old), __double2hiint(v), CTRL,
                                               ROW_MASK, 0xf, false);
    return __hiloint2double(hi, lo);
}
DEV double lane63(double v) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
    return __hiloint2double(hi, lo);
}
DEV double wave_sum(double v) {
    v += dpp_move<0xB1>(v, 0.0);        // quad_perm [1,0,3,2]
    v += dpp_move<0x4E>(v, 0.0);        // quad_perm [2,3,0,1]
    v += dpp_move<0x141>(v, 0.0);       // row_half_mirror
    v += dpp_move<0x140>(v, 0.0);       // row_mirror
    v += dpp_move<0x142, 0xa>(v, 0.0);  // row_bcast:15 into rows 1 and 3
    v += dpp_move<0x143, 0xc>(v, 0.0);  // row_bcast:31 into rows 2 and 3
    return lane63(v);
}
DEV double wave_max(double v) {
    v = smax(v, dpp_move<0xB1>(v, v));
    v = smax(v, dpp_move<0x4E>(v, v));
    v = smax(v, dpp_move<0x141>(v, v));
    v = smax(v, dpp_move<0x140>(v, v));
    v = smax(v, dpp_move<0x142, 0xa>(v, v));
    v = smax(v, dpp_move<0x143, 0xc>(v, v));
    return lane63(v);
}

template <int CTRL, int ROW_MASK = 0xf>
DEV float dpp_movef(float v, float old) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), CTRL,
                                                      ROW_MASK, 0xf, false));
}
DEV float wave_sumf(float v) {
    v += dpp_movef<0xB1>(v, 0.0f);
    v += dpp_movef<0x4E>(v, 0.0f);
    v += dpp_movef<0x141>(v, 0.0f);
    v += dpp_movef<0x140>(v, 0.0f);
    v += dpp_movef<0x142, 0xa>(v, 0.0f);
    v += dpp_movef<0x143, 0xc>(v, 0.0f);
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

// Grid barrier of the workgroups of one draw. Arrivals go to one 64-bit
// counter that is never reset: barrier k of the launch is passed when the
// counter reaches nblk * k. Every wait is bounded: past kGridSpinLimit polls
// the waiting workgroup raises abort, and every workgroup then leaves every
// later barrier at once (the draw ends as Unknown), so no wave is left
// waiting.
//
// Two visibility regimes. In general the workgroups may sit on different
// XCDs, whose L2 caches are not coherent with each other, so the barrier is
// an agent-scope release/acquire pair (the compiler emits an L2 write-back
// before arriving and an L2 invalidate after leaving). When every workgroup
// of the draw has been seen on the same XCD (GridLocal::local), they share
// one L2, which then holds the latest value of every line they write (the
// vector L1 is write-through). It is enough that each wave's stores have
// completed before the arrival (s_waitcnt vmcnt(0)); the L2 keeps its
// contents across the barrier. Stale lines can only sit in a CU's L1, so a
// multi-workgroup draw reads every value another workgroup may have written
// with device-scope loads, which miss the L1 (GRef below).
struct GridLocal {
    unsigned long long target;  // arrivals that end the next barrier (thread 0)
    int flag;                   // 1: the last barrier completed; 0: aborted
    int local;                  // 1: all workgroups of the draw share one XCD
    long long wait, count;      // shader clocks thread 0 spent in barriers, barriers
};

constexpr long long kGridSpinLimit = 1LL << 25;
// s_getreg_b32 operand of HW_REG_XCC_ID (register 20, bits 0..3): the XCD
// the wave runs on
constexpr int kHwRegXccId = 20 | (3 << 11);
DEV bool grid_barrier(GridState *g, int nblk, GridLocal *gl) {
    const bool local = gl->local != 0;
    if (local)
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    else
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (threadIdx.x == 0) {
        const long long t0 = clock64();
        int ok = 1;
        const unsigned long long target = gl->target + (unsigned long long)nblk;
        gl->target = target;
        if (__hip_atomic_load(&g->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
            ok = 0;
        } else {
            __hip_atomic_fetch_add(&g->count, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            long long spins = 0;
            while ((long long)(__hip_atomic_load(&g->count, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT) - target) < 0) {
                if (__hip_atomic_load(&g->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                    ok = 0;
                    break;
                }
                if (++spins > kGridSpinLimit) {
                    __hip_atomic_store(&g->abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    ok = 0;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        gl->flag = ok;
        gl->wait += clock64() - t0;
        gl->count += 1;
    }
    __syncthreads();
    if (!local) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return gl->flag != 0;
}

// A double in global memory that other workgroups of a multi-workgroup draw
// write: reads are device-scope loads (they miss the CU's L1 and are served
// by the L2), writes are plain stores (the L1 is write-through).
typedef __attribute__((address_space(1))) double GlobalDouble;
struct GRef {
    GlobalDouble *p;
    DEV operator double() const {
        return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    DEV const GRef &operator=(double v) const {
        *p = v;
        return *this;
    }
    DEV const GRef &operator=(const GRef &o) const { return *this = double(o); }
    DEV const GRef &operator+=(double v) const { return *this = double(*this) + v; }
    DEV const GRef &operator-=(double v) const { return *this = double(*this) - v; }
    DEV const GRef &operator*=(double v) const { return *this = double(*this) * v; }
    DEV const GRef &operator/=(double v) const { return *this = double(*this) / v; }
};
struct GPtr {
    GlobalDouble *p;
    DEV GRef operator[](int i) const { return GRef{p + i}; }
};
// The same for the FP32 Krylov vectors of a multi-workgroup draw.
typedef __attribute__((address_space(1))) float GlobalFloat;
struct GRefF {
    GlobalFloat *p;
    DEV operator float() const {
        return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    DEV const GRefF &operator=(float v) const {
        *p = v;
        return *this;
    }
    DEV const GRefF &operator=(const GRefF &o) const { return *this = float(o); }
    DEV const GRefF &operator+=(float v) const { return *this = float(*this) + v; }
};

template <int BS>
struct Reducer {
    static constexpr int NW = BS / 64;
    static constexpr int kStride = 8;  // values per wave in one reduction
    double *buf;  // LDS scratch [2][NW][kStride]
    int parity;
    // multi-workgroup draws: block totals meet in the draw's GridState and
    // every workgroup adds them in rank order (the same bits everywhere)
    GridState *grid = nullptr;
    int nblk = 1, rank = 0, gparity = 0;
    GridLocal *gl = nullptr;

    template <int NV, bool MAX>
    DEV void run(double (&v)[NV]) {
        static_assert(NV >= 1 && NV <= kStride, "a reduction takes at most 8 values");
#pragma unroll
        for (int i = 0; i < NV; ++i) v[i] = MAX ? wave_max(v[i]) : wave_sum(v[i]);
        if constexpr (NW > 1) {
            const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
            double *b = buf + parity * (NW * kStride);
            if (lane == 0)
#pragma unroll
                for (int i = 0; i < NV; ++i) b[wave * kStride + i] = v[i];
            __syncthreads();
#pragma unroll
            for (int i = 0; i < NV; ++i) {
                double a = b[i];
                for (int w = 1; w < NW; ++w)
                    a = MAX ? smax(a, b[w * kStride + i]) : a + b[w * kStride + i];
                v[i] = a;
            }
            parity ^= 1;
        } else {
            __syncthreads();  // orders the phase's LDS writes before later reads
        }
        if (grid) {
            // the grid exchange carries 4 values per workgroup at a time
#pragma unroll
            for (int i0 = 0; i0 < NV; i0 += 4) {
                double *pt = &grid->partials[gparity][0][0];
                if (threadIdx.x == 0)
#pragma unroll
                    for (int i = i0; i < NV && i < i0 + 4; ++i) pt[rank * 4 + i - i0] = v[i];
                grid_barrier(grid, nblk, gl);
                // device-scope loads: the partials come from other workgroups
                auto ld = [&](int k) {
                    return __hip_atomic_load(pt + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                };
#pragma unroll
                for (int i = i0; i < NV && i < i0 + 4; ++i) {
                    double a = ld(i - i0);
                    for (int r = 1; r < nblk; ++r)
                        a = MAX ? smax(a, ld(r * 4 + i - i0)) : a + ld(r * 4 + i - i0);
                    v[i] = a;
                }
                gparity ^= 1;
            }
        }
    }
    template <int NV>
    DEV void sum(double (&v)[NV]) { run<NV, false>(v); }
    template <int NV>
    DEV void max(double (&v)[NV]) { run<NV, true>(v); }
};

// Arena pointers carry their address space explicitly (LDS or global), so
// every access compiles to ds_* or global_* instructions rather than flat.
typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v2i __attribute__((ext_vector_type(2)));
template <bool kLDS>
struct Space;
template <>
struct Space<true> {
    typedef __attribute__((address_space(3))) double D;
    typedef __attribute__((address_space(3))) float F;
    typedef __attribute__((address_space(3))) const int I;
    typedef __attribute__((address_space(3))) const v4i I4;
    typedef __attribute__((address_space(3))) const v2i I2;
};
template <>
struct Space<false> {
    typedef __attribute__((address_space(1))) double D;
    typedef __attribute__((address_space(1))) float F;
    typedef __attribute__((address_space(1))) const int I;
    typedef __attribute__((address_space(1))) const v4i I4;
    typedef __attribute__((address_space(1))) const v2i I2;
};

// slot code: face id in the low 30 bits, side (0 owner, 1 neighbour) in bit 30
DEV int slot_face(int code) { return code & 0x3fffffff; }
DEV bool slot_own(int code) { return (code >> 30) == 0; }

// scalar bookkeeping of one draw, owned by thread 0
struct Book {
    int done, status, error, finalIter, stalled, haveQPrev, turbEst, pad;
    double lastK, lastOm, normUx0, normUy0, normP0, pfRef, worst, rate;
    int nEst[4], pos[4];  // samples pushed, ring slot of the newest sample
    long long kry[5];
    long long tlast;
    long long cyc[kPhases];
};

struct Krylov {
    int iterations, converged;
    double initial;
};

// kServe: the serving kernel, whose problems choose their Krylov precision
// at run time (the batch kernels are compiled for one precision each).
template <int BS, bool kLDS, bool kGrid = false, bool kF32 = false, bool kServe = false>
struct Draw {
    // lanes per cell of the momentum, k and omega row assembly: two in the
    // 256-thread LDS kernels (an LDS arena holds at most about 130 cells, so
    // one lane per cell would leave half of those threads idle), one elsewhere
    static constexpr int QL = (kLDS && !kGrid && BS == 256) ? 2 : 1;
    const DevMesh &M;
    const DevSettings &S;
    const double *c;  // 11 coefficients (LDS)
    Book *bk;
    Krylov *kres;  // LDS slot broadcasting a wave-resident solve's result
    double *amgw;  // multigrid workspace of this draw (global memory) or null
    double *cbuf = nullptr;  // LDS, kMaxAmgCoarse doubles: the coarsest right-hand side
    Reducer<BS> red;
    int nc, nif, nf, nb, nbs;  // nbs = max(nb, 1): boundary arrays' length
    bool planar, has_outlet;
    const int tid = threadIdx.x;
    // multi-lane rows: this lane's place in its cell's group, and the first
    // cell and cell stride of the groups
    const int qi = threadIdx.x % QL;
    const int ct = threadIdx.x / QL, cs = BS / QL;
    // cells, faces and matrix rows are distributed over gs threads; a thread's
    // first index is gt (in one-workgroup mode: the workgroup's threads)
    int gt = threadIdx.x, gs = BS;
    // cell variable the wave-resident Krylov solvers store their solution in
    int xdst = V_kx;
    // multi-workgroup draws: this workgroup's rank among nblk
    int rank = 0, nblk = 1;
    GridState *grid = nullptr;
    GridLocal *gl = nullptr;
    DEV void sync() {
        if constexpr (kGrid) {
            if (!grid_barrier(grid, nblk, gl) && tid == 0 && !bk->done) {
                bk->done = 1;
                bk->status = ST_UNKNOWN;
                bk->error = ERR_GRID_TIMEOUT;
            }
        } else {
            __syncthreads();
        }
    }
    using D = typename Space<kLDS>::D;
    D *Cb, *Fb, *Bb;                            // record bases
    typename Space<kLDS>::I4 *slot4;            // [nc][2]: slot codes, neighbour cells
    typename Space<kLDS>::I2 *fon;              // [nf]: owner, neighbour
    typename Space<kLDS>::I *bflag;             // [nb]
    // contraction rings and monitor values: per workgroup (LDS) in
    // multi-workgroup draws, so each workgroup's bookkeeping replica is private
    using DL = typename Space<kLDS || kGrid>::D;
    DL *ringb, *q, *qprev;
    D *Ob, *Onb;
    typename Space<kLDS>::I *obi, *onbi;  // probe ints, stencil cell/face pairs
    int nobs;

    static DEV int nobs_of(const DevMesh &m) { return m.nobs; }
    DEV Draw(const DevMesh &m, const DevSettings &s, const double *coeffs, double *A, Book *b,
             double *redbuf, Krylov *kr, double *amg, double *aux = nullptr)
        : M(m), S(s), c(coeffs), bk(b), kres(kr), amgw(amg) {
        red.buf = redbuf;
        red.parity = 0;
        nc = M.nc;
        nif = M.nif;
        nf = M.nf;
        nb = M.nb;
        nbs = nb > 1 ? nb : 1;
        planar = M.planar;
        has_outlet = M.has_outlet;
        const ArenaLayout &L = M.lay;
        Cb = (D *)A;
        Fb = (D *)(A + L.face);
        Bb = (D *)(A + L.bnd);
        const int *ib = reinterpret_cast<const int *>(A + L.ibase);
        slot4 = (typename Space<kLDS>::I4 *)(ib + L.islot);
        fon = (typename Space<kLDS>::I2 *)(ib + L.iface);
        bflag = (typename Space<kLDS>::I *)(ib + L.ibflag);
        if constexpr (kGrid) {
            ringb = (DL *)aux;
            q = ringb + 8 * (int64_t)S.convergence_window;
            qprev = q + (nobs_of(M) > 0 ? nobs_of(M) : 1);
        } else {
            ringb = (DL *)(A + L.ring);
            q = (DL *)(A + L.q);
            qprev = (DL *)(A + L.qprev);
        }
        Ob = (D *)(A + L.obs);
        Onb = (D *)(A + L.onb);
        obi = (typename Space<kLDS>::I *)(ib + L.iobs);
        onbi = (typename Space<kLDS>::I *)(ib + L.ionb);
        nobs = M.nobs;

    }

    // record accessors: cell, face and boundary-face variables
    // (record indices stay below 2^23, so 24-bit multiplies are exact)
    // (multi-workgroup draws: coherent views, since any record may have been
    // written by another workgroup since the last barrier)
    using Ref = typename std::conditional<kGrid, GRef, D &>::type;
    DEV Ref at(D *base, int k) const {
        if constexpr (kGrid) return GRef{(GlobalDouble *)(base + k)};
        else return base[k];
    }
    // LDS arenas hold records (odd strides spread them over the banks); HBM
    // arenas hold one array per variable (nc, nf and max(nb, 1) long), so a
    // wave's accesses to one variable of consecutive cells are coalesced
    DEV Ref C(int i, int v) const {
        if constexpr (kLDS) return at(Cb, __mul24(i, kNA) + v);
        else return at(Cb, v * nc + i);
    }
    DEV Ref Fc(int f, int v) const {
        if constexpr (kLDS) return at(Fb, __mul24(f, kNF) + v);
        else return at(Fb, v * nf + f);
    }
    DEV Ref Bd(int b, int v) const {
        if constexpr (kLDS) return at(Bb, __mul24(b, kNB) + v);
        else return at(Bb, v * nbs + b);
    }

    // off-diagonal coefficient of slot s of cell ci
    DEV Ref off(int ci, int s) const { return C(ci, V_off0 + s); }

    // multigrid workspace arrays (global memory, shared by the workgroups of
    // a multi-workgroup draw)
    using WPtr = typename std::conditional<kGrid, GPtr, double *>::type;
    DEV WPtr wp(int64_t o) const {
        if constexpr (kGrid) return GPtr{(GlobalDouble *)(amgw + o)};
        else return amgw + o;
    }

    // ---------------------------------------------------------------- SST
    // SSTModel pointwise functions (core/src/SSTModel.cpp:15-108)
    DEV double sst_f1(double kk, double omega, double y, double CDkwv) const {
        const double nu = S.nu;
        double sqrtK = sqrt(smax(kk, 0.0));
        double omSafe = smax(omega, 1e-20);
        double ySafe = smax(y, 1e-20);
        double y2s = ySafe * ySafe;
        double term1 = sqrtK / (c[8] * omSafe * ySafe);
        double term2 = 500.0 * nu / (y2s * omSafe);
        double CDpos = smax(CDkwv, 1e-10);
        double term3 = 4.0 * c[5] * smax(kk, 0.0) / (CDpos * y2s);
        double arg1 = smin(smax(term1, term2), term3);
        double a4 = arg1 * arg1 * arg1 * arg1;
        return tanh(a4);
    }
    DEV double sst_f2(double kk, double omega, double y) const {
        const double nu = S.nu;
        double sqrtK = sqrt(smax(kk, 0.0));
        double omSafe = smax(omega, 1e-20);
        double ySafe = smax(y, 1e-20);
        double y2 = ySafe * ySafe;
        double term1 = 2.0 * sqrtK / (c[8] * omSafe * ySafe);
        double term2 = 500.0 * nu / (y2 * omSafe);
        double arg2 = smax(term1, term2);
        return tanh(arg2 * arg2);
    }
    DEV double sst_nut(double kk, double omega, double Sv, double f2) const {
        double kSafe = smax(kk, 0.0);
        if (S.sst_variant == 1) return kSafe / smax(omega, 1e-20);
        double denom = smax(c[9] * omega, Sv * f2);
        denom = smax(denom, 1e-20);
        return c[9] * kSafe / denom;
    }
    DEV double sst_pk(double nt, double Sv, double kk, double omega) const {
        double raw = nt * Sv * Sv;
        double lim = 10.0 * c[8] * smax(kk, 0.0) * smax(omega, 1e-20);
        return smin(raw, lim);
    }
    DEV double sst_pomega(double nt, double Sv, double kk, double omega) const {
        double S2 = Sv * Sv;
        double lim = 10.0 * c[8] * smax(kk, 0.0) * smax(omega, 1e-20) / smax(nt, 1e-30);
        return smin(S2, lim);
    }
    DEV double blend(double a, double b, double f1) const { return f1 * a + (1.0 - f1) * b; }
    // SIMPLESolver::resolvedWallOmega (SIMPLESolver.cpp:155-161)
    DEV double resolved_wall_omega(double y) const {
        const double coefficient = planar ? 6.0 : 60.0;
        return coefficient * S.nu / (c[2] * y * y);
    }

    // ----------------------------------------------------------- boundary
    // applyVelocityBC / applyPressureBC / applyKBC / applyOmegaBC
    // (core/include/BoundaryCondition.hpp:279-384)
    DEV void bc_velocity() {
        for (int b = gt; b < nb; b += gs) {
            const int f = nif + b, o = fon[f].x, fl = bflag[b];
            switch (bc_type(fl, 0)) {
            case BC_WALL_NOSLIP:
                Bd(b, B_Ux) = 0.0;
                Bd(b, B_Uy) = 0.0;
                break;
            case BC_WALL_MOVING:
            case BC_INLET:
            case BC_DIRICHLET:
                Bd(b, B_Ux) = Bd(b, B_uvx);
                Bd(b, B_Uy) = Bd(b, B_uvy);
                break;
            case BC_SYMMETRY: {
                const double nx = Fc(f, W_nx), ny = Fc(f, W_ny);
                const double d = C(o, V_Ux) * nx + C(o, V_Uy) * ny;
                Bd(b, B_Ux) = C(o, V_Ux) - nx * d;
                Bd(b, B_Uy) = C(o, V_Uy) - ny * d;
                break;
            }
            default:
                Bd(b, B_Ux) = C(o, V_Ux);
                Bd(b, B_Uy) = C(o, V_Uy);
            }
        }
    }
    DEV void bc_pressure() {
        for (int b = gt; b < nb; b += gs) {
            const int f = nif + b, t = bc_type(bflag[b], 1);
            Bd(b, B_p) = (t == BC_OUTLET || t == BC_DIRICHLET) ? Bd(b, B_pv) : C(fon[f].x, V_p);
        }
    }
    DEV void bc_k() {
        for (int b = gt; b < nb; b += gs) {
            const int f = nif + b, t = bc_type(bflag[b], 2);
            Bd(b, B_k) = t == BC_WALL_KOMEGA ? 0.0 : (t == BC_DIRICHLET ? Bd(b, B_kv) : C(fon[f].x, V_k));
        }
    }
    DEV void bc_omega(double beta1) {
        for (int b = gt; b < nb; b += gs) {
            const int f = nif + b, t = bc_type(bflag[b], 3);
            double v;
            if (t == BC_WALL_KOMEGA) {
                const double y1 = Fc(f, W_dlt);
                v = 60.0 * S.nu / (beta1 * y1 * y1);
            } else if (t == BC_DIRICHLET) {
                v = Bd(b, B_wv);
            } else {
                v = C(fon[f].x, V_w);
            }
            Bd(b, B_w) = v;
        }
    }

    DEV void slots(int ci, int (&code)[kSlots], int (&nb_)[kSlots]) const {
        const v4i a = slot4[2 * ci], b = slot4[2 * ci + 1];
        code[0] = a.x; code[1] = a.y; code[2] = a.z; code[3] = a.w;
        nb_[0] = b.x; nb_[1] = b.y; nb_[2] = b.z; nb_[3] = b.w;
    }

    // A cell's slots as a branch-free set: every index is clamped to a record
    // that exists (face 0, boundary face 0, the cell itself), so a row loads
    // the data of all its slots at once and the slots' arithmetic interleaves;
    // the contribution of a slot that does not apply is dropped with a select.
    // Every sum keeps its slot order and its operations, so the results are
    // bitwise those of a loop that branches per slot. Internal faces need no
    // owner/neighbour lookup: the face's other cell is the slot's neighbour,
    // and the slot side says which of the two is the owner.
    struct SlotSet {
        int f[kSlots], b[kSlots], n[kSlots];
        bool valid[kSlots], inner[kSlots], own[kSlots];
    };
    DEV void slot_set(int ci, SlotSet &q) const {
        int code[kSlots], nbv[kSlots];
        slots(ci, code, nbv);
#pragma unroll
        for (int s = 0; s < kSlots; ++s) {
            q.valid[s] = code[s] >= 0;
            const int f = q.valid[s] ? slot_face(code[s]) : 0;
            q.f[s] = f;
            q.inner[s] = nbv[s] >= 0;
            q.own[s] = slot_own(code[s]);
            q.b[s] = (q.valid[s] && !q.inner[s]) ? f - nif : 0;
            q.n[s] = q.inner[s] ? nbv[s] : ci;
        }
    }

    // ----------------------------------------------------------- gradients
    // greenGaussGrad (core/include/Field.hpp:178-212) of NF fields at one
    // cell, its faces gathered in face order.
    template <int NF>
    DEV void gg_cell(int ci, const int (&ph)[NF], const int (&phb)[NF], double (&gx)[NF],
                     double (&gy)[NF]) const {
        SlotSet q;
        slot_set(ci, q);
        double xc[NF];
#pragma unroll
        for (int i = 0; i < NF; ++i) {
            gx[i] = gy[i] = 0.0;
            xc[i] = C(ci, ph[i]);
        }
#pragma unroll
        for (int s = 0; s < kSlots; ++s) {
            const int f = q.f[s];
            const double Sx = Fc(f, W_Sx), Sy = Fc(f, W_Sy);
            const double wt = Fc(f, W_w), wmv = Fc(f, W_wm);
            const bool own = q.own[s], in = q.inner[s], on = q.valid[s];
            // boundary values only where the slot is a boundary face (few
            // cells have one): a masked load, with no arithmetic under it
            double pb[NF];
#pragma unroll
            for (int i = 0; i < NF; ++i) pb[i] = 0.0;
            if (on && !in)
#pragma unroll
                for (int i = 0; i < NF; ++i) pb[i] = Bd(q.b[s], phb[i]);
#pragma unroll
            for (int i = 0; i < NF; ++i) {
                const double xn = C(q.n[s], ph[i]);
                const double pbv = pb[i];
                // internal: owner value * w + neighbour value * (1 - w)
                const double phiF = (own ? xc[i] : xn) * wt + (own ? xn : xc[i]) * wmv;
                const double v = in ? phiF : pbv;
                const bool add = !in || own;
                const double ax = add ? gx[i] + Sx * v : gx[i] - Sx * v;
                const double ay = add ? gy[i] + Sy * v : gy[i] - Sy * v;
                gx[i] = on ? ax : gx[i];
                gy[i] = on ? ay : gy[i];
            }
        }
        const double V = C(ci, V_vol);
        if (V > 1e-30)
#pragma unroll
            for (int i = 0; i < NF; ++i) {
                gx[i] = gx[i] / V;
                gy[i] = gy[i] / V;
            }
    }

    // U_f . S_f (SIMPLESolver.cpp:842-852)
    DEV double interp_flux(int f) const {
        if (f < nif) {
            const v2i on = fon[f];
            const double wt = Fc(f, W_w), wmv = Fc(f, W_wm);
            const double ufx = C(on.x, V_Ux) * wt + C(on.y, V_Ux) * wmv;
            const double ufy = C(on.x, V_Uy) * wt + C(on.y, V_Uy) * wmv;
            return (ufx * Fc(f, W_nx) + ufy * Fc(f, W_ny)) * Fc(f, W_A);
        }
        const int b = f - nif;
        return (Bd(b, B_Ux) * Fc(f, W_nx) + Bd(b, B_Uy) * Fc(f, W_ny)) * Fc(f, W_A);
    }

    // F (phi_f^HO - phi_U) of the deferred convection correction at an
    // internal face (SIMPLESolver.cpp:495-532), 0 when the face flux is 0.
    // The host stores x_f - x_U in dO/dN for linear upwind and x_D - x_U
    // for the limited scheme. Branch-free over the slot: the values of the
    // cell (c) and of the slot's neighbour (n) are loaded, own says whether
    // the cell is the face's owner.
    DEV double convection_corr(int f, bool own, double Fl, double phc, double gxc, double gyc,
                               int n, int ph, int gx, int gy, int scheme) const {
        const bool ownerUp = Fl > 0.0;
        const bool upIsCell = own == ownerUp;  // the upwind cell is this cell
        const double phn = C(n, ph), gxn = C(n, gx), gyn = C(n, gy);
        const double dOx = Fc(f, W_dOx), dOy = Fc(f, W_dOy), dNx = Fc(f, W_dNx), dNy = Fc(f, W_dNy);
        const double dx = ownerUp ? dOx : dNx;
        const double dy = ownerUp ? dOy : dNy;
        const double pu = upIsCell ? phc : phn;
        const double gux = upIsCell ? gxc : gxn, guy = upIsCell ? gyc : gyn;
        double phiF;
        if (scheme == CONV_LINEAR_UPWIND) {
            phiF = pu + (gux * dx + guy * dy);
        } else {
            const double pd = upIsCell ? phn : phc;
            const double dphi = pd - pu;
            const double r = 2.0 * (gux * dx + guy * dy) / dphi - 1.0;
            const double psiR = r > 0.0 ? (r * r + r) / (r * r + 1.0) : 0.0;
            const double psi = fabs(dphi) > 1e-300 ? psiR : 0.0;
            phiF = pu + 0.5 * psi * dphi;
        }
        return Fl == 0.0 ? 0.0 : Fl * (phiF - pu);
    }

    // y = A x for row ci (LinearSystem::matvec order: diagonal, then faces)
    DEV double spmv(int ci, int x) const {
        const v4i nb4 = slot4[2 * ci + 1];
        const int nbv[4] = {nb4.x, nb4.y, nb4.z, nb4.w};
        double y = C(ci, V_diag) * C(ci, x);
#pragma unroll
        for (int s = 0; s < kSlots; ++s) {
            // every slot's loads issue together; unused slots are dropped
            const double a = off(ci, s), xn = C(nbv[s] >= 0 ? nbv[s] : ci, x);
            y = nbv[s] >= 0 ? y + a * xn : y;
        }
        return y;
    }

    DEV bool row_finite(int ci, double dg, double sr) const {
        const v4i nb4 = slot4[2 * ci + 1];
        const int nbv[4] = {nb4.x, nb4.y, nb4.z, nb4.w};
        bool ok = isfinite(dg) && isfinite(sr);
#pragma unroll
        for (int s = 0; s < kSlots; ++s) {
            const double a = off(ci, s);
            ok = ok && (nbv[s] < 0 || isfinite(a));
        }
        return ok;
    }

    // ------------------------------------------------------ momentum row
    // SIMPLESolver::assembleMomentum (SIMPLESolver.cpp:303-482) for one cell;
    // returns the row's residual at the current component values, the
    // unrelaxed diagonal in aPu and the admission flag in ok.
    DEV double momentum_row(int ci, int comp, double &aPu, bool &ok) {
        const double nu = S.nu;
        const int Ubv = comp == 0 ? B_Ux : B_Uy;
        const int Uc = comp == 0 ? V_Ux : V_Uy;
        SlotSet q;
        slot_set(ci, q);
        double dg = 0.0, sr = 0.0, offv[kSlots] = {0.0, 0.0, 0.0, 0.0};
        const double nuTc = C(ci, V_nuT);
        double Sf[kSlots], mFlux[kSlots], wt[kSlots], wmv[kSlots], nuTn[kSlots];
        bool wall[kSlots];
#pragma unroll
        for (int s = 0; s < kSlots; ++s) {
            const int f = q.f[s];
            Sf[s] = Fc(f, W_A);
            const double delta = Fc(f, W_dlt);
            mFlux[s] = Fc(f, W_phi);
            wt[s] = Fc(f, W_w);
            wmv[s] = Fc(f, W_wm);
            nuTn[s] = C(q.n[s], V_nuT);
            wall[s] = bflag[q.b[s]] & kFlagWall;
            const double bU = Bd(q.b[s], Ubv);
            const bool own = q.own[s], in = q.inner[s];
            // internal face: nuEff interpolated from owner and neighbour;
            // boundary face: nu at a wall, nu + nuT of the cell elsewhere
            const double nuEff_o = nu + (own ? nuTc : nuTn[s]);
            const double nuEff_n = nu + (own ? nuTn[s] : nuTc);
            const double nuEff_f = wt[s] * nuEff_o + wmv[s] * nuEff_n;
            const double nuEffB = wall[s] ? nu : nu + nuTc;
            const double D = (in ? nuEff_f : nuEffB) * Sf[s] / delta;
            const double m = mFlux[s];
            const double cPos = smax(m, 0.0), cNeg = smax(-m, 0.0);
            const double dIn = own ? D + cPos : D + cNeg;
            const double oIn = -(own ? D + cNeg : D + cPos);
            const bool pos = m >= 0;
            const double dB = pos ? m + D : D;
            const double sB = (pos ? D : D - m) * bU;
            dg = q.valid[s] ? dg + (in ? dIn : dB) : dg;
            sr = (q.valid[s] && !in) ? sr + sB : sr;
            offv[s] = in ? oIn : 0.0;
        }
        if constexpr (!kGrid) if (M.folded) folded_faces(ci, nu + nuTc, C(ci, Uc), dg, sr);
        const double V = C(ci, V_vol);
        const double dpdx = comp == 0 ? C(ci, V_gPx) : C(ci, V_gPy);
        sr -= dpdx * V;
        const double fb = comp == 0 ? S.body_force_x : S.body_force_y;
        if (fb != 0.0) sr += fb * V;
        // transpose stress div(nuT (grad U)^T), zero wall coefficient
        // (core/include/StressOperators.hpp:31-81)
        {
            const int ga = comp == 0 ? V_gUxx : V_gUxy;
            const int gb = comp == 0 ? V_gUyx : V_gUyy;
            const double gac = C(ci, ga), gbc = C(ci, gb);
            double tsrc = 0.0;
#pragma unroll
            for (int s = 0; s < kSlots; ++s) {
                const int f = q.f[s];
                const bool own = q.own[s], in = q.inner[s];
                const double gan_ = C(q.n[s], ga), gbn_ = C(q.n[s], gb);
                const double nx = Fc(f, W_nx), ny = Fc(f, W_ny);
                const double nuo = own ? nuTc : nuTn[s], nun = own ? nuTn[s] : nuTc;
                const double gao = own ? gac : gan_, gan = own ? gan_ : gac;
                const double gbo = own ? gbc : gbn_, gbn = own ? gbn_ : gbc;
                const double cfI = wt[s] * nuo + wmv[s] * nun;
                const double gxf = gao * wt[s] + gan * wmv[s];
                const double gyf = gbo * wt[s] + gbn * wmv[s];
                const double cf = in ? cfI : (wall[s] ? 0.0 : nuTc);
                const double gX = in ? gxf : gac, gY = in ? gyf : gbc;
                const double flux = cf * (gX * nx + gY * ny) * Sf[s];
                const double t = (!in || own) ? tsrc + flux : tsrc - flux;
                tsrc = q.valid[s] ? t : tsrc;
            }
            sr += tsrc;
        }
        const double uc = C(ci, Uc);
        if (S.momentum_convection != CONV_UPWIND) {
            const int gx = comp == 0 ? V_gUxx : V_gUyx;
            const int gy = comp == 0 ? V_gUxy : V_gUyy;
            const double gxc = C(ci, gx), gyc = C(ci, gy);
#pragma unroll
            for (int s = 0; s < kSlots; ++s) {
                const double corr = convection_corr(q.f[s], q.own[s], mFlux[s], uc, gxc, gyc,
                                                    q.n[s], Uc, gx, gy, S.momentum_convection);
                const double t = q.own[s] ? sr - corr : sr + corr;
                sr = q.inner[s] ? t : sr;
            }
        }
        const double alphaU = S.alpha_u;
        aPu = dg;
        sr += (1.0 - alphaU) / alphaU * dg * uc;
        dg /= alphaU;
        C(ci, V_diag) = dg;
        C(ci, V_src) = sr;
        if (comp == 0) C(ci, V_aPs) = dg;
        ok = isfinite(dg) && isfinite(sr) && isfinite(uc);
        double y = dg * uc;
#pragma unroll
        for (int s = 0; s < kSlots; ++s) {
            const double xn = C(q.n[s], Uc);
            if (q.inner[s]) off(ci, s) = offv[s];
            ok = ok && (!q.inner[s] || isfinite(offv[s]));
            y = q.inner[s] ? y + offv[s] * xn : y;
        }
        return sr - y;
    }

    // The front and back faces of a one-layer extruded mesh, folded out of
    // the slots by the host: each is a zero-gradient boundary face with no
    // flux, so it adds D = nuEff S / delta to the diagonal and D phi_b to the
    // source of its owner's row, phi_b being the cell's current value (the
    // boundary-face loops of SIMPLESolver.cpp:360-382, 983-1001, 1092-1113,
    // after the cell's in-plane faces). nuEff is the row's boundary-face
    // diffusivity. Not compiled into the multi-workgroup kernels; the host
    // runs such meshes on one workgroup per draw.
    DEV void folded_faces(int ci, double nuEff, double phi, double &dg, double &sr) const {
        const double *fp = M.folded + 4 * (int64_t)ci;
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            const double A = fp[2 * e], dl = fp[2 * e + 1];
            const double D = nuEff * A / dl;
            dg = A > 0.0 ? dg + D : dg;
            sr = A > 0.0 ? sr + D * phi : sr;
        }
    }

    // ---------------------------------------------------------- k / omega rows
    // Diffusion and upwind convection of one k or omega row over its slots
    // (the face loops of assembleKEquation / assembleOmegaEquation): sig0 and
    // sig1 are the inner and outer diffusion coefficients blended by F1,
    // Bv the boundary variable. Fills dg, sr and offv; Sf, delta, nuEff of
    // internal faces and the face fluxes are returned for the callers' other
    // slot terms.
    struct TurbSlots {
        double Sf[kSlots], delta[kSlots], mFlux[kSlots], nuEff[kSlots];
    };
    DEV void turb_faces(int ci, const SlotSet &q, double sig0, double sig1, int Bv, int Cv,
                        double F1c, double nuTc, double &dg, double &sr, double (&offv)[kSlots],
                        TurbSlots &t) const {
        const double nu = S.nu;
        // boundary faces: the cell's blended coefficient
        const double sgB = blend(sig0, sig1, F1c);
        const double nuEffB = nu + sgB * nuTc;
#pragma unroll
        for (int s = 0; s < kSlots; ++s) {
            const int f = q.f[s];
            const double Sf = Fc(f, W_A);
            const double delta = Fc(f, W_dlt);
            const double m = Fc(f, W_phi);
            const double wt = Fc(f, W_w), wmv = Fc(f, W_wm);
            const double F1n = C(q.n[s], V_F1), ntn_ = C(q.n[s], V_nuT);
            const double bv = Bd(q.b[s], Bv);
            const bool own = q.own[s], in = q.inner[s];
            const double F1o = own ? F1c : F1n, F1nb = own ? F1n : F1c;
            const double nto = own ? nuTc : ntn_, ntn = own ? ntn_ : nuTc;
            const double F1f = wt * F1o + wmv * F1nb;
            const double sg = blend(sig0, sig1, F1f);
            const double nuEff_o = nu + sg * nto;
            const double nuEff_n = nu + sg * ntn;
            const double nuEff_f = wt * nuEff_o + wmv * nuEff_n;
            const double D = (in ? nuEff_f : nuEffB) * Sf / delta;
            const double cPos = smax(m, 0.0), cNeg = smax(-m, 0.0);
            const double dIn = own ? D + cPos : D + cNeg;
            const double oIn = -(own ? D + cNeg : D + cPos);
            const bool pos = m >= 0;
            const double dB = pos ? m + D : D;
            const double sB = (pos ? D : D - m) * bv;
            dg = q.valid[s] ? dg + (in ? dIn : dB) : dg;
            sr = (q.valid[s] && !in) ? sr + sB : sr;
            offv[s] = in ? oIn : 0.0;
            t.Sf[s] = Sf;
            t.delta[s] = delta;
            t.mFlux[s] = m;
            t.nuEff[s] = nuEff_f;
        }
        if constexpr (!kGrid) if (M.folded) folded_faces(ci, nuEffB, C(ci, Cv), dg, sr);
    }

    // deferred convection correction of a k or omega row (internal slots)
    DEV double turb_convection(int ci, const SlotSet &q, const TurbSlots &t, double sr, int ph,
                               int gx, int gy) const {
        const double phc = C(ci, ph), gxc = C(ci, gx), gyc = C(ci, gy);
#pragma unroll
        for (int s = 0; s < kSlots; ++s) {
            const double corr = convection_corr(q.f[s], q.own[s], t.mFlux[s], phc, gxc, gyc,
                                                q.n[s], ph, gx, gy, S.turbulence_convection);
            const double v = q.own[s] ? sr - corr : sr + corr;
            sr = q.inner[s] ? v : sr;
        }
        return sr;
    }

    // SIMPLESolver::assembleKEquation (SIMPLESolver.cpp:947-1023); on return
    // kr holds the row's residual b - A k.
    DEV void k_row(int ci, bool &ok, double &resid) {
        SlotSet q;
        slot_set(ci, q);
        double dg = 0.0, sr = 0.0, offv[kSlots] = {0.0, 0.0, 0.0, 0.0};
        const double F1c = C(ci, V_F1), nuTc = C(ci, V_nuT);
        TurbSlots t;
        turb_faces(ci, q, c[0], c[4], B_k, V_k, F1c, nuTc, dg, sr, offv, t);
        const double V = C(ci, V_vol);
        sr += C(ci, V_Pk) * V;
        const double destruction = c[8] * smax(C(ci, V_w), 1e-20);
        dg += destruction * V;
        if (S.turbulence_convection != CONV_UPWIND) sr = turb_convection(ci, q, t, sr, V_k, V_gKx, V_gKy);
        const double aK = S.alpha_k;
        const double kc = C(ci, V_k);
        sr += (1.0 - aK) / aK * dg * kc;
        dg /= aK;
        C(ci, V_diag) = dg;
        C(ci, V_src) = sr;
        ok = isfinite(dg) && isfinite(sr) && isfinite(kc);
        // b - A x at x = k (LinearSystem::matvec order: diagonal, then faces)
        double y = dg * kc;
#pragma unroll
        for (int s = 0; s < kSlots; ++s) {
            const double xn = C(q.n[s], V_k);
            if (q.inner[s]) off(ci, s) = offv[s];
            ok = ok && (!q.inner[s] || isfinite(offv[s]));
            y = q.inner[s] ? y + offv[s] * xn : y;
        }
        resid = sr - y;
    }

    // SIMPLESolver::assembleOmegaEquation (SIMPLESolver.cpp:1030-1169) and
    // constrainOmega (SIMPLESolver.cpp:686-704); returns the constrained
    // initial value x0 of the omega solve and the row's residual b - A x0
    // (every cell's x0 is its wall value where it has one, else omega, so
    // the neighbours' x0 are known without a barrier).
    DEV double omega_row(int ci, bool &ok, double &resid) {
        const double nu = S.nu;
        SlotSet q;
        slot_set(ci, q);
        double dg = 0.0, sr = 0.0, balance = 0.0, offv[kSlots] = {0.0, 0.0, 0.0, 0.0};
        const double amplitude = 6 * nu / c[2];
        const double F1c = C(ci, V_F1), nuTc = C(ci, V_nuT);
        TurbSlots t;
        turb_faces(ci, q, c[1], c[5], B_w, V_w, F1c, nuTc, dg, sr, offv, t);
        if (planar) {
            const double omsc = C(ci, V_oms);
#pragma unroll
            for (int s = 0; s < kSlots; ++s) {
                const bool own = q.own[s];
                const double omsn = C(q.n[s], V_oms);
                const double omg = Fc(q.f[s], W_omg);
                const double so = own ? omsc : omsn, sn = own ? omsn : omsc;
                const double difference = (sn - so) / t.delta[s];
                const double nuEff_f = t.nuEff[s];
                const double correction =
                    amplitude * t.Sf[s] * ((nuEff_f - nu) * omg - nuEff_f * difference);
                const double v = own ? balance + correction : balance - correction;
                balance = q.inner[s] ? v : balance;
            }
        }
        const double V = C(ci, V_vol);
        const double wc = C(ci, V_w);
        const double omC = smax(wc, 1e-20);
        const double alphaB = blend(c[3], c[7], F1c);
        const double production = sst_pomega(nuTc, C(ci, V_S), C(ci, V_k), wc);
        sr += alphaB * production * V;
        const double betaB = blend(c[2], c[6], F1c);
        dg += betaB * omC * V;
        sr += (1.0 - F1c) * C(ci, V_CDkw) * V;
        if (planar) {
            const double correction = balance + nu * (6 * nu / c[2]) * C(ci, V_oml) * V;
            if (correction < 0) dg -= correction / omC;
            else sr += correction;
        }
        if (S.turbulence_convection != CONV_UPWIND) sr = turb_convection(ci, q, t, sr, V_w, V_gWx, V_gWy);
        const double aW = S.alpha_omega;
        sr += (1.0 - aW) / aW * dg * wc;
        dg /= aW;
        const double wvc = C(ci, V_wallv);
        double x0n[kSlots];
#pragma unroll
        for (int s = 0; s < kSlots; ++s) {
            const bool own = q.own[s], in = q.inner[s];
            const double vn_ = C(q.n[s], V_wallv);
            const double wn_ = C(q.n[s], V_w);
            x0n[s] = isfinite(vn_) ? vn_ : wn_;
            const double vo = own ? wvc : vn_, vn = own ? vn_ : wvc;
            const bool fo = isfinite(vo), fn = isfinite(vn);
            // a fixed value on the owner side (else the neighbour side) moves
            // the coefficient into the source of the other side's row
            const bool sub = in && (fo ? !own : (fn && own));
            const double v = fo ? vo : vn;
            sr = sub ? sr - offv[s] * v : sr;
            offv[s] = (in && (fo || fn)) ? 0.0 : offv[s];
        }
        if (isfinite(wvc)) sr = dg * wvc;
        C(ci, V_diag) = dg;
        C(ci, V_src) = sr;
        const double x0 = isfinite(wvc) ? wvc : wc;
        ok = isfinite(dg) && isfinite(sr) && isfinite(x0);
        double y = dg * x0;
#pragma unroll
        for (int s = 0; s < kSlots; ++s) {
            if (q.inner[s]) off(ci, s) = offv[s];
            ok = ok && (!q.inner[s] || isfinite(offv[s]));
            y = q.inner[s] ? y + offv[s] * x0n[s] : y;
        }
        resid = sr - y;
        return x0;
    }

    // ------------------------------------------------ several lanes per cell
    // The 256-thread LDS kernels assemble the momentum, k and omega rows with
    // QL = 2 lanes per cell: lane qi of the pair evaluates the face terms of
    // slots qi and qi + 2, the pair shares them with DPP broadcasts, and each
    // lane adds them in slot order with the same operations as the one-lane
    // rows above, so a row's values are bitwise the same. The row's own
    // values are written by lane 0 of the pair; each lane writes the
    // off-diagonals of its slots.
    static constexpr int NJ = kSlots / QL;  // slots per lane
    template <int L>
    static DEV int lane_bc(int v) {  // lane L of the cell's group, in every lane of it
        constexpr int ctrl = QL == 4 ? (L | (L << 2) | (L << 4) | (L << 6))
                                     : (L | (L << 2) | ((2 + L) << 4) | ((2 + L) << 6));
        return __builtin_amdgcn_mov_dpp(v, ctrl, 0xf, 0xf, false);
    }
    template <int L>
    static DEV double lane_bcd(double v) {
        const int lo = lane_bc<L>(__double2loint(v));
        const int hi = lane_bc<L>(__double2hiint(v));
        return __hiloint2double(hi, lo);
    }
    // the per-slot values a of every slot from each lane's own slots v
    // (slot s lives in lane s % QL as its entry s / QL)
    template <int S>
    static DEV double slot_of(const double (&v)[NJ]) { return lane_bcd<S % QL>(v[S / QL]); }
    template <int S>
    static DEV int slot_ofi(const int (&v)[NJ]) { return lane_bc<S % QL>(v[S / QL]); }
    static DEV void gatherq(const double (&v)[NJ], double (&a)[kSlots]) {
        a[0] = slot_of<0>(v);
        a[1] = slot_of<1>(v);
        a[2] = slot_of<2>(v);
        a[3] = slot_of<3>(v);
    }
    static DEV void gatherqi(const int (&v)[NJ], int (&a)[kSlots]) {
        a[0] = slot_ofi<0>(v);
        a[1] = slot_ofi<1>(v);
        a[2] = slot_ofi<2>(v);
        a[3] = slot_ofi<3>(v);
    }
    template <class T>
    static DEV T pick(const T (&a)[kSlots], int s) {
        return s == 0 ? a[0] : s == 1 ? a[1] : s == 2 ? a[2] : a[3];
    }
    // this lane's slots (qi + QL j) of a cell's slot set
    struct OwnSlots {
        int s[NJ], f[NJ], b[NJ], n[NJ];
        bool valid[NJ], in[NJ], own[NJ];
    };
    DEV void own_slots(const SlotSet &q, OwnSlots &o) const {
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            const int s = qi + QL * j;
            o.s[j] = s;
            o.f[j] = pick(q.f, s);
            o.b[j] = pick(q.b, s);
            o.n[j] = pick(q.n, s);
            o.valid[j] = pick(q.valid, s);
            o.in[j] = pick(q.inner, s);
            o.own[j] = pick(q.own, s);
        }
    }
    // y = diag x + sum over internal slots of off x_n and the row's
    // finiteness (the residual loop of every row); writes this lane's
    // off-diagonals
    DEV double row_residual_q(int ci, const SlotSet &q, const OwnSlots &o,
                              const double (&offv)[NJ], const double (&xn)[NJ], double y,
                              bool &ok) {
        double pr[NJ];
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            if (o.in[j]) off(ci, o.s[j]) = offv[j];
            pr[j] = offv[j] * xn[j];
        }
        double of4[kSlots], pr4[kSlots];
        gatherq(offv, of4);
        gatherq(pr, pr4);
#pragma unroll
        for (int s = 0; s < kSlots; ++s) {
            ok = ok && (!q.inner[s] || isfinite(of4[s]));
            y = q.inner[s] ? y + pr4[s] : y;
        }
        return y;
    }

    // momentum_row with QL lanes per cell (the same operations, in the same order)
    DEV double momentum_row_q(int ci, int comp, double &aPu, bool &ok) {
        const double nu = S.nu;
        const int Ubv = comp == 0 ? B_Ux : B_Uy;
        const int Uc = comp == 0 ? V_Ux : V_Uy;
        SlotSet q;
        slot_set(ci, q);
        OwnSlots o;
        own_slots(q, o);
        const double nuTc = C(ci, V_nuT);
        double Sf[NJ], mFlux[NJ], wt[NJ], wmv[NJ], nuTn[NJ], offv[NJ], cd[NJ], cb[NJ];
        bool wall[NJ];
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            const int f = o.f[j];
            Sf[j] = Fc(f, W_A);
            const double delta = Fc(f, W_dlt);
            mFlux[j] = Fc(f, W_phi);
            wt[j] = Fc(f, W_w);
            wmv[j] = Fc(f, W_wm);
            nuTn[j] = C(o.n[j], V_nuT);
            wall[j] = bflag[o.b[j]] & kFlagWall;
            const double bU = Bd(o.b[j], Ubv);
            const bool own = o.own[j], in = o.in[j];
            const double nuEff_o = nu + (own ? nuTc : nuTn[j]);
            const double nuEff_n = nu + (own ? nuTn[j] : nuTc);
            const double nuEff_f = wt[j] * nuEff_o + wmv[j] * nuEff_n;
            const double nuEffB = wall[j] ? nu : nu + nuTc;
            const double D = (in ? nuEff_f : nuEffB) * Sf[j] / delta;
            const double m = mFlux[j];
            const double cPos = smax(m, 0.0), cNeg = smax(-m, 0.0);
            const double dIn = own ? D + cPos : D + cNeg;
            const double oIn = -(own ? D + cNeg : D + cPos);
            const bool pos = m >= 0;
            const double dB = pos ? m + D : D;
            cb[j] = (pos ? D : D - m) * bU;
            cd[j] = in ? dIn : dB;
            offv[j] = in ? oIn : 0.0;
        }
        double cd4[kSlots], cb4[kSlots];
        gatherq(cd, cd4);
        gatherq(cb, cb4);
        double dg = 0.0, sr = 0.0;
#pragma unroll
        for (int s = 0; s < kSlots; ++s) {
            dg = q.valid[s] ? dg + cd4[s] : dg;
            sr = (q.valid[s] && !q.inner[s]) ? sr + cb4[s] : sr;
        }
        if (M.folded) folded_faces(ci, nu + nuTc, C(ci, Uc), dg, sr);
        const double V = C(ci, V_vol);
        const double dpdx = comp == 0 ? C(ci, V_gPx) : C(ci, V_gPy);
        sr -= dpdx * V;
        const double fb = comp == 0 ? S.body_force_x : S.body_force_y;
        if (fb != 0.0) sr += fb * V;
        {
            const int ga = comp == 0 ? V_gUxx : V_gUxy;
            const int gb = comp == 0 ? V_gUyx : V_gUyy;
            const double gac = C(ci, ga), gbc = C(ci, gb);
            double fl[NJ];
#pragma unroll
            for (int j = 0; j < NJ; ++j) {
                const int f = o.f[j];
                const bool own = o.own[j], in = o.in[j];
                const double gan_ = C(o.n[j], ga), gbn_ = C(o.n[j], gb);
                const double nx = Fc(f, W_nx), ny = Fc(f, W_ny);
                const double nuo = own ? nuTc : nuTn[j], nun = own ? nuTn[j] : nuTc;
                const double gao = own ? gac : gan_, gan = own ? gan_ : gac;
                const double gbo = own ? gbc : gbn_, gbn = own ? gbn_ : gbc;
                const double cfI = wt[j] * nuo + wmv[j] * nun;
                const double gxf = gao * wt[j] + gan * wmv[j];
                const double gyf = gbo * wt[j] + gbn * wmv[j];
                const double cf = in ? cfI : (wall[j] ? 0.0 : nuTc);
                const double gX = in ? gxf : gac, gY = in ? gyf : gbc;
                fl[j] = cf * (gX * nx + gY * ny) * Sf[j];
            }
            double fl4[kSlots];
            gatherq(fl, fl4);
            double tsrc = 0.0;
#pragma unroll
            for (int s = 0; s < kSlots; ++s) {
                const double t = (!q.inner[s] || q.own[s]) ? tsrc + fl4[s] : tsrc - fl4[s];
                tsrc = q.valid[s] ? t : tsrc;
            }
            sr += tsrc;
        }
        const double uc = C(ci, Uc);
        if (S.momentum_convection != CONV_UPWIND) {
            const int gx = comp == 0 ? V_gUxx : V_gUyx;
            const int gy = comp == 0 ? V_gUxy : V_gUyy;
            const double gxc = C(ci, gx), gyc = C(ci, gy);
            double co[NJ];
#pragma unroll
            for (int j = 0; j < NJ; ++j)
                co[j] = convection_corr(o.f[j], o.own[j], mFlux[j], uc, gxc, gyc, o.n[j], Uc, gx,
                                        gy, S.momentum_convection);
            double co4[kSlots];
            gatherq(co, co4);
#pragma unroll
            for (int s = 0; s < kSlots; ++s) {
                const double t = q.own[s] ? sr - co4[s] : sr + co4[s];
                sr = q.inner[s] ? t : sr;
            }
        }
        const double alphaU = S.alpha_u;
        aPu = dg;
        sr += (1.0 - alphaU) / alphaU * dg * uc;
        dg /= alphaU;
        if (qi == 0) {
            C(ci, V_diag) = dg;
            C(ci, V_src) = sr;
            if (comp == 0) C(ci, V_aPs) = dg;
        }
        ok = isfinite(dg) && isfinite(sr) && isfinite(uc);
        double xn[NJ];
#pragma unroll
        for (int j = 0; j < NJ; ++j) xn[j] = C(o.n[j], Uc);
        const double y = row_residual_q(ci, q, o, offv, xn, dg * uc, ok);
        return sr - y;
    }

    // turb_faces with QL lanes per cell; each lane keeps its slots' terms
    struct TurbOwn {
        double Sf[NJ], delta[NJ], mFlux[NJ], nuEff[NJ], offv[NJ];
    };
    DEV void turb_faces_q(int ci, const SlotSet &q, const OwnSlots &o, double sig0, double sig1,
                          int Bv, int Cv, double F1c, double nuTc, double &dg, double &sr,
                          TurbOwn &t) const {
        const double nu = S.nu;
        const double sgB = blend(sig0, sig1, F1c);
        const double nuEffB = nu + sgB * nuTc;
        double cd[NJ], cb[NJ];
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            const int f = o.f[j];
            const double Sf = Fc(f, W_A);
            const double delta = Fc(f, W_dlt);
            const double m = Fc(f, W_phi);
            const double wt = Fc(f, W_w), wmv = Fc(f, W_wm);
            const double F1n = C(o.n[j], V_F1), ntn_ = C(o.n[j], V_nuT);
            const double bv = Bd(o.b[j], Bv);
            const bool own = o.own[j], in = o.in[j];
            const double F1o = own ? F1c : F1n, F1nb = own ? F1n : F1c;
            const double nto = own ? nuTc : ntn_, ntn = own ? ntn_ : nuTc;
            const double F1f = wt * F1o + wmv * F1nb;
            const double sg = blend(sig0, sig1, F1f);
            const double nuEff_o = nu + sg * nto;
            const double nuEff_n = nu + sg * ntn;
            const double nuEff_f = wt * nuEff_o + wmv * nuEff_n;
            const double D = (in ? nuEff_f : nuEffB) * Sf / delta;
            const double cPos = smax(m, 0.0), cNeg = smax(-m, 0.0);
            const double dIn = own ? D + cPos : D + cNeg;
            const double oIn = -(own ? D + cNeg : D + cPos);
            const bool pos = m >= 0;
            const double dB = pos ? m + D : D;
            cb[j] = (pos ? D : D - m) * bv;
            cd[j] = in ? dIn : dB;
            t.offv[j] = in ? oIn : 0.0;
            t.Sf[j] = Sf;
            t.delta[j] = delta;
            t.mFlux[j] = m;
            t.nuEff[j] = nuEff_f;
        }
        double cd4[kSlots], cb4[kSlots];
        gatherq(cd, cd4);
        gatherq(cb, cb4);
#pragma unroll
        for (int s = 0; s < kSlots; ++s) {
            dg = q.valid[s] ? dg + cd4[s] : dg;
            sr = (q.valid[s] && !q.inner[s]) ? sr + cb4[s] : sr;
        }
        if (M.folded) folded_faces(ci, nuEffB, C(ci, Cv), dg, sr);
    }
    DEV double turb_convection_q(int ci, const SlotSet &q, const OwnSlots &o, const TurbOwn &t,
                                 double sr, int ph, int gx, int gy) const {
        const double phc = C(ci, ph), gxc = C(ci, gx), gyc = C(ci, gy);
        double co[NJ];
#pragma unroll
        for (int j = 0; j < NJ; ++j)
            co[j] = convection_corr(o.f[j], o.own[j], t.mFlux[j], phc, gxc, gyc, o.n[j], ph, gx,
                                    gy, S.turbulence_convection);
        double co4[kSlots];
        gatherq(co, co4);
#pragma unroll
        for (int s = 0; s < kSlots; ++s) {
            const double v = q.own[s] ? sr - co4[s] : sr + co4[s];
            sr = q.inner[s] ? v : sr;
        }
        return sr;
    }

    // The gradients and SST closure of a turbulence update with the two lanes
    // of a pair (SSTModel::computeFields, SSTModel.cpp:208-240; the startup
    // nuT floor, SIMPLESolver.cpp:1333-1337): lane 0 takes the Green-Gauss
    // gradients of Ux, Uy and p, lane 1 those of k and omega (each field's
    // sum is the one-lane sum); the F1 and F2 arguments share their common
    // terms and one tanh evaluates F1 in lane 0 and F2 in lane 1. Every value
    // is computed with the operations of the one-lane loop.
    DEV void turbulence_fields_pair(int ci, double nuTMin) {
        const bool l0 = qi == 0;
        const int ph[3] = {l0 ? V_Ux : V_k, l0 ? V_Uy : V_w, l0 ? V_p : V_k};
        const int pb[3] = {l0 ? B_Ux : B_k, l0 ? B_Uy : B_w, l0 ? B_p : B_k};
        double gx[3], gy[3];
        gg_cell<3>(ci, ph, pb, gx, gy);
        // lane 0 receives lane 1's k and omega gradients
        const double gkx = lane_bcd<1>(gx[0]), gky = lane_bcd<1>(gy[0]);
        const double gwx = lane_bcd<1>(gx[1]), gwy = lane_bcd<1>(gy[1]);
        if (l0) {
            C(ci, V_gUxx) = gx[0];
            C(ci, V_gUxy) = gy[0];
            C(ci, V_gUyx) = gx[1];
            C(ci, V_gUyy) = gy[1];
            C(ci, V_gPx) = gx[2];
            C(ci, V_gPy) = gy[2];
        } else {
            C(ci, V_gKx) = gx[0];
            C(ci, V_gKy) = gy[0];
            C(ci, V_gWx) = gx[1];
            C(ci, V_gWy) = gy[1];
        }
        const double nu = S.nu;
        const double kc = smax(C(ci, V_k), 0.0);
        const double wc = smax(C(ci, V_w), 1e-20);
        const double y = smax(C(ci, V_wd), 1e-20);
        const double omSafe = smax(wc, 1e-20);
        const double cd = 2.0 * c[5] / omSafe * (gkx * gwx + gky * gwy);
        // sst_f1 in lane 0, sst_f2 in lane 1, term by term
        const double sqrtK = sqrt(smax(kc, 0.0));
        const double om2 = smax(wc, 1e-20);
        const double ySafe = smax(y, 1e-20);
        const double y2s = ySafe * ySafe;
        const double term1 = (l0 ? sqrtK : 2.0 * sqrtK) / (c[8] * om2 * ySafe);
        const double term2 = 500.0 * nu / (y2s * om2);
        const double CDpos = smax(cd, 1e-10);
        const double term3 = 4.0 * c[5] * smax(kc, 0.0) / (CDpos * y2s);
        const double arg1 = smin(smax(term1, term2), term3);
        const double arg2 = smax(term1, term2);
        const double a = l0 ? arg1 * arg1 * arg1 * arg1 : arg2 * arg2;
        const double t = tanh(a);
        const double f1 = t;                 // lane 0's
        const double f2 = lane_bcd<1>(t);    // lane 1's
        if (l0) {
            const double S11 = gx[0], S22 = gy[1];
            const double S12 = 0.5 * (gy[0] + gx[1]);
            const double Sc = sqrt(2.0 * (S11 * S11 + S22 * S22 + 2.0 * (S12 * S12)));
            C(ci, V_S) = Sc;
            C(ci, V_CDkw) = cd;
            C(ci, V_F1) = (S.sst_variant == 2) ? 1.0 : f1;
            C(ci, V_F2) = f2;
            const double nt = sst_nut(kc, wc, Sc, f2);
            C(ci, V_Pk) = sst_pk(nt, Sc, kc, wc);
            C(ci, V_nuT) = smax(nt, nuTMin);
        }
    }

    // k_row with QL lanes per cell
    DEV void k_row_q(int ci, bool &ok, double &resid) {
        SlotSet q;
        slot_set(ci, q);
        OwnSlots o;
        own_slots(q, o);
        double dg = 0.0, sr = 0.0;
        const double F1c = C(ci, V_F1), nuTc = C(ci, V_nuT);
        TurbOwn t;
        turb_faces_q(ci, q, o, c[0], c[4], B_k, V_k, F1c, nuTc, dg, sr, t);
        const double V = C(ci, V_vol);
        sr += C(ci, V_Pk) * V;
        const double destruction = c[8] * smax(C(ci, V_w), 1e-20);
        dg += destruction * V;
        if (S.turbulence_convection != CONV_UPWIND)
            sr = turb_convection_q(ci, q, o, t, sr, V_k, V_gKx, V_gKy);
        const double aK = S.alpha_k;
        const double kc = C(ci, V_k);
        sr += (1.0 - aK) / aK * dg * kc;
        dg /= aK;
        if (qi == 0) {
            C(ci, V_diag) = dg;
            C(ci, V_src) = sr;
        }
        ok = isfinite(dg) && isfinite(sr) && isfinite(kc);
        double xn[NJ];
#pragma unroll
        for (int j = 0; j < NJ; ++j) xn[j] = C(o.n[j], V_k);
        const double y = row_residual_q(ci, q, o, t.offv, xn, dg * kc, ok);
        resid = sr - y;
    }

    // omega_row with QL lanes per cell
    DEV double omega_row_q(int ci, bool &ok, double &resid) {
        const double nu = S.nu;
        SlotSet q;
        slot_set(ci, q);
        OwnSlots o;
        own_slots(q, o);
        double dg = 0.0, sr = 0.0, balance = 0.0;
        const double amplitude = 6 * nu / c[2];
        const double F1c = C(ci, V_F1), nuTc = C(ci, V_nuT);
        TurbOwn t;
        turb_faces_q(ci, q, o, c[1], c[5], B_w, V_w, F1c, nuTc, dg, sr, t);
        if (planar) {
            const double omsc = C(ci, V_oms);
            double co[NJ];
#pragma unroll
            for (int j = 0; j < NJ; ++j) {
                const bool own = o.own[j];
                const double omsn = C(o.n[j], V_oms);
                const double omg = Fc(o.f[j], W_omg);
                const double so = own ? omsc : omsn, sn = own ? omsn : omsc;
                const double difference = (sn - so) / t.delta[j];
                const double nuEff_f = t.nuEff[j];
                co[j] = amplitude * t.Sf[j] * ((nuEff_f - nu) * omg - nuEff_f * difference);
            }
            double co4[kSlots];
            gatherq(co, co4);
#pragma unroll
            for (int s = 0; s < kSlots; ++s) {
                const double v = q.own[s] ? balance + co4[s] : balance - co4[s];
                balance = q.inner[s] ? v : balance;
            }
        }
        const double V = C(ci, V_vol);
        const double wc = C(ci, V_w);
        const double omC = smax(wc, 1e-20);
        const double alphaB = blend(c[3], c[7], F1c);
        const double production = sst_pomega(nuTc, C(ci, V_S), C(ci, V_k), wc);
        sr += alphaB * production * V;
        const double betaB = blend(c[2], c[6], F1c);
        dg += betaB * omC * V;
        sr += (1.0 - F1c) * C(ci, V_CDkw) * V;
        if (planar) {
            const double correction = balance + nu * (6 * nu / c[2]) * C(ci, V_oml) * V;
            if (correction < 0) dg -= correction / omC;
            else sr += correction;
        }
        if (S.turbulence_convection != CONV_UPWIND)
            sr = turb_convection_q(ci, q, o, t, sr, V_w, V_gWx, V_gWy);
        const double aW = S.alpha_omega;
        sr += (1.0 - aW) / aW * dg * wc;
        dg /= aW;
        const double wvc = C(ci, V_wallv);
        double x0n[NJ], offv[NJ], pv[NJ];
        int sb[NJ];
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            const bool own = o.own[j], in = o.in[j];
            const double vn_ = C(o.n[j], V_wallv);
            const double wn_ = C(o.n[j], V_w);
            x0n[j] = isfinite(vn_) ? vn_ : wn_;
            const double vo = own ? wvc : vn_, vn = own ? vn_ : wvc;
            const bool fo = isfinite(vo), fn = isfinite(vn);
            const bool sub = in && (fo ? !own : (fn && own));
            const double v = fo ? vo : vn;
            pv[j] = t.offv[j] * v;
            sb[j] = sub ? 1 : 0;
            offv[j] = (in && (fo || fn)) ? 0.0 : t.offv[j];
        }
        double pv4[kSlots];
        int sb4[kSlots];
        gatherq(pv, pv4);
        gatherqi(sb, sb4);
#pragma unroll
        for (int s = 0; s < kSlots; ++s) sr = sb4[s] ? sr - pv4[s] : sr;
        if (isfinite(wvc)) sr = dg * wvc;
        if (qi == 0) {
            C(ci, V_diag) = dg;
            C(ci, V_src) = sr;
        }
        const double x0 = isfinite(wvc) ? wvc : wc;
        ok = isfinite(dg) && isfinite(sr) && isfinite(x0);
        const double y = row_residual_q(ci, q, o, offv, x0n, dg * x0, ok);
        resid = sr - y;
        return x0;
    }

    // ------------------------------------------------------------ Krylov
    DEV double jac(double r, double d) const { return (fabs(d) > 1e-30) ? r / d : r; }

    // BiCGSTABSolver::solve (LinearSolver.hpp:259-343). On entry kr holds
    // b - A x and kx holds x; rr = r.r.
    DEV Krylov bicgstab(double rr, int maxIter, double tol) {
        Krylov res{0, 0, 0.0};
        const double r0 = sqrt(rr);
        res.initial = r0;
        if (r0 < 1e-30) {
            res.converged = 1;
            return res;
        }
        for (int ci = gt; ci < nc; ci += gs) {
            C(ci, V_krh) = C(ci, V_kr);
            C(ci, V_kp) = 0.0;
            C(ci, V_kv) = 0.0;
        }
        double rho = 1, alpha = 1, omega = 1;
        double rhoN = rr;  // rh == r on entry, so rh.r == r.r
        for (int it = 0; it < maxIter; ++it) {
            if (fabs(rhoN) < 1e-30) break;
            const double beta = (rhoN / (rho + 1e-30)) * (alpha / (omega + 1e-30));
            for (int ci = gt; ci < nc; ci += gs) {
                const double pv = C(ci, V_kr) + beta * (C(ci, V_kp) - omega * C(ci, V_kv));
                C(ci, V_kp) = pv;
                C(ci, V_kph) = jac(pv, C(ci, V_diag));
            }
            sync();
            double a1[1] = {0.0};
            for (int ci = gt; ci < nc; ci += gs) {
                const double vv = spmv(ci, V_kph);
                C(ci, V_kv) = vv;
                a1[0] += C(ci, V_krh) * vv;
            }
            red.sum(a1);
            const double rv = a1[0];
            if (fabs(rv) < 1e-30) break;
            alpha = rhoN / rv;
            double a2[1] = {0.0};
            for (int ci = gt; ci < nc; ci += gs) {
                const double sv = C(ci, V_kr) - alpha * C(ci, V_kv);
                C(ci, V_ks) = sv;
                C(ci, V_ksh) = jac(sv, C(ci, V_diag));
                a2[0] += sv * sv;
            }
            red.sum(a2);
            const double sn = sqrt(a2[0]);
            if (sn / r0 < tol) {
                for (int ci = gt; ci < nc; ci += gs) C(ci, V_kx) += alpha * C(ci, V_kph);
                sync();
                res.iterations = it + 1;
                res.converged = 1;
                return res;
            }
            double a3[2] = {0.0, 0.0};
            for (int ci = gt; ci < nc; ci += gs) {
                const double tv = spmv(ci, V_ksh);
                C(ci, V_kt) = tv;
                a3[0] += tv * tv;
                a3[1] += tv * C(ci, V_ks);
            }
            red.sum(a3);
            const double tt = a3[0];
            omega = (tt > 1e-30) ? a3[1] / tt : 0.0;
            double a4[2] = {0.0, 0.0};
            for (int ci = gt; ci < nc; ci += gs) {
                C(ci, V_kx) += alpha * C(ci, V_kph) + omega * C(ci, V_ksh);
                const double rv2 = C(ci, V_ks) - omega * C(ci, V_kt);
                C(ci, V_kr) = rv2;
                a4[0] += rv2 * rv2;
                a4[1] += C(ci, V_krh) * rv2;
            }
            red.sum(a4);
            const double rn = sqrt(a4[0]);
            res.iterations = it + 1;
            if (rn / r0 < tol) {
                res.converged = 1;
                return res;
            }
            if (fabs(omega) < 1e-30) break;
            rho = rhoN;
            rhoN = a4[1];
        }
        return res;
    }

    // PCGSolver::solve (LinearSolver.hpp:191-244) from x = 0. On entry kr
    // holds b and rr = b.b.
    DEV Krylov pcg(double rr, int maxIter, double tol) {
        const int z = V_krh, pv = V_kp, Ap = V_kv;
        Krylov res{0, 0, 0.0};
        const double r0 = sqrt(rr);
        res.initial = r0;
        if (r0 < 1e-30) {
            res.converged = 1;
            return res;
        }
        double a0[1] = {0.0};
        for (int ci = gt; ci < nc; ci += gs) {
            const double zv = jac(C(ci, V_kr), C(ci, V_diag));
            C(ci, z) = zv;
            C(ci, pv) = zv;
            a0[0] += C(ci, V_kr) * zv;
        }
        red.sum(a0);
        double rz = a0[0];
        int pc = pv, pn = V_kph;  // p, and the buffer the next p is written to
        double beta = 0.0;
        for (int it = 0; it < maxIter; ++it) {
            double a1[1] = {0.0};
            pcg_product(it, beta, pc, pn, z, a1);
            red.sum(a1);
            const double pAp = a1[0];
            if (fabs(pAp) < 1e-30) break;
            const double alpha = rz / pAp;
            double a2[2] = {0.0, 0.0};
            for (int ci = gt; ci < nc; ci += gs) {
                C(ci, V_kx) += alpha * C(ci, pc);
                const double rv = C(ci, V_kr) - alpha * C(ci, Ap);
                C(ci, V_kr) = rv;
                const double zv = jac(rv, C(ci, V_diag));
                C(ci, z) = zv;
                a2[0] += rv * rv;
                a2[1] += rv * zv;
            }
            red.sum(a2);
            const double rn = sqrt(a2[0]);
            res.iterations = it + 1;
            if (rn / r0 < tol) {
                res.converged = 1;
                return res;
            }
            const double rzNew = a2[1];
            beta = rzNew / (rz + 1e-30);
            rz = rzNew;
        }
        return res;
    }

    // ------------------------------------------- wave-resident Krylov
    // For meshes of at most 64*CPL cells one wave runs the whole Krylov
    // solve: lane l owns cells l, l+64, ..., the vectors stay in registers,
    // only the vector the matrix multiplies goes through the arena (so the
    // neighbours' values can be gathered), and dot products are single-wave
    // DPP reductions with no block barrier. Each per-cell operation is the one
    // of the block-wide solvers above, in the same order.
    DEV static void wave_sync() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }

    template <int CPL>
    struct Rows {
        double dg[CPL], of[CPL][kSlots];
        int nbr[CPL][kSlots];
        bool act[CPL];
    };

    template <int CPL>
    DEV void load_rows(Rows<CPL> &R) const {
#pragma unroll
        for (int j = 0; j < CPL; ++j) {
            const int ci = tid + 64 * j;
            R.act[j] = ci < nc;
            const v4i raw = slot4[2 * (R.act[j] ? ci : 0) + 1];
            const v4i nb4 = R.act[j] ? raw : v4i{-1, -1, -1, -1};
            R.nbr[j][0] = nb4.x;
            R.nbr[j][1] = nb4.y;
            R.nbr[j][2] = nb4.z;
            R.nbr[j][3] = nb4.w;
            const int cc = R.act[j] ? ci : 0;
            const double dgv = C(cc, V_diag);
            R.dg[j] = R.act[j] ? dgv : 0.0;
#pragma unroll
            for (int s = 0; s < kSlots; ++s) {
                const double ov = off(cc, s);
                R.of[j][s] = R.nbr[j][s] >= 0 ? ov : 0.0;
            }
        }
    }

    // y = A x for the lane's rows; xs holds the lane's own entries of x and
    // the arena variable xv all of x (LinearSystem::matvec order)
    template <int CPL>
    DEV void spmv_wave(const Rows<CPL> &R, const double (&xs)[CPL], int xv, double (&y)[CPL]) const {
#pragma unroll
        for (int j = 0; j < CPL; ++j) {
            double a = R.dg[j] * xs[j];
#pragma unroll
            for (int s = 0; s < kSlots; ++s) {
                const double xn = C(R.nbr[j][s] >= 0 ? R.nbr[j][s] : 0, xv);
                a = R.nbr[j][s] >= 0 ? a + R.of[j][s] * xn : a;
            }
            y[j] = a;
        }
    }

    template <int CPL>
    DEV void publish(const Rows<CPL> &R, const double (&xs)[CPL], int xv) {
#pragma unroll
        for (int j = 0; j < CPL; ++j)
            if (R.act[j]) C(tid + 64 * j, xv) = xs[j];
        wave_sync();
    }

    template <int CPL>
    DEV void store_x(const Rows<CPL> &R, const double (&x)[CPL]) {
#pragma unroll
        for (int j = 0; j < CPL; ++j)
            if (R.act[j]) C(tid + 64 * j, xdst) = x[j];
    }

    // BiCGSTABSolver::solve (LinearSolver.hpp:259-343), wave 0 only
    template <int CPL>
    DEV Krylov bicgstab_wave(double rr, int maxIter, double tol) {
        Krylov res{0, 0, 0.0};
        const double r0 = sqrt(rr);
        res.initial = r0;
        if (r0 < 1e-30) {
            res.converged = 1;
            return res;
        }
        Rows<CPL> R;
        load_rows(R);
        double r[CPL], rh[CPL], p[CPL], v[CPL], x[CPL];
#pragma unroll
        for (int j = 0; j < CPL; ++j) {
            const int ci = tid + 64 * j;
            r[j] = R.act[j] ? C(ci, V_kr) : 0.0;
            x[j] = R.act[j] ? C(ci, V_kx) : 0.0;
            rh[j] = r[j];
            p[j] = 0.0;
            v[j] = 0.0;
        }
        double rho = 1, alpha = 1, omega = 1;
        double rhoN = rr;
        for (int it = 0; it < maxIter; ++it) {
            if (fabs(rhoN) < 1e-30) break;
            const double beta = (rhoN / (rho + 1e-30)) * (alpha / (omega + 1e-30));
            double ph[CPL];
#pragma unroll
            for (int j = 0; j < CPL; ++j) {
                p[j] = r[j] + beta * (p[j] - omega * v[j]);
                ph[j] = jac(p[j], R.dg[j]);
            }
            publish(R, ph, V_kph);
            spmv_wave(R, ph, V_kph, v);
            double a1 = 0.0;
#pragma unroll
            for (int j = 0; j < CPL; ++j) a1 += rh[j] * v[j];
            const double rv = wave_sum(a1);
            if (fabs(rv) < 1e-30) break;
            alpha = rhoN / rv;
            double s[CPL], sh[CPL], a2 = 0.0;
#pragma unroll
            for (int j = 0; j < CPL; ++j) {
                s[j] = r[j] - alpha * v[j];
                sh[j] = jac(s[j], R.dg[j]);
                a2 += s[j] * s[j];
            }
            const double sn = sqrt(wave_sum(a2));
            if (sn / r0 < tol) {
#pragma unroll
                for (int j = 0; j < CPL; ++j) x[j] += alpha * ph[j];
                store_x(R, x);
                res.iterations = it + 1;
                res.converged = 1;
                return res;
            }
            publish(R, sh, V_ksh);
            double t[CPL];
            spmv_wave(R, sh, V_ksh, t);
            double tt = 0.0, ts = 0.0;
#pragma unroll
            for (int j = 0; j < CPL; ++j) {
                tt += t[j] * t[j];
                ts += t[j] * s[j];
            }
            tt = wave_sum(tt);
            ts = wave_sum(ts);
            omega = (tt > 1e-30) ? ts / tt : 0.0;
            double rr2 = 0.0, rhr = 0.0;
#pragma unroll
            for (int j = 0; j < CPL; ++j) {
                x[j] += alpha * ph[j] + omega * sh[j];
                r[j] = s[j] - omega * t[j];
                rr2 += r[j] * r[j];
                rhr += rh[j] * r[j];
            }
            const double rn = sqrt(wave_sum(rr2));
            rhr = wave_sum(rhr);
            res.iterations = it + 1;
            if (rn / r0 < tol) {
                store_x(R, x);
                res.converged = 1;
                return res;
            }
            if (fabs(omega) < 1e-30) break;
            rho = rhoN;
            rhoN = rhr;
        }
        store_x(R, x);
        return res;
    }

    // PCGSolver::solve (LinearSolver.hpp:191-244) from x = 0, wave 0 only
    template <int CPL>
    DEV Krylov pcg_wave(double rr, int maxIter, double tol) {
        Krylov res{0, 0, 0.0};
        const double r0 = sqrt(rr);
        res.initial = r0;
        if (r0 < 1e-30) {
            res.converged = 1;
            return res;
        }
        Rows<CPL> R;
        load_rows(R);
        double r[CPL], z[CPL], p[CPL], x[CPL], a0 = 0.0;
#pragma unroll
        for (int j = 0; j < CPL; ++j) {
            const int ci = tid + 64 * j;
            r[j] = R.act[j] ? C(ci, V_kr) : 0.0;
            x[j] = R.act[j] ? C(ci, V_kx) : 0.0;
            z[j] = jac(r[j], R.dg[j]);
            p[j] = z[j];
            a0 += r[j] * z[j];
        }
        double rz = wave_sum(a0);
        for (int it = 0; it < maxIter; ++it) {
            publish(R, p, V_kp);
            double ap[CPL];
            spmv_wave(R, p, V_kp, ap);
            double a1 = 0.0;
#pragma unroll
            for (int j = 0; j < CPL; ++j) a1 += p[j] * ap[j];
            const double pAp = wave_sum(a1);
            if (fabs(pAp) < 1e-30) break;
            const double alpha = rz / pAp;
            double rr2 = 0.0, rzn = 0.0;
#pragma unroll
            for (int j = 0; j < CPL; ++j) {
                x[j] += alpha * p[j];
                r[j] = r[j] - alpha * ap[j];
                z[j] = jac(r[j], R.dg[j]);
                rr2 += r[j] * r[j];
                rzn += r[j] * z[j];
            }
            const double rn = sqrt(wave_sum(rr2));
            const double rzNew = wave_sum(rzn);
            res.iterations = it + 1;
            if (rn / r0 < tol) {
                store_x(R, x);
                res.converged = 1;
                return res;
            }
            const double beta = rzNew / (rz + 1e-30);
#pragma unroll
            for (int j = 0; j < CPL; ++j) p[j] = z[j] + beta * p[j];
            rz = rzNew;
        }
        store_x(R, x);
        return res;
    }

    // -------------------------------------- mixed precision (wave-resident)
    // FP32 Krylov on the defect equation A d = r0 / |r0| with the FP64
    // residual r0 = b - A x0 already in kr; x = x0 + |r0| d is formed in FP64.
    // The matrix, the vectors and the dot products are FP32; the stopping
    // rule is the solver's own (|r| / |r0| < tol). The breakdown guards act on
    // the normalised quantities.
    template <int CPL>
    struct Rows32 {
        float dg[CPL], of[CPL][kSlots];
        int nbr[CPL][kSlots];
        bool act[CPL];
    };

    template <int CPL>
    DEV void load_rows32(Rows32<CPL> &R) const {
#pragma unroll
        for (int j = 0; j < CPL; ++j) {
            const int ci = tid + 64 * j;
            R.act[j] = ci < nc;
            const v4i raw = slot4[2 * (R.act[j] ? ci : 0) + 1];
            const v4i nb4 = R.act[j] ? raw : v4i{-1, -1, -1, -1};
            R.nbr[j][0] = nb4.x;
            R.nbr[j][1] = nb4.y;
            R.nbr[j][2] = nb4.z;
            R.nbr[j][3] = nb4.w;
            const int cc = R.act[j] ? ci : 0;
            const float dgv = (float)C(cc, V_diag);
            R.dg[j] = R.act[j] ? dgv : 0.0f;
#pragma unroll
            for (int s = 0; s < kSlots; ++s) {
                const float ov = (float)off(cc, s);
                R.of[j][s] = R.nbr[j][s] >= 0 ? ov : 0.0f;
            }
        }
    }

    template <int CPL>
    DEV void spmv_wave32(const Rows32<CPL> &R, const float (&xs)[CPL], int xv, float (&y)[CPL]) const {
#pragma unroll
        for (int j = 0; j < CPL; ++j) {
            float a = R.dg[j] * xs[j];
#pragma unroll
            for (int s = 0; s < kSlots; ++s) {
                const float xn = (float)C(R.nbr[j][s] >= 0 ? R.nbr[j][s] : 0, xv);
                a = R.nbr[j][s] >= 0 ? a + R.of[j][s] * xn : a;
            }
            y[j] = a;
        }
    }

    template <int CPL>
    DEV void publish32(const Rows32<CPL> &R, const float (&xs)[CPL], int xv) {
#pragma unroll
        for (int j = 0; j < CPL; ++j)
            if (R.act[j]) C(tid + 64 * j, xv) = (double)xs[j];
        wave_sync();
    }

    DEV static float jacf(float r, float d) { return (fabsf(d) > 1e-30f) ? r / d : r; }

    template <int CPL>
    DEV void apply_defect(const Rows32<CPL> &R, double scale, const float (&d)[CPL]) {
#pragma unroll
        for (int j = 0; j < CPL; ++j)
            if (R.act[j]) {
                const int ci = tid + 64 * j;
                C(ci, xdst) = C(ci, V_kx) + scale * (double)d[j];
            }
    }

    template <int CPL>
    DEV Krylov bicgstab_wave32(double rr, int maxIter, double tol64) {
        Krylov res{0, 0, 0.0};
        const double r0 = sqrt(rr);
        res.initial = r0;
        if (r0 < 1e-30) {
            res.converged = 1;
            return res;
        }
        const float tol = (float)tol64;
        const double inv = 1.0 / r0;
        Rows32<CPL> R;
        load_rows32(R);
        float r[CPL], rh[CPL], p[CPL], v[CPL], x[CPL];
#pragma unroll
        for (int j = 0; j < CPL; ++j) {
            const int ci = tid + 64 * j;
            r[j] = R.act[j] ? (float)(C(ci, V_kr) * inv) : 0.0f;
            x[j] = 0.0f;
            rh[j] = r[j];
            p[j] = 0.0f;
            v[j] = 0.0f;
        }
        float rho = 1, alpha = 1, omega = 1;
        float rhoN = 0.0f;
#pragma unroll
        for (int j = 0; j < CPL; ++j) rhoN += r[j] * r[j];
        rhoN = wave_sumf(rhoN);
        for (int it = 0; it < maxIter; ++it) {
            if (fabsf(rhoN) < 1e-30f) break;
            const float beta = (rhoN / (rho + 1e-30f)) * (alpha / (omega + 1e-30f));
            float ph[CPL];
#pragma unroll
            for (int j = 0; j < CPL; ++j) {
                p[j] = r[j] + beta * (p[j] - omega * v[j]);
                ph[j] = jacf(p[j], R.dg[j]);
            }
            publish32(R, ph, V_kph);
            spmv_wave32(R, ph, V_kph, v);
            float a1 = 0.0f;
#pragma unroll
            for (int j = 0; j < CPL; ++j) a1 += rh[j] * v[j];
            const float rv = wave_sumf(a1);
            if (fabsf(rv) < 1e-30f) break;
            alpha = rhoN / rv;
            float s[CPL], sh[CPL], a2 = 0.0f;
#pragma unroll
            for (int j = 0; j < CPL; ++j) {
                s[j] = r[j] - alpha * v[j];
                sh[j] = jacf(s[j], R.dg[j]);
                a2 += s[j] * s[j];
            }
            const float sn = sqrtf(wave_sumf(a2));
            if (sn < tol) {
#pragma unroll
                for (int j = 0; j < CPL; ++j) x[j] += alpha * ph[j];
                apply_defect(R, r0, x);
                res.iterations = it + 1;
                res.converged = 1;
                return res;
            }
            publish32(R, sh, V_ksh);
            float t[CPL];
            spmv_wave32(R, sh, V_ksh, t);
            float tt = 0.0f, ts = 0.0f;
#pragma unroll
            for (int j = 0; j < CPL; ++j) {
                tt += t[j] * t[j];
                ts += t[j] * s[j];
            }
            tt = wave_sumf(tt);
            ts = wave_sumf(ts);
            omega = (tt > 1e-30f) ? ts / tt : 0.0f;
            float rr2 = 0.0f, rhr = 0.0f;
#pragma unroll
            for (int j = 0; j < CPL; ++j) {
                x[j] += alpha * ph[j] + omega * sh[j];
                r[j] = s[j] - omega * t[j];
                rr2 += r[j] * r[j];
                rhr += rh[j] * r[j];
            }
            const float rn = sqrtf(wave_sumf(rr2));
            rhr = wave_sumf(rhr);
            res.iterations = it + 1;
            if (rn < tol) {
                apply_defect(R, r0, x);
                res.converged = 1;
                return res;
            }
            if (fabsf(omega) < 1e-30f) break;
            rho = rhoN;
            rhoN = rhr;
        }
        apply_defect(R, r0, x);
        return res;
    }

    template <int CPL>
    DEV Krylov pcg_wave32(double rr, int maxIter, double tol64) {
        Krylov res{0, 0, 0.0};
        const double r0 = sqrt(rr);
        res.initial = r0;
        if (r0 < 1e-30) {
            res.converged = 1;
            return res;
        }
        const float tol = (float)tol64;
        const double inv = 1.0 / r0;
        Rows32<CPL> R;
        load_rows32(R);
        float r[CPL], z[CPL], p[CPL], x[CPL], a0 = 0.0f;
#pragma unroll
        for (int j = 0; j < CPL; ++j) {
            const int ci = tid + 64 * j;
            r[j] = R.act[j] ? (float)(C(ci, V_kr) * inv) : 0.0f;
            x[j] = 0.0f;
            z[j] = jacf(r[j], R.dg[j]);
            p[j] = z[j];
            a0 += r[j] * z[j];
        }
        float rz = wave_sumf(a0);
        for (int it = 0; it < maxIter; ++it) {
            publish32(R, p, V_kp);
            float ap[CPL];
            spmv_wave32(R, p, V_kp, ap);
            float a1 = 0.0f;
#pragma unroll
            for (int j = 0; j < CPL; ++j) a1 += p[j] * ap[j];
            const float pAp = wave_sumf(a1);
            if (fabsf(pAp) < 1e-30f) break;
            const float alpha = rz / pAp;
            float rr2 = 0.0f, rzn = 0.0f;
#pragma unroll
            for (int j = 0; j < CPL; ++j) {
                x[j] += alpha * p[j];
                r[j] = r[j] - alpha * ap[j];
                z[j] = jacf(r[j], R.dg[j]);
                rr2 += r[j] * r[j];
                rzn += r[j] * z[j];
            }
            const float rn = sqrtf(wave_sumf(rr2));
            const float rzNew = wave_sumf(rzn);
            res.iterations = it + 1;
            if (rn < tol) {
                apply_defect(R, r0, x);
                res.converged = 1;
                return res;
            }
            const float beta = rzNew / (rz + 1e-30f);
#pragma unroll
            for (int j = 0; j < CPL; ++j) p[j] = z[j] + beta * p[j];
            rz = rzNew;
        }
        apply_defect(R, r0, x);
        return res;
    }

    template <int CPL>
    DEV Krylov solve_wave(int kind, double rr, bool zeroStart) {
        bool f32 = kF32;
        if constexpr (kServe) f32 = S.mixed_precision != 0;
        if (f32) {
            if (kind == SOLVER_PCG && zeroStart)
                return pcg_wave32<CPL>(rr, S.inner_iterations, S.inner_tolerance);
            return bicgstab_wave32<CPL>(rr, S.inner_iterations, S.inner_tolerance);
        }
        if (kind == SOLVER_PCG && zeroStart)
            return pcg_wave<CPL>(rr, S.inner_iterations, S.inner_tolerance);
        return bicgstab_wave<CPL>(rr, S.inner_iterations, S.inner_tolerance);
    }

    // Every thread calls; the result is the same in every thread.
    // ------------------------------------------------------------ multigrid
    // AMG-preconditioned CG for the pressure correction (AMGSolver::solve,
    // LinearSolver.hpp:908-970) with a frozen aggregation hierarchy: every
    // solve refreshes the Galerkin values of all levels from the current
    // level-0 matrix (sumCoarse, LinearSolver.hpp:699-706, as fixed-order
    // gathers) and inverts the coarsest level densely. The V-cycle smooths
    // with damped Jacobi (2 sweeps before and after, omega = 2/3), so the
    // preconditioner is symmetric and every step is parallel.
    static constexpr double kJacobiOmega = 2.0 / 3.0;

    DEV double lev0_value(int src) const {
        const int ci = src >> 3, which = src & 7;
        return which == 0 ? C(ci, V_diag) : off(ci, which - 1);
    }

    // Multigrid workspace pointers. In a multi-workgroup draw, the levels
    // from amg_ls() down are read and written by rank 0 alone (their
    // Galerkin values too), so rank 0 reads them with plain loads, which may
    // hit its L1: its own stores keep that L1 current. Everything else another
    // workgroup writes is read with device-scope loads (WPtr).
    template <bool kPlain>
    using PPtr = typename std::conditional<kPlain, double *, WPtr>::type;
    template <bool kPlain>
    DEV PPtr<kPlain> pp(int64_t o) const {
        if constexpr (kPlain) return amgw + o;
        else return wp(o);
    }
    // first level run by rank 0 alone (levels [0, amg_ls()) are grid levels)
    DEV int amg_ls() const {
        const AmgMesh &G = M.amg;
        int ls = 1;
        if constexpr (kGrid)
            while (ls < G.levels - 1 && G.lev[ls].n > S.amg_block_rows) ++ls;
        return ls;
    }

    // Galerkin values and inverse diagonal of level l, rows e0, e0 + es, ...
    template <bool kPlain>
    DEV void refresh_level(int l, int e0, int es) {
        const AmgMesh &G = M.amg;
        const AmgLevel &F = G.lev[l - 1], &L = G.lev[l];
        const PPtr<kPlain> val = pp<kPlain>(L.val), dinv = pp<kPlain>(L.dinv);
        const WPtr fval = wp(F.val);
        for (int e = e0; e < L.nnz; e += es) {
            double a = 0.0;
            for (int q = F.gptr[e]; q < F.gptr[e + 1]; ++q) {
                const int fe = F.gidx[q];
                a += l == 1 ? lev0_value(G.src0[fe]) : double(fval[fe]);
            }
            val[e] = a;
        }
        kPlain ? bsync() : sync();
        for (int i = e0; i < L.n; i += es) {
            const double d = val[L.diagpos[i]];
            dinv[i] = fabs(d) > 1e-300 ? 1.0 / d : 0.0;
        }
    }

    DEV void amg_refresh() {
        const AmgMesh &G = M.amg;
        const int ls = amg_ls();
        for (int l = 1; l < G.levels; ++l) {
            if (kGrid && l >= ls) {
                // the tail of the hierarchy: rank 0 alone
                if (rank != 0) break;
                refresh_level<kGrid>(l, tid, BS);
                bsync();
            } else {
                refresh_level<false>(l, gt, gs);
            }
        }
        sync();
        // dense Gauss-Jordan inverse of the coarsest level (symmetric positive
        // definite: no pivoting); A in inv[0, n^2), its inverse after it. One
        // workgroup computes it (rank 0 of a multi-workgroup draw), so within
        // it A, X and fac are read back by the workgroup that wrote them.
        if (rank == 0) {
            const int n = G.coarse_n;
            const AmgLevel &L = G.lev[G.levels - 1];
            double *A = amgw + G.inv, *X = A + n * n, *fac = X + n * n;
            for (int e = tid; e < n * n; e += BS) {
                A[e] = 0.0;
                X[e] = (e / n == e % n) ? 1.0 : 0.0;
            }
            __syncthreads();
            const WPtr lval = wp(L.val);
            // (+=: the padding entries of a coarse row add zeros to its diagonal)
            for (int i = tid; i < n; i += BS)
                for (int q2 = L.rowptr[i]; q2 < L.rowptr[i + 1]; ++q2)
                    A[i * n + L.col[q2]] +=
                        G.levels == 1 ? lev0_value(G.src0[q2]) : double(lval[q2]);
            __syncthreads();
            for (int kk = 0; kk < n; ++kk) {
                const double piv = A[kk * n + kk];
                const double ip = fabs(piv) > 1e-300 ? 1.0 / piv : 1.0;
                __syncthreads();
                for (int j = tid; j < n; j += BS) {
                    A[kk * n + j] *= ip;
                    X[kk * n + j] *= ip;
                }
                for (int i = tid; i < n; i += BS) fac[i] = A[i * n + kk];
                __syncthreads();
                for (int e = tid; e < n * n; e += BS) {
                    const int i = e / n, j = e % n;
                    if (i == kk) continue;
                    A[e] -= fac[i] * A[kk * n + j];
                    X[e] -= fac[i] * X[kk * n + j];
                }
                __syncthreads();
            }
        }
        sync();
    }

    // y = A x on a coarse level (CSR); x(j) gives the vector's entries
    // (coarse rows are padded to a multiple of 4 entries, so a row is read
    // 4 entries at a time and the loads of a chunk are in flight together)
    template <bool kPlain = false, class F>
    DEV double row_f(const AmgLevel &L, int i, F x) const {
        double a = 0.0;
        const PPtr<kPlain> val = pp<kPlain>(L.val);
        for (int q = L.rowptr[i]; q < L.rowptr[i + 1]; q += 4) {
            const v4i c = *reinterpret_cast<const v4i *>(L.col + q);
            const double v0 = val[q], v1 = val[q + 1], v2 = val[q + 2], v3 = val[q + 3];
            const double x0 = x(c.x), x1 = x(c.y), x2 = x(c.z), x3 = x(c.w);
            a += v0 * x0;
            a += v1 * x1;
            a += v2 * x2;
            a += v3 * x3;
        }
        return a;
    }
    template <bool kPlain = false>
    DEV double coarse_row(const AmgLevel &L, int i, PPtr<kPlain> x) const {
        return row_f<kPlain>(L, i, [&](int j) { return double(x[j]); });
    }
    // level-0 product with the vector given entry by entry
    template <class F>
    DEV double spmv_f(int ci, F x) const {
        const v4i nb4 = slot4[2 * ci + 1];
        const int nbv[4] = {nb4.x, nb4.y, nb4.z, nb4.w};
        double y = C(ci, V_diag) * x(ci);
#pragma unroll
        for (int s = 0; s < kSlots; ++s) {
            const double a = off(ci, s), xn = x(nbv[s] >= 0 ? nbv[s] : ci);
            y = nbv[s] >= 0 ? y + a * xn : y;
        }
        return y;
    }

    // Workgroup barrier for work one workgroup does alone. In a
    // multi-workgroup draw its reads are device-scope loads, which bypass
    // the L1, so the workgroup's stores must have reached the L2 first.
    DEV void bsync() const {
        if constexpr (kGrid) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
    }

    // Levels with more than S.amg_block_rows rows are worked by every
    // workgroup of a multi-workgroup draw; smaller ones by rank 0 alone,
    // where a workgroup barrier replaces a grid barrier. Either way each row
    // is computed by the same arithmetic, so the results do not depend on it.

    // One level on the way down: two damped-Jacobi sweeps from zero, the
    // residual, and its restriction (aggregate sums) into the next level's
    // right-hand side. The first sweep's iterate w D^-1 b is formed where it
    // is read, so both sweeps take one pass; the arithmetic is the two
    // sweeps' own.
    template <bool kPlain = false>
    DEV void amg_down(int l, bool g) {
        const AmgMesh &G = M.amg;
        const AmgLevel &L = G.lev[l];
        const double w = kJacobiOmega;
        const int i0 = g ? gt : tid, is = g ? gs : BS;
        if (l == 0) {
            auto x0 = [&](int j) {
                const double d = C(j, V_diag);
                return fabs(d) > 1e-300 ? w * C(j, V_kr) / d : 0.0;
            };
            for (int ci = i0; ci < nc; ci += is) {
                const double d = C(ci, V_diag), xi = x0(ci);
                const double r0 = C(ci, V_kr) - spmv_f(ci, x0);
                C(ci, V_krh) = fabs(d) > 1e-300 ? xi + w * r0 / d : xi;
            }
            g ? sync() : bsync();
            for (int ci = i0; ci < nc; ci += is) C(ci, V_ks) = C(ci, V_kr) - spmv(ci, V_krh);
        } else {
            const PPtr<kPlain> x = pp<kPlain>(L.x), b = pp<kPlain>(L.b), r = pp<kPlain>(L.r),
                               di = pp<kPlain>(L.dinv);
            auto x0 = [&](int j) { return w * b[j] * di[j]; };
            for (int i = i0; i < L.n; i += is) {
                const double xi = x0(i), r0 = b[i] - row_f<kPlain>(L, i, x0);
                x[i] = xi + w * r0 * di[i];
            }
            g ? sync() : bsync();
            for (int i = i0; i < L.n; i += is) r[i] = b[i] - coarse_row<kPlain>(L, i, x);
        }
        // the residual is a pass of its own: formed inside the restriction,
        // each thread would run several members' products one after another
        // (measured slower than the extra barrier)
        g ? sync() : bsync();
        const PPtr<kPlain> rf = pp<kPlain>(L.r);
        amg_restrict<kPlain>(l, g,
                             [&](int m) { return l == 0 ? double(C(m, V_ks)) : double(rf[m]); });
    }

    // Restriction of level l's residual (res(m) of member m) into the next
    // level's right-hand side: aggregate sums in member order.
    template <bool kPlain = false, class F>
    DEV void amg_restrict(int l, bool g, F res) {
        const AmgMesh &G = M.amg;
        const AmgLevel &L = G.lev[l];
        const int i0 = g ? gt : tid, is = g ? gs : BS;
        const AmgLevel &Cl = G.lev[l + 1];
        const PPtr<kPlain> bc = pp<kPlain>(Cl.b);
        // member lists are padded to a multiple of 4 with -1
        for (int I = i0; I < Cl.n; I += is) {
            double a = 0.0;
            for (int q = L.mptr[I]; q < L.mptr[I + 1]; q += 4) {
                const v4i m = *reinterpret_cast<const v4i *>(L.midx + q);
                const double r0 = res(m.x);
                const double r1 = m.y >= 0 ? res(m.y) : 0.0;
                const double r2 = m.z >= 0 ? res(m.z) : 0.0;
                const double r3 = m.w >= 0 ? res(m.w) : 0.0;
                a += r0;
                if (m.y >= 0) a += r1;
                if (m.z >= 0) a += r2;
                if (m.w >= 0) a += r3;
            }
            bc[I] = a;
        }
        g ? sync() : bsync();
    }

    // One level on the way up: prolongation (x += xc of the cell's
    // aggregate) and two damped-Jacobi sweeps. The prolonged iterate is
    // formed where it is read and each sweep writes a second buffer (the
    // level's residual array, or V_kt on level 0), so each sweep is one
    // pass. The last pass of level 0 is not followed by a barrier: every
    // caller reads only its own rows next.
    template <bool kPlain = false>
    DEV void amg_up(int l, bool g) {
        const AmgMesh &G = M.amg;
        const AmgLevel &L = G.lev[l];
        const double w = kJacobiOmega;
        const int i0 = g ? gt : tid, is = g ? gs : BS;
        const PPtr<kPlain> xc = pp<kPlain>(G.lev[l + 1].x);
        if (l == 0) {
            auto xp = [&](int j) { return C(j, V_krh) + xc[L.agg[j]]; };
            for (int ci = i0; ci < nc; ci += is) {
                const double d = C(ci, V_diag), xi = xp(ci);
                const double r = C(ci, V_kr) - spmv_f(ci, xp);
                C(ci, V_kt) = fabs(d) > 1e-300 ? xi + w * r / d : xi;
            }
            g ? sync() : bsync();
            for (int ci = i0; ci < nc; ci += is) {
                const double d = C(ci, V_diag), xi = C(ci, V_kt);
                const double r = C(ci, V_kr) - spmv(ci, V_kt);
                C(ci, V_krh) = fabs(d) > 1e-300 ? xi + w * r / d : xi;
            }
        } else {
            const PPtr<kPlain> x = pp<kPlain>(L.x), b = pp<kPlain>(L.b), t = pp<kPlain>(L.r),
                               di = pp<kPlain>(L.dinv);
            auto xp = [&](int j) { return x[j] + xc[L.agg[j]]; };
            for (int i = i0; i < L.n; i += is) {
                const double xi = xp(i), r = b[i] - row_f<kPlain>(L, i, xp);
                t[i] = xi + w * r * di[i];
            }
            g ? sync() : bsync();
            for (int i = i0; i < L.n; i += is) {
                const double xi = t[i], r = b[i] - coarse_row<kPlain>(L, i, t);
                x[i] = xi + w * r * di[i];
            }
            g ? sync() : bsync();
        }
    }

    // z = M^-1 r: level-0 b is V_kr, x is V_krh, the residual V_ks, the
    // second sweep's buffer V_kt (FP32: the float vectors F_R, F_Z, F_S and
    // F_T2; the coarser levels stay FP64)
    template <bool F32, bool kPlain = false>
    DEV void amg_down_any(int l, bool g) {
        if constexpr (F32) {
            if (l == 0) return amg_down0_32(g);
        }
        amg_down<kPlain>(l, g);
    }
    template <bool F32, bool kPlain = false>
    DEV void amg_up_any(int l, bool g) {
        if constexpr (F32) {
            if (l == 0) return amg_up0_32(g);
        }
        amg_up<kPlain>(l, g);
    }
    template <bool F32 = false>
    DEV void amg_vcycle() {
        const AmgMesh &G = M.amg;
        const int nl = G.levels;
        if (nl == 1) {
            const int n = G.coarse_n;
            const WPtr X = wp(G.inv + (int64_t)n * n);
            for (int i = gt; i < n; i += gs) {
                double a = 0.0;
                if constexpr (F32) {
                    for (int j = 0; j < n; ++j) a += X[i * n + j] * (double)float(Fv(j, F_R));
                    Fv(i, F_Z) = (float)a;
                } else {
                    for (int j = 0; j < n; ++j) a += X[i * n + j] * C(j, V_kr);
                    C(i, V_krh) = a;
                }
            }
            sync();
            return;
        }
        // levels [0, ls) are grid levels (kGrid), the rest rank 0's
        const int ls = amg_ls();
        for (int l = 0; l < ls; ++l) amg_down_any<F32>(l, kGrid);
        if (!kGrid || rank == 0) {
            if constexpr (kGrid) {
                // level ls's right-hand side was restricted by every workgroup:
                // device-scope loads give its current values, and storing them
                // again from this workgroup keeps its L1 current for the plain
                // loads of rank 0's levels below
                const AmgLevel &Lt = G.lev[ls];
                double *bt = amgw + Lt.b;
                for (int i = tid; i < Lt.n; i += BS) {
                    const double v =
                        __hip_atomic_load(bt + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    __hip_atomic_store(bt + i, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                }
                bsync();
            }
            for (int l = ls; l < nl - 1; ++l) amg_down_any<F32, kGrid>(l, false);
            // coarsest: x = A^-1 b, with b staged in LDS; the inverse was
            // written by this workgroup, so its rows are read with plain loads
            const int n = G.coarse_n;
            const double *X = amgw + G.inv + (int64_t)n * n;
            const AmgLevel &L = G.lev[nl - 1];
            const PPtr<kGrid> x = pp<kGrid>(L.x), b = pp<kGrid>(L.b);
            for (int i = tid; i < n; i += BS) cbuf[i] = b[i];
            __syncthreads();
            for (int i = tid; i < n; i += BS) {
                double a = 0.0;
                for (int j = 0; j < n; ++j) a += X[i * n + j] * cbuf[j];
                x[i] = a;
            }
            bsync();
            for (int l = nl - 2; l >= ls; --l) amg_up_any<F32, kGrid>(l, false);
        }
        if constexpr (kGrid) sync();  // rank 0's levels are visible to every workgroup
        for (int l = ls - 1; l >= 0; --l) amg_up_any<F32>(l, kGrid);
    }

    // AMGSolver::solve from x = 0 (LinearSolver.hpp:908-970): kr holds b.
    DEV Krylov amg_pcg(double rr, int maxIter, double tol) {
        Krylov res{0, 0, 0.0};
        const double r0 = sqrt(rr);
        res.initial = r0;
        if (r0 < 1e-30) {
            res.converged = 1;
            return res;
        }
        amg_refresh();
        amg_vcycle();
        double a0[1] = {0.0};
        for (int ci = gt; ci < nc; ci += gs) {
            C(ci, V_kp) = C(ci, V_krh);
            a0[0] += C(ci, V_kr) * C(ci, V_krh);
        }
        red.sum(a0);
        double rz = a0[0];
        // p, and the buffer the next p is written to (the V-cycle uses krh, ks
        // and kt on level 0)
        int pc = V_kp, pn = V_kph;
        double beta = 0.0;
        for (int it = 0; it < maxIter; ++it) {
            double a1[1] = {0.0};
            pcg_product(it, beta, pc, pn, V_krh, a1);
            red.sum(a1);
            const double pAp = a1[0];
            if (fabs(pAp) < 1e-30) break;
            const double alpha = rz / pAp;
            double a2[1] = {0.0};
            for (int ci = gt; ci < nc; ci += gs) {
                C(ci, V_kx) += alpha * C(ci, pc);
                const double rv = C(ci, V_kr) - alpha * C(ci, V_kv);
                C(ci, V_kr) = rv;
                a2[0] += rv * rv;
            }
            red.sum(a2);
            const double rn = sqrt(a2[0]);
            res.iterations = it + 1;
            if (rn / r0 < tol) {
                res.converged = 1;
                return res;
            }
            amg_vcycle();
            double a3[1] = {0.0};
            for (int ci = gt; ci < nc; ci += gs) a3[0] += C(ci, V_kr) * C(ci, V_krh);
            red.sum(a3);
            const double rzNew = a3[0];
            beta = rzNew / (rz + 1e-30);
            rz = rzNew;
        }
        return res;
    }

    // Ap of a CG iteration into kv, p.Ap into a[0]. From the second
    // iteration on, p = z + beta p_old is formed where it is read (own row
    // and neighbours, the same expression) and written to the other buffer,
    // so the update needs no pass and no barrier of its own; the buffers
    // then swap. Every value has the bits of a stored update.
    DEV void pcg_product(int it, double beta, int &pc, int &pn, int z, double (&a)[1]) {
        if (it == 0) {
            for (int ci = gt; ci < nc; ci += gs) {
                const double av = spmv(ci, pc);
                C(ci, V_kv) = av;
                a[0] += C(ci, pc) * av;
            }
            return;
        }
        const int po = pc;
        auto pnew = [&](int j) { return C(j, z) + beta * C(j, po); };
        for (int ci = gt; ci < nc; ci += gs) {
            const double pv = pnew(ci);
            C(ci, pn) = pv;
            const double av = spmv_f(ci, pnew);
            C(ci, V_kv) = av;
            a[0] += pv * av;
        }
        const int written = pn;
        pn = pc;
        pc = written;
    }

    // -------------------------------------- mixed precision (block)
    // FP32 Krylov for meshes beyond the wave-resident path (arenas in HBM,
    // multi-workgroup draws, and LDS arenas of more than 256 cells). As in
    // the wave-resident form, the FP32 iteration solves the defect equation
    // A d = r0 / |r0| from d = 0, with r0 = b - A x0 the FP64 residual in kr,
    // and x = x0 + |r0| d is formed in FP64; the stopping rule is the
    // solver's own (|r| / |r0| < tol). Matrix (rounded from the FP64
    // assembly), vectors and per-thread dot products are FP32; the partial
    // sums meet in the FP64 reduction. Storage: each of the seven double
    // Krylov slots after kr holds two float vectors (in LDS records, the two
    // halves of the slot's double; in HBM, the slot's nc doubles read as
    // 2 nc floats). Multigrid levels below level 0 stay FP64.
    enum F32Vec {
        F_DG = 0, F_O0 = 1,  // matrix: diagonal, slot off-diagonals (1..4)
        F_R = 5, F_RH = 6, F_P = 7, F_V = 8, F_T = 9, F_PH = 10, F_SH = 11, F_X = 12,
        F_Z = F_RH, F_AP = F_V, F_S = F_T, F_T2 = F_PH
    };
    using FD = typename Space<kLDS>::F;
    using RefF = typename std::conditional<kGrid, GRefF, FD &>::type;
    DEV RefF Fv(int i, int k) const {
        // float k of the record is float 2 V_krh + k of the record; in HBM the
        // float vectors after V_krh's doubles are consecutive nc-long arrays
        FD *p;
        if constexpr (kLDS) p = (FD *)(Cb + V_krh) + (__mul24(i, 2 * kNA) + k);
        else p = (FD *)(Cb + (int64_t)V_krh * nc) + (k * nc + i);
        if constexpr (kGrid) return GRefF{(GlobalFloat *)p};
        else return *p;
    }
    DEV static float jacf32(float r, float d) { return (fabsf(d) > 1e-30f) ? r / d : r; }

    // every workgroup converts its own rows
    DEV void f32_matrix() {
        for (int ci = gt; ci < nc; ci += gs) {
            const v4i nb4 = slot4[2 * ci + 1];
            const int nbv[4] = {nb4.x, nb4.y, nb4.z, nb4.w};
            Fv(ci, F_DG) = (float)C(ci, V_diag);
#pragma unroll
            for (int s = 0; s < kSlots; ++s)
                Fv(ci, F_O0 + s) = nbv[s] >= 0 ? (float)off(ci, s) : 0.0f;
        }
    }
    template <class F>
    DEV float spmv32_f(int ci, F x) const {
        const v4i nb4 = slot4[2 * ci + 1];
        const int nbv[4] = {nb4.x, nb4.y, nb4.z, nb4.w};
        float y = float(Fv(ci, F_DG)) * x(ci);
#pragma unroll
        for (int s = 0; s < kSlots; ++s)
            if (nbv[s] >= 0) y += float(Fv(ci, F_O0 + s)) * x(nbv[s]);
        return y;
    }
    DEV float spmv32(int ci, int xv) const {
        return spmv32_f(ci, [&](int j) { return float(Fv(j, xv)); });
    }
    // x = x0 + |r0| d, then a barrier (callers read their own rows of x)
    DEV void defect32(double r0) {
        for (int ci = gt; ci < nc; ci += gs)
            C(ci, V_kx) = C(ci, V_kx) + r0 * (double)float(Fv(ci, F_X));
        sync();
    }
    // r = r0 / |r0| into F_R (and F_RH when rh), d = 0; returns r.r
    DEV double f32_start(double inv, bool rh) {
        float a = 0.0f;
        for (int ci = gt; ci < nc; ci += gs) {
            const float r = (float)(C(ci, V_kr) * inv);
            Fv(ci, F_R) = r;
            if (rh) Fv(ci, F_RH) = r;
            Fv(ci, F_X) = 0.0f;
            a += r * r;
        }
        return (double)a;
    }

    DEV Krylov bicgstab32(double rr, int maxIter, double tol64) {
        Krylov res{0, 0, 0.0};
        const double r0 = sqrt(rr);
        res.initial = r0;
        if (r0 < 1e-30) {
            res.converged = 1;
            return res;
        }
        const float tol = (float)tol64;
        f32_matrix();
        double a0[1] = {f32_start(1.0 / r0, true)};
        for (int ci = gt; ci < nc; ci += gs) {
            Fv(ci, F_P) = 0.0f;
            Fv(ci, F_V) = 0.0f;
        }
        red.sum(a0);
        float rho = 1, alpha = 1, omega = 1;
        float rhoN = (float)a0[0];
        for (int it = 0; it < maxIter; ++it) {
            if (fabsf(rhoN) < 1e-30f) break;
            const float beta = (rhoN / (rho + 1e-30f)) * (alpha / (omega + 1e-30f));
            for (int ci = gt; ci < nc; ci += gs) {
                const float pv = float(Fv(ci, F_R)) +
                                 beta * (float(Fv(ci, F_P)) - omega * float(Fv(ci, F_V)));
                Fv(ci, F_P) = pv;
                Fv(ci, F_PH) = jacf32(pv, Fv(ci, F_DG));
            }
            sync();
            float s1 = 0.0f;
            for (int ci = gt; ci < nc; ci += gs) {
                const float vv = spmv32(ci, F_PH);
                Fv(ci, F_V) = vv;
                s1 += float(Fv(ci, F_RH)) * vv;
            }
            double a1[1] = {(double)s1};
            red.sum(a1);
            const float rv = (float)a1[0];
            if (fabsf(rv) < 1e-30f) break;
            alpha = rhoN / rv;
            // s overwrites r (r is not read again before r = s - omega t)
            float s2 = 0.0f;
            for (int ci = gt; ci < nc; ci += gs) {
                const float sv = float(Fv(ci, F_R)) - alpha * float(Fv(ci, F_V));
                Fv(ci, F_R) = sv;
                Fv(ci, F_SH) = jacf32(sv, Fv(ci, F_DG));
                s2 += sv * sv;
            }
            double a2[1] = {(double)s2};
            red.sum(a2);
            const float sn = sqrtf((float)a2[0]);
            if (sn < tol) {
                for (int ci = gt; ci < nc; ci += gs)
                    Fv(ci, F_X) = float(Fv(ci, F_X)) + alpha * float(Fv(ci, F_PH));
                defect32(r0);
                res.iterations = it + 1;
                res.converged = 1;
                return res;
            }
            float t0 = 0.0f, t1 = 0.0f;
            for (int ci = gt; ci < nc; ci += gs) {
                const float tv = spmv32(ci, F_SH);
                Fv(ci, F_T) = tv;
                t0 += tv * tv;
                t1 += tv * float(Fv(ci, F_R));
            }
            double a3[2] = {(double)t0, (double)t1};
            red.sum(a3);
            const float tt = (float)a3[0];
            omega = (tt > 1e-30f) ? (float)a3[1] / tt : 0.0f;
            float u0 = 0.0f, u1 = 0.0f;
            for (int ci = gt; ci < nc; ci += gs) {
                Fv(ci, F_X) = float(Fv(ci, F_X)) + alpha * float(Fv(ci, F_PH)) +
                              omega * float(Fv(ci, F_SH));
                const float rv2 = float(Fv(ci, F_R)) - omega * float(Fv(ci, F_T));
                Fv(ci, F_R) = rv2;
                u0 += rv2 * rv2;
                u1 += float(Fv(ci, F_RH)) * rv2;
            }
            double a4[2] = {(double)u0, (double)u1};
            red.sum(a4);
            const float rn = sqrtf((float)a4[0]);
            res.iterations = it + 1;
            if (rn < tol) {
                defect32(r0);
                res.converged = 1;
                return res;
            }
            if (fabsf(omega) < 1e-30f) break;
            rho = rhoN;
            rhoN = (float)a4[1];
        }
        defect32(r0);
        return res;
    }

    // Jacobi PCG from x = 0 (kr holds b); AMG-preconditioned when amg
    template <bool kAmg>
    DEV Krylov pcg32(double rr, int maxIter, double tol64) {
        Krylov res{0, 0, 0.0};
        const double r0 = sqrt(rr);
        res.initial = r0;
        if (r0 < 1e-30) {
            res.converged = 1;
            return res;
        }
        const float tol = (float)tol64;
        if constexpr (kAmg) amg_refresh();
        f32_matrix();
        const double rs = f32_start(1.0 / r0, false);
        (void)rs;
        float s0 = 0.0f;
        if constexpr (kAmg) {
            sync();  // the V-cycle reads neighbours' rows of r and the matrix
            amg_vcycle<true>();
            for (int ci = gt; ci < nc; ci += gs) {
                const float z = Fv(ci, F_Z);
                Fv(ci, F_P) = z;
                s0 += float(Fv(ci, F_R)) * z;
            }
        } else {
            for (int ci = gt; ci < nc; ci += gs) {
                const float r = Fv(ci, F_R);
                const float z = jacf32(r, Fv(ci, F_DG));
                Fv(ci, F_Z) = z;
                Fv(ci, F_P) = z;
                s0 += r * z;
            }
        }
        double a0[1] = {(double)s0};
        red.sum(a0);
        float rz = (float)a0[0];
        for (int it = 0; it < maxIter; ++it) {
            float s1 = 0.0f;
            for (int ci = gt; ci < nc; ci += gs) {
                const float av = spmv32(ci, F_P);
                Fv(ci, F_AP) = av;
                s1 += float(Fv(ci, F_P)) * av;
            }
            double a1[1] = {(double)s1};
            red.sum(a1);
            const float pAp = (float)a1[0];
            if (fabsf(pAp) < 1e-30f) break;
            const float alpha = rz / pAp;
            float s2 = 0.0f, s3 = 0.0f;
            for (int ci = gt; ci < nc; ci += gs) {
                Fv(ci, F_X) = float(Fv(ci, F_X)) + alpha * float(Fv(ci, F_P));
                const float rv = float(Fv(ci, F_R)) - alpha * float(Fv(ci, F_AP));
                Fv(ci, F_R) = rv;
                s2 += rv * rv;
                if constexpr (!kAmg) {
                    const float zv = jacf32(rv, Fv(ci, F_DG));
                    Fv(ci, F_Z) = zv;
                    s3 += rv * zv;
                }
            }
            double a2[2] = {(double)s2, (double)s3};
            red.sum(a2);
            const float rn = sqrtf((float)a2[0]);
            res.iterations = it + 1;
            if (rn < tol) {
                defect32(r0);
                res.converged = 1;
                return res;
            }
            float rzNew = (float)a2[1];
            if constexpr (kAmg) {
                amg_vcycle<true>();
                float s4 = 0.0f;
                for (int ci = gt; ci < nc; ci += gs) s4 += float(Fv(ci, F_R)) * float(Fv(ci, F_Z));
                double a3[1] = {(double)s4};
                red.sum(a3);
                rzNew = (float)a3[0];
            }
            const float beta = rzNew / (rz + 1e-30f);
            for (int ci = gt; ci < nc; ci += gs)
                Fv(ci, F_P) = float(Fv(ci, F_Z)) + beta * float(Fv(ci, F_P));
            sync();
            rz = rzNew;
        }
        defect32(r0);
        return res;
    }

    // level 0 of the FP32 V-cycle: amg_down / amg_up on the float vectors
    // (b = F_R, x = F_Z, residual F_S, second sweep buffer F_T2)
    DEV void amg_down0_32(bool g) {
        const float w = (float)kJacobiOmega;
        const int i0 = g ? gt : tid, is = g ? gs : BS;
        auto x0 = [&](int j) {
            const float d = Fv(j, F_DG);
            return fabsf(d) > 1e-30f ? w * float(Fv(j, F_R)) / d : 0.0f;
        };
        for (int ci = i0; ci < nc; ci += is) {
            const float d = Fv(ci, F_DG), xi = x0(ci);
            const float r0 = float(Fv(ci, F_R)) - spmv32_f(ci, x0);
            Fv(ci, F_Z) = fabsf(d) > 1e-30f ? xi + w * r0 / d : xi;
        }
        g ? sync() : bsync();
        for (int ci = i0; ci < nc; ci += is) Fv(ci, F_S) = float(Fv(ci, F_R)) - spmv32(ci, F_Z);
        g ? sync() : bsync();
        amg_restrict(0, g, [&](int m) { return (double)float(Fv(m, F_S)); });
    }
    DEV void amg_up0_32(bool g) {
        const AmgMesh &G = M.amg;
        const AmgLevel &L = G.lev[0];
        const float w = (float)kJacobiOmega;
        const int i0 = g ? gt : tid, is = g ? gs : BS;
        const WPtr xc = wp(G.lev[1].x);
        auto xp = [&](int j) { return float(Fv(j, F_Z)) + (float)double(xc[L.agg[j]]); };
        for (int ci = i0; ci < nc; ci += is) {
            const float d = Fv(ci, F_DG), xi = xp(ci);
            const float r = float(Fv(ci, F_R)) - spmv32_f(ci, xp);
            Fv(ci, F_T2) = fabsf(d) > 1e-30f ? xi + w * r / d : xi;
        }
        g ? sync() : bsync();
        for (int ci = i0; ci < nc; ci += is) {
            const float d = Fv(ci, F_DG), xi = Fv(ci, F_T2);
            const float r = float(Fv(ci, F_R)) - spmv32(ci, F_T2);
            Fv(ci, F_Z) = fabsf(d) > 1e-30f ? xi + w * r / d : xi;
        }
    }

    // The wave-resident path is compiled into the launch shapes that small
    // meshes use (up to 256 threads).
    // The solution is left in kx and, when dst names another cell variable,
    // also in dst (the wave-resident solvers store it there directly).
    DEV Krylov solve(int kind, double rr, bool zeroStart, int dst = V_kx) {
        if constexpr (BS <= 256 && !kGrid) {
            if (nc <= 256 && !(kind == SOLVER_AMG && zeroStart && M.amg.levels > 0)) {
                // Every wave-resident solver stops before its first iteration
                // when |r| < 1e-30 (x keeps its start value); rr is the block
                // reduction's value, the same in every thread, so the whole
                // workgroup takes that exit here without the detour through
                // wave 0.
                const double r0 = sqrt(rr);
                if (r0 < 1e-30) {
                    if (dst != V_kx) {
                        for (int ci = tid; ci < nc; ci += BS) C(ci, dst) = C(ci, V_kx);
                        __syncthreads();
                    }
                    return Krylov{0, 1, r0};
                }
                // past that exit every solver iterates and stores x in dst
                xdst = dst;
                if (tid < 64) {
                    Krylov r;
                    if (nc <= 64) r = solve_wave<1>(kind, rr, zeroStart);
                    else if (nc <= 128) r = solve_wave<2>(kind, rr, zeroStart);
                    else r = solve_wave<4>(kind, rr, zeroStart);
                    if (tid == 0) *kres = r;
                }
                xdst = V_kx;
                __syncthreads();
                return *kres;
            }
        }
        const Krylov r = solve_block(kind, rr, zeroStart);
        if (dst != V_kx) {
            for (int ci = gt; ci < nc; ci += gs) C(ci, dst) = C(ci, V_kx);
            sync();
        }
        return r;
    }
    DEV Krylov solve_block(int kind, double rr, bool zeroStart) {
        if (kind == SOLVER_AMG && zeroStart && M.amg.levels > 0) {
            if constexpr (kF32) return pcg32<true>(rr, S.inner_iterations, S.inner_tolerance);
            else return amg_pcg(rr, S.inner_iterations, S.inner_tolerance);
        }
        if constexpr (kF32) {
            if (kind == SOLVER_PCG && zeroStart)
                return pcg32<false>(rr, S.inner_iterations, S.inner_tolerance);
            return bicgstab32(rr, S.inner_iterations, S.inner_tolerance);
        }
        if (kind == SOLVER_PCG && zeroStart) return pcg(rr, S.inner_iterations, S.inner_tolerance);
        return bicgstab(rr, S.inner_iterations, S.inner_tolerance);
    }

    // ------------------------------------------------------ contraction
    // ContractionEstimate (SIMPLESolver.cpp:32-74): each estimator keeps its
    // last 2m samples in an LDS ring (older samples never enter its rate or
    // remainder). Window maxima are taken across the lanes of wave 0; a NaN
    // sample counts as 0, as in the reference's max fold starting from 0.
    DEV DL *ring(int e) const { return ringb + e * 2 * S.convergence_window; }
    DEV void ce_push(int e, double d) {
        const int m2 = 2 * S.convergence_window;
        const int p = bk->nEst[e] == 0 ? 0 : (bk->pos[e] + 1 == m2 ? 0 : bk->pos[e] + 1);
        ring(e)[p] = d;
        bk->pos[e] = p;
        bk->nEst[e]++;
    }
    DEV double ce_last(int e) const { return bk->nEst[e] == 0 ? INFINITY : ring(e)[bk->pos[e]]; }
    // wave 0 only: maxima of the newest and the previous window (a full ring)
    DEV void ce_windows(int e, double &dNew, double &dOld) const {
        const int p = bk->pos[e];
        const int m = S.convergence_window, m2 = 2 * m;
        double nw = 0.0, ol = 0.0;
        for (int j = tid; j < m2; j += 64) {
            int age = p - j;
            age = age < 0 ? age + m2 : age;
            double v = ring(e)[j];
            v = (v != v) ? 0.0 : v;
            if (age < m) nw = smax(nw, v);
            else ol = smax(ol, v);
        }
        dNew = wave_max(nw);
        dOld = wave_max(ol);
    }
    DEV double ce_rate(int n, double dNew, double dOld) const {
        const int m = S.convergence_window;
        if (n < 2 * m) return NAN;
        if (dOld <= 0.0) return dNew > 0.0 ? INFINITY : 0.0;
        return pow(dNew / dOld, 1.0 / (double)m);
    }
    DEV double ce_remaining(int n, double dNew, double rate, double negligible) const {
        const int m = S.convergence_window;
        if (n < 2 * m) return INFINITY;
        if (dNew <= negligible) return dNew;
        if (!(rate < 1.0)) return INFINITY;
        return dNew * rate / (1.0 - rate);
    }

    // ------------------------------------------------------- observation
    // PreparedObservation::evaluate (ObservationOperator.hpp:678-772)
    DEV double observe(int i) const {
        const D *R = Ob + __mul24(i, kNO);
        const typename Space<kLDS>::I *Q = obi + 8 * i;
        const int type = Q[0] & 15;
        if (type == OBS_VELOCITY || type == OBS_PRESSURE) {
            const bool pressure = type == OBS_PRESSURE;
            const int comp = Q[0] >> 4;
            const bool zero = !pressure && comp == 2;
            const int cell = pressure ? V_p : (comp == 0 ? V_Ux : V_Uy);
            const int bnd = pressure ? B_p : (comp == 0 ? B_Ux : B_Uy);
            const double center = zero ? 0.0 : C(Q[1], cell);
            double bx = 0, by = 0;
            for (int qq = Q[2]; qq < Q[3]; ++qq) {
                const int nbc = onbi[2 * qq], nbf = onbi[2 * qq + 1];
                const double ndx = Onb[3 * qq], ndy = Onb[3 * qq + 1], nw = Onb[3 * qq + 2];
                const double nbv = zero ? 0.0 : (nbc < 0 ? Bd(nbf - nif, bnd) : C(nbc, cell));
                const double difference = nbv - center;
                bx += nw * ndx * difference;
                by += nw * ndy * difference;
            }
            double gxv = 0, gyv = 0;
            const double det = R[O_det], axx = R[O_axx], axy = R[O_axy], ayy = R[O_ayy];
            if (fabs(det) > 1e-14 * (axx * ayy + 1e-300)) {
                gxv = (ayy * bx - axy * by) / det;
                gyv = (axx * by - axy * bx) / det;
            }
            const double value = center + gxv * R[O_offx] + gyv * R[O_offy];
            const double refv = R[O_refv];
            const double dynP = .5 * R[O_refrho] * refv * refv;
            return pressure ? (value - R[O_refp]) / smax(dynP, 1e-20) : value;
        }
        // The multi-workgroup kernels carry no drag observation, and the
        // mixed-precision one no wall-crossing observations either: the host
        // runs problems that have them on one workgroup per draw.
        if constexpr (!kGrid) if (type == OBS_DRAG) {
            // pressure and viscous x force over the patch faces, as Cd
            double force = 0.0;
            for (int qq = Q[2]; qq < Q[3]; ++qq) {
                const int f = onbi[2 * qq + 1];
                const int ow = fon[f].x;
                const double area = Fc(f, W_A);
                force += C(ow, V_p) * Fc(f, W_nx) * area;
                const double tau = S.nu * (C(ow, V_Ux) - Bd(f - nif, B_Ux)) / Fc(f, W_dlt);
                force += tau * area;
            }
            const double refv = R[O_refv];
            const double dynP = .5 * refv * refv;
            return force / (dynP * R[O_offx]);
        }
        if constexpr (!(kGrid && kF32)) if (type == OBS_SEPARATION || type == OBS_REATTACHMENT) {
            // first sign change of the molecular streamwise wall shear along
            // the patch, interpolated linearly between face centres
            const bool sep = type == OBS_SEPARATION;
            const int q0 = Q[2], q1 = Q[3];
            double value = (sep || q1 == q0) ? 0.0 : Onb[3 * (q1 - 1)];
            bool have = false;
            double prevTau = sep ? 1.0 : -1.0, prevX = 0.0;
            for (int qq = q0; qq < q1; ++qq) {
                const int f = onbi[2 * qq + 1];
                const int ow = fon[f].x;
                const double nx = Fc(f, W_nx), ny = Fc(f, W_ny);
                const double ux = C(ow, V_Ux);
                const double un = ux * nx + C(ow, V_Uy) * ny;
                const double tau = S.nu * (ux - nx * un) / Fc(f, W_dlt);
                const double x = Onb[3 * qq];
                if (have && (sep ? (prevTau > 0 && tau <= 0) : (prevTau < 0 && tau >= 0))) {
                    const double den = sep ? prevTau - tau : tau - prevTau;
                    const double frac = fabs(den) > 1e-30 ? (sep ? prevTau : -prevTau) / den : 0.0;
                    value = prevX + frac * (x - prevX);
                    break;
                }
                prevTau = tau;
                prevX = x;
                have = true;
            }
            return value;
        }
        double value = 0.0;
        for (int pass = 0; pass < 2; ++pass) {
            const int f = pass == 0 ? Q[4] : Q[5];
            if (f < 0) break;
            const int ow = fon[f].x;
            const double delta = Fc(f, W_dlt);
            const double nx = Fc(f, W_nx), ny = Fc(f, W_ny);
            const double un = C(ow, V_Ux) * nx + C(ow, V_Uy) * ny;
            const double utx = C(ow, V_Ux) - nx * un, uty = C(ow, V_Uy) - ny * un;
            const double tau = S.nu * sqrt(utx * utx + uty * uty) / delta;
            const double refv = R[O_refv];
            const double dynP = 0.5 * refv * refv;
            const double cf = tau / smax(dynP, 1e-20);
            value = pass == 0 ? cf : (1.0 - R[O_t]) * value + R[O_t] * cf;
        }
        return value;
    }

    // per-phase clock accounting (thread 0, between block-wide barriers)
    DEV void mark(int phase) {
        if (S.profile && tid == 0) {
            const long long now = clock64();
            bk->cyc[phase] += now - bk->tlast;
            bk->tlast = now;
        }
    }

    DEV void fail(int code) {
        if (tid == 0) {
            bk->done = 1;
            bk->status = ST_UNKNOWN;
            bk->error = code;
        }
        __syncthreads();
    }

    // ----------------------------------------------------------- the solve
    // SIMPLESolver::solve (SIMPLESolver.cpp:1193-1716) after initUniform
    // (SIMPLESolver.cpp:267-298).
    // warm start: the cached state replaces the uniform initialisation and
    // the boundary-condition pass (ForwardModel.cpp:85-88: a warm-started
    // solve begins from the stored fields, boundary values and turbulence
    // flag as they are)
    const double *warm = nullptr;
    bool warm_bcs = false;
    DEV void load_state() {
        const int64_t nbo = (int64_t)kStateInterior * nc;
        const int vars[kStateInterior] = {V_Ux, V_Uy, V_p, V_k, V_w, V_nuT, V_F1, V_F2, V_Pk, V_CDkw};
        const int bvars[kStateBoundary] = {B_Ux, B_Uy, B_p, B_k, B_w};
        for (int ci = gt; ci < nc; ci += gs) {
#pragma unroll
            for (int a = 0; a < kStateInterior; ++a) C(ci, vars[a]) = warm[(int64_t)a * nc + ci];
            const double wy = M.wall_y[ci];
            C(ci, V_wallv) = isfinite(wy) ? resolved_wall_omega(wy) : NAN;
        }
        for (int b = gt; b < nb; b += gs)
#pragma unroll
            for (int a = 0; a < kStateBoundary; ++a) Bd(b, bvars[a]) = warm[nbo + (int64_t)a * nbs + b];
    }

    DEV void run() {
        const double nu = S.nu;
        if (warm) {
            load_state();
        } else {
            for (int ci = gt; ci < nc; ci += gs) {
                C(ci, V_Ux) = S.init_ux;
                C(ci, V_Uy) = S.init_uy;
                C(ci, V_p) = S.init_p;
                C(ci, V_k) = S.init_k;
                C(ci, V_F1) = 1.0;
                C(ci, V_F2) = 1.0;
                C(ci, V_Pk) = 0.0;
                C(ci, V_CDkw) = 0.0;
                const double y = smax(C(ci, V_wd), 1e-20);
                const double om = smax(resolved_wall_omega(y), S.init_w);
                C(ci, V_w) = om;
                C(ci, V_nuT) = c[9] * S.init_k / smax(c[9] * om, 1e-20);
                const double wy = M.wall_y[ci];
                C(ci, V_wallv) = isfinite(wy) ? resolved_wall_omega(wy) : NAN;
            }
        }
        if (tid == 0) {
            bk->done = 0;
            bk->status = ST_UNCONVERGED;
            bk->error = ERR_NONE;
            bk->finalIter = 0;
            bk->stalled = 0;
            bk->haveQPrev = 0;
            bk->turbEst = warm ? (warm[(int64_t)kStateInterior * nc +
                                       (int64_t)kStateBoundary * nbs] != 0.0) : 0;
            bk->lastK = bk->lastOm = 1.0;
            bk->normUx0 = bk->normUy0 = bk->normP0 = 1.0;
            bk->pfRef = 0.0;
            bk->worst = 0.0;
            bk->rate = 0.0;
            for (int e = 0; e < 4; ++e) bk->nEst[e] = bk->pos[e] = 0;
            for (int e = 0; e < 5; ++e) bk->kry[e] = 0;
            for (int e = 0; e < kPhases; ++e) bk->cyc[e] = 0;
            bk->tlast = clock64();
        }
        sync();
        if (!warm) {
            bc_velocity();
            bc_pressure();
            bc_k();
            bc_omega(0.075);
            sync();
        } else if (warm_bcs) {  // with the row's own beta1
            bc_velocity();
            bc_pressure();
            bc_k();
            bc_omega(c[2]);
            sync();
        }
        for (int f = gt; f < nf; f += gs) Fc(f, W_phi) = interp_flux(f);
        sync();

        const double tol = S.convergence_tol;
        const double negligible = 0.01 * tol;
        int turbEst = bk->turbEst;
        for (int iter = 0; iter < S.max_iterations; ++iter) {
            for (int ci = gt; ci < nc; ci += gs) {
                C(ci, V_UxP) = C(ci, V_Ux);
                C(ci, V_UyP) = C(ci, V_Uy);
            }
            for (int f = gt; f < nif; f += gs) Fc(f, W_phiBar) = interp_flux(f);
            const bool turbActive = iter >= S.turb_start_iter;
            const bool turbUpdate =
                turbActive && ((iter - S.turb_start_iter) % S.turb_update_interval == 0);
            if (turbUpdate) {
                if (!turbEst && iter >= S.turb_start_iter + S.nut_floor_iters) turbEst = 1;
                const double nuTMin = turbEst ? 0.0 : 0.1 * nu;
                if constexpr (QL == 2) {
                    for (int ci = ct; ci < nc; ci += cs) turbulence_fields_pair(ci, nuTMin);
                } else
                for (int ci = gt; ci < nc; ci += gs) {
                    const int ph5[5] = {V_Ux, V_Uy, V_k, V_w, V_p};
                    const int pb5[5] = {B_Ux, B_Uy, B_k, B_w, B_p};
                    double gx[5], gy[5];
                    gg_cell<5>(ci, ph5, pb5, gx, gy);
                    C(ci, V_gUxx) = gx[0];
                    C(ci, V_gUxy) = gy[0];
                    C(ci, V_gUyx) = gx[1];
                    C(ci, V_gUyy) = gy[1];
                    C(ci, V_gKx) = gx[2];
                    C(ci, V_gKy) = gy[2];
                    C(ci, V_gWx) = gx[3];
                    C(ci, V_gWy) = gy[3];
                    C(ci, V_gPx) = gx[4];
                    C(ci, V_gPy) = gy[4];
                    // SSTModel::computeFields (SSTModel.cpp:208-240) and the
                    // startup nuT floor (SIMPLESolver.cpp:1333-1337)
                    const double S11 = gx[0], S22 = gy[1];
                    const double S12 = 0.5 * (gy[0] + gx[1]);
                    const double Sc = sqrt(2.0 * (S11 * S11 + S22 * S22 + 2.0 * (S12 * S12)));
                    C(ci, V_S) = Sc;
                    const double kc = smax(C(ci, V_k), 0.0);
                    const double wc = smax(C(ci, V_w), 1e-20);
                    const double y = smax(C(ci, V_wd), 1e-20);
                    const double omSafe = smax(wc, 1e-20);
                    const double cd = 2.0 * c[5] / omSafe * (gx[2] * gx[3] + gy[2] * gy[3]);
                    C(ci, V_CDkw) = cd;
                    C(ci, V_F1) = (S.sst_variant == 2) ? 1.0 : sst_f1(kc, wc, y, cd);
                    const double f2 = sst_f2(kc, wc, y);
                    C(ci, V_F2) = f2;
                    const double nt = sst_nut(kc, wc, Sc, f2);
                    C(ci, V_Pk) = sst_pk(nt, Sc, kc, wc);
                    C(ci, V_nuT) = smax(nt, nuTMin);
                }
            } else {
                for (int ci = gt; ci < nc; ci += gs) {
                    const int ph3[3] = {V_Ux, V_Uy, V_p};
                    const int pb3[3] = {B_Ux, B_Uy, B_p};
                    double gx[3], gy[3];
                    gg_cell<3>(ci, ph3, pb3, gx, gy);
                    C(ci, V_gUxx) = gx[0];
                    C(ci, V_gUxy) = gy[0];
                    C(ci, V_gUyx) = gx[1];
                    C(ci, V_gUyy) = gy[1];
                    C(ci, V_gPx) = gx[2];
                    C(ci, V_gPy) = gy[2];
                }
            }
            sync();
            mark(0);

            // ---- x momentum (SIMPLESolver.cpp:1342-1379)
            double momResid, momScale;
            Krylov rUx, rUy;
            {
                double a[4] = {0.0, 0.0, 0.0, 0.0};
                auto xrow = [&](int ci, double r, double aPu, bool ok) {
                    const double ux = C(ci, V_Ux), uy = C(ci, V_Uy);
                    C(ci, V_kr) = r;
                    C(ci, V_kx) = ux;
                    a[0] += fabs(r);
                    const double contribution = aPu * sqrt(ux * ux + uy * uy);
                    a[1] += contribution;
                    a[2] += ok ? 0.0 : 1.0;
                    a[3] += r * r;
                };
                if constexpr (QL > 1) {
                    for (int ci = ct; ci < nc; ci += cs) {
                        double aPu;
                        bool ok;
                        const double r = momentum_row_q(ci, 0, aPu, ok);
                        if (qi == 0) xrow(ci, r, aPu, ok);
                    }
                } else {
                    for (int ci = gt; ci < nc; ci += gs) {
                        double aPu;
                        bool ok;
                        const double r = momentum_row(ci, 0, aPu, ok);
                        xrow(ci, r, aPu, ok);
                    }
                }
                red.sum(a);
                momResid = a[0];
                momScale = a[1];
                mark(1);
                if (a[2] != 0.0) {
                    fail(ERR_NONFINITE_SYSTEM);
                    break;
                }
                rUx = solve(S.momentum_solver, a[3], false, V_Ux);
                mark(2);
            }
            // ---- y momentum: its transpose stress sees the new Ux
            {
                for (int ci = gt; ci < nc; ci += gs) {
                    const int ph1[1] = {V_Ux};
                    const int pb1[1] = {B_Ux};
                    double gx[1], gy[1];
                    gg_cell<1>(ci, ph1, pb1, gx, gy);
                    C(ci, V_gUxx) = gx[0];
                    C(ci, V_gUxy) = gy[0];
                }
                sync();
                double a[3] = {0.0, 0.0, 0.0};
                auto yrow = [&](int ci, double r, bool ok) {
                    C(ci, V_kr) = r;
                    C(ci, V_kx) = C(ci, V_Uy);
                    a[0] += fabs(r);
                    a[1] += ok ? 0.0 : 1.0;
                    a[2] += r * r;
                };
                if constexpr (QL > 1) {
                    for (int ci = ct; ci < nc; ci += cs) {
                        double aPu;
                        bool ok;
                        const double r = momentum_row_q(ci, 1, aPu, ok);
                        if (qi == 0) yrow(ci, r, ok);
                    }
                } else {
                    for (int ci = gt; ci < nc; ci += gs) {
                        double aPu;
                        bool ok;
                        const double r = momentum_row(ci, 1, aPu, ok);
                        yrow(ci, r, ok);
                    }
                }
                red.sum(a);
                momResid = smax(momResid, a[0]);
                mark(3);
                if (a[1] != 0.0) {
                    fail(ERR_NONFINITE_SYSTEM);
                    break;
                }
                rUy = solve(S.momentum_solver, a[2], false, V_Uy);
                mark(4);
            }
            bc_velocity();
            sync();

            // ---- predicted Rhie-Chow flux and the pressure correction
            //      (SIMPLESolver.cpp:1410-1432, 717-830, 870-884)
            const double relaxCorr = 1.0 - S.alpha_u;
            double ms[1] = {0.0};
            for (int f = gt; f < nf; f += gs) {
                double m = interp_flux(f);
                double coeff = 0.0;
                if (f < nif) {
                    const v2i on = fon[f];
                    const int o = on.x, n = on.y;
                    const double delta = Fc(f, W_dlt);
                    const double ao = C(o, V_aPs), an = C(n, V_aPs);
                    const double dP_o = (fabs(ao) > 1e-30) ? C(o, V_vol) / ao : 0.0;
                    const double dP_n = (fabs(an) > 1e-30) ? C(n, V_vol) / an : 0.0;
                    const double wt = Fc(f, W_w), wmv = Fc(f, W_wm);
                    const double dP_f = wt * dP_o + wmv * dP_n;
                    const double gbx = C(o, V_gPx) * wt + C(n, V_gPx) * wmv;
                    const double gby = C(o, V_gPy) * wt + C(n, V_gPy) * wmv;
                    const double Af = Fc(f, W_A);
                    m += -dP_f * Af * ((C(n, V_p) - C(o, V_p)) / delta - (gbx * Fc(f, W_ex) + gby * Fc(f, W_ey)));
                    m += relaxCorr * (Fc(f, W_phi) - Fc(f, W_phiBar));
                    coeff = dP_f * Af / delta;
                } else if (bflag[f - nif] & 2) {
                    const int o = fon[f].x;
                    const double ao = C(o, V_aPs);
                    const double dP_o = (fabs(ao) > 1e-30) ? C(o, V_vol) / ao : 0.0;
                    coeff = dP_o * Fc(f, W_A) / Fc(f, W_dlt);
                }
                Fc(f, W_phiPred) = m;
                Fc(f, W_pCoeff) = coeff;
                ms[0] += fabs(m);
            }
            sync();
            double a8[3] = {0.0, 0.0, 0.0};
            for (int ci = gt; ci < nc; ci += gs) {
                SlotSet q;
                slot_set(ci, q);
                double dg = 0.0, sr = 0.0, offv[kSlots];
#pragma unroll
                for (int s = 0; s < kSlots; ++s) {
                    const int f = q.f[s];
                    const double coeff = Fc(f, W_pCoeff), pred = Fc(f, W_phiPred);
                    const bool outlet = bflag[q.b[s]] & kFlagOutlet;
                    const bool in = q.inner[s];
                    // internal faces and outlet faces add their coefficient;
                    // the predicted flux leaves the owner and enters the
                    // neighbour, and leaves through every boundary face
                    dg = (q.valid[s] && (in || outlet)) ? dg + coeff : dg;
                    const double t = (!in || q.own[s]) ? sr - pred : sr + pred;
                    sr = q.valid[s] ? t : sr;
                    offv[s] = -coeff;
                    if (in) off(ci, s) = offv[s];
                }
                if (!has_outlet && ci == 0) dg += fabs(dg) > 1e-30 ? dg : 1.0;
                C(ci, V_diag) = dg;
                C(ci, V_src) = sr;
                C(ci, V_kr) = sr;  // b - A x at x = 0
                C(ci, V_kx) = 0.0;
                bool okp = isfinite(dg) && isfinite(sr);
#pragma unroll
                for (int s = 0; s < kSlots; ++s)
                    okp = okp && (!q.inner[s] || isfinite(offv[s]));
                a8[0] += fabs(sr);
                a8[1] += sr * sr;
                a8[2] += okp ? 0.0 : 1.0;
            }
            // one reduction for the flux scale and the p' row sums
            double a4[4] = {ms[0], a8[0], a8[1], a8[2]};
            red.sum(a4);
            ms[0] = a4[0];
            a8[0] = a4[1];
            a8[1] = a4[2];
            a8[2] = a4[3];
            const double massFluxScale = ms[0];
            mark(5);
            const double contScaled = a8[0] / smax(massFluxScale, 1e-300);
            if (a8[2] != 0.0) {
                fail(ERR_NONFINITE_SYSTEM);
                break;
            }
            const Krylov rP = solve(S.pressure_solver, a8[1], true, V_pP);
            mark(6);
            for (int b = gt; b < nb; b += gs) Bd(b, B_pP) = (bflag[b] & 2) ? 0.0 : C(fon[nif + b].x, V_pP);
            sync();
            // ---- correctVelocity / correctPressure (SIMPLESolver.cpp:916-940)
            for (int ci = gt; ci < nc; ci += gs) {
                const int ph1[1] = {V_pP};
                const int pb1[1] = {B_pP};
                double gx[1], gy[1];
                gg_cell<1>(ci, ph1, pb1, gx, gy);
                const double ac = C(ci, V_aPs);
                const double rAP = (fabs(ac) > 1e-30) ? 1.0 / ac : 0.0;
                const double V = C(ci, V_vol);
                C(ci, V_Ux) -= rAP * gx[0] * V;
                C(ci, V_Uy) -= rAP * gy[0] * V;
                C(ci, V_p) += S.alpha_p * C(ci, V_pP);
            }
            sync();
            bc_velocity();
            bc_pressure();
            for (int f = gt; f < nf; f += gs) {
                const v2i on = fon[f];
                if (f >= nif) Fc(f, W_phi) = Fc(f, W_phiPred) + Fc(f, W_pCoeff) * C(on.x, V_pP);
                else Fc(f, W_phi) = Fc(f, W_phiPred) - Fc(f, W_pCoeff) * (C(on.y, V_pP) - C(on.x, V_pP));
            }
            sync();
            mark(7);

            // velocity maxima of the convergence norms (U is final for this
            // iteration: the turbulence transport below does not change it)
            double mu[2] = {1e-300, 0.0};
            for (int ci = gt; ci < nc; ci += gs) {
                const double ux = C(ci, V_Ux), uy = C(ci, V_Uy);
                mu[0] = smax(mu[0], sqrt(ux * ux + uy * uy));
                const double dx = ux - C(ci, V_UxP), dy = uy - C(ci, V_UyP);
                mu[1] = smax(mu[1], sqrt(dx * dx + dy * dy));
            }

            // ---- turbulence transport (SIMPLESolver.cpp:1476-1538)
            if (turbUpdate) {
                double a9[2] = {0.0, 0.0};
                auto krow = [&](int ci, double r, bool ok) {
                    C(ci, V_kx) = C(ci, V_k);
                    C(ci, V_kr) = r;
                    a9[1] += ok ? 0.0 : 1.0;
                    a9[0] += r * r;
                };
                if constexpr (QL > 1) {
                    for (int ci = ct; ci < nc; ci += cs) {
                        bool ok;
                        double r;
                        k_row_q(ci, ok, r);
                        if (qi == 0) krow(ci, r, ok);
                    }
                } else {
                    for (int ci = gt; ci < nc; ci += gs) {
                        bool ok;
                        double r;
                        k_row(ci, ok, r);
                        krow(ci, r, ok);
                    }
                }
                red.sum(a9);
                mark(8);
                if (a9[1] != 0.0) {
                    fail(ERR_NONFINITE_SYSTEM);
                    break;
                }
                const Krylov rK = solve(S.turbulence_solver, a9[0], false);
                mark(9);
                double mk[2] = {1e-30, 0.0};
                for (int ci = gt; ci < nc; ci += gs) {
                    const double kn = smax(S.k_min, smin(S.k_max, C(ci, V_kx)));
                    mk[0] = smax(mk[0], fabs(kn));
                    mk[1] = smax(mk[1], fabs(kn - C(ci, V_k)));
                    C(ci, V_k) = kn;
                }
                sync();
                bc_k();
                sync();
                double a10[2] = {0.0, 0.0};
                auto wrow = [&](int ci, double x0, double r, bool ok) {
                    C(ci, V_kx) = x0;
                    C(ci, V_kr) = r;
                    a10[1] += ok ? 0.0 : 1.0;
                    a10[0] += r * r;
                };
                if constexpr (QL > 1) {
                    for (int ci = ct; ci < nc; ci += cs) {
                        bool ok;
                        double r;
                        const double x0 = omega_row_q(ci, ok, r);
                        if (qi == 0) wrow(ci, x0, r, ok);
                    }
                } else {
                    for (int ci = gt; ci < nc; ci += gs) {
                        bool ok;
                        double r;
                        const double x0 = omega_row(ci, ok, r);
                        wrow(ci, x0, r, ok);
                    }
                }
                red.sum(a10);
                mark(10);
                if (a10[1] != 0.0) {
                    fail(ERR_NONFINITE_SYSTEM);
                    break;
                }
                const Krylov rW = solve(S.turbulence_solver, a10[0], false);
                mark(11);
                double mw[2] = {1e-30, 0.0};
                for (int ci = gt; ci < nc; ci += gs) {
                    double wn = smax(S.omega_min, smin(1e15, C(ci, V_kx)));
                    const double wvc = C(ci, V_wallv);
                    if (isfinite(wvc)) wn = wvc;
                    mw[0] = smax(mw[0], fabs(wn));
                    mw[1] = smax(mw[1], fabs(wn - C(ci, V_w)));
                    C(ci, V_w) = wn;
                }
                sync();
                bc_omega(c[2]);
                // one reduction for the k/omega change maxima and the velocity norms
                double m4[6] = {mk[0], mk[1], mw[0], mw[1], mu[0], mu[1]};
                red.max(m4);
                mu[0] = m4[4];
                mu[1] = m4[5];
                if (tid == 0) {
                    const double kChange = m4[1] / m4[0], omChange = m4[3] / m4[2];
                    bk->lastK = kChange;
                    bk->lastOm = omChange;
                    ce_push(1, kChange);
                    ce_push(2, omChange);
                    bk->kry[3] += rK.iterations;
                    bk->kry[4] += rW.iterations;
                }
            }

            // ---- convergence bookkeeping (SIMPLESolver.cpp:1545-1705)
            if (!turbUpdate) red.max(mu);
            mark(12);
            const bool monitor =
                nobs > 0 && iter % (S.monitor_interval > 1 ? S.monitor_interval : 1) == 0;
            if (monitor) {
                for (int i = tid; i < nobs; i += BS) q[i] = observe(i);
                __syncthreads();
            }
            if (tid == 0) {
                bk->turbEst = turbEst;
                bk->kry[0] += rUx.iterations;
                bk->kry[1] += rUy.iterations;
                bk->kry[2] += rP.iterations;
            }
            bookkeeping(iter, turbActive, momResid, momScale, contScaled, massFluxScale,
                        rUx.initial, rUy.initial, rP.initial, mu[0], mu[1], monitor, negligible,
                        tol);
            mark(13);
            if (bk->done) break;
        }
        if (tid == 0 && !bk->done) {
            bk->finalIter = S.max_iterations;
            bk->stalled = isfinite(bk->rate) && bk->rate >= 1.0;
            bk->status = ST_UNCONVERGED;
        }
        sync();
    }

    // The convergence decision of SIMPLESolver::solve (SIMPLESolver.cpp:
    // 1545-1705) for this iteration. Called by every thread; wave 0 does the
    // work (lane-parallel window maxima and monitor norms, scalar logic in
    // every lane, thread 0 writes the state).
    DEV void bookkeeping(int iter, bool turbActive, double momResid, double momScale,
                         double contScaled, double massFluxScale, double eUx, double eUy,
                         double eP, double uMax, double duMax, bool monitor, double negligible,
                         double tol) {
        if (tid < 64) {
            const bool lane0 = tid == 0;
            const double momentumScaled = momResid / smax(momScale, 1e-300);
            const double changeU = duMax / uMax;
            if (lane0) ce_push(0, changeU);
            bool qbad = false;
            if (monitor) {
                                const bool have = bk->haveQPrev;
                double bad = 0.0, dq = 0.0;
                for (int i = tid; i < nobs; i += 64) {
                    const double qi = q[i];
                    if (!isfinite(qi)) bad = 1.0;
                    const double den = smax(smax(fabs(qi), Ob[__mul24(i, kNO) + O_floor]), 1e-300);
                    const double x = fabs(qi - qprev[i]) / den;
                    if (have) dq = smax(dq, (x != x) ? 0.0 : x);
                }
                bad = wave_max(bad);
                dq = wave_max(dq);
                if (bad != 0.0) {
                    qbad = true;
                    if (lane0) {
                        bk->done = 1;
                        bk->status = ST_UNKNOWN;
                        bk->error = ERR_NONFINITE_MONITOR;
                    }
                } else {
                    if (have && lane0) ce_push(3, dq);
                    for (int i = tid; i < nobs; i += 64) qprev[i] = q[i];
                    if (lane0) bk->haveQPrev = 1;
                }
            }
            // lane 0's ring pushes must be visible to the other lanes of the wave
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
            if (!qbad) {
                // The window maxima and rates only decide anything when every
                // last-sample term is already below tol (the worst measure is
                // their maximum, or non-finite), and they are reported only at
                // the final iteration: otherwise the full measure is skipped.
                const double lastK = bk->lastK, lastOm = bk->lastOm;
                const double lastQ = nobs > 0 ? ce_last(3) : 0.0;
                double cheap = smax(smax(momentumScaled, contScaled), changeU);
                if (turbActive) cheap = smax(smax(cheap, lastK), lastOm);
                if (nobs > 0) cheap = smax(cheap, lastQ);
                double normUx0 = bk->normUx0, normUy0 = bk->normUy0, normP0 = bk->normP0;
                double pfRef = bk->pfRef;
                if (iter == 0) {
                    normUx0 = smax(eUx, 1e-30);
                    normUy0 = smax(eUy, 1e-30);
                    normP0 = smax(eP, 1e-30);
                    pfRef = massFluxScale;
                } else if (iter < 50) {
                    normP0 = smax(normP0, eP);
                    pfRef = smax(pfRef, massFluxScale);
                }
                const double normMom0 = smax(normUx0, normUy0);
                const double nUx = normMom0 < 1e-20 ? 0.0 : eUx / normMom0;
                const double nUy = normMom0 < 1e-20 ? 0.0 : eUy / normMom0;
                const double pref = smax(normP0, pfRef);
                const double nP = pref < 1e-20 ? 0.0 : eP / pref;
                double maxRes = smax(smax(nUx, nUy), nP);
                const double eK = turbActive ? lastK : 0.0, eOm = turbActive ? lastOm : 0.0;
                if (turbActive) maxRes = smax(smax(maxRes, lastK), lastOm);
                const bool nonfinite =
                    !isfinite(eUx) || !isfinite(eUy) || !isfinite(eP) || !isfinite(eK) ||
                    !isfinite(eOm) || !isfinite(lastK) || !isfinite(lastOm) ||
                    !isfinite(momentumScaled) || !isfinite(contScaled) || !isfinite(changeU);
                if (nonfinite) maxRes = INFINITY;
                const bool diverging = isnan(maxRes) || isinf(maxRes) || maxRes > S.divergence_limit;
                const bool last = iter + 1 >= S.max_iterations;
                double worst = INFINITY, rate0 = bk->rate;
                if (!(cheap >= tol) || diverging || last) {
                    double dNew[4], dOld[4];
                    int n[4];
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        n[e] = bk->nEst[e];
                        if (n[e] >= 2 * S.convergence_window) ce_windows(e, dNew[e], dOld[e]);
                        else dNew[e] = dOld[e] = 0.0;
                    }
                    rate0 = ce_rate(n[0], dNew[0], dOld[0]);
                    const double remainingU = ce_remaining(n[0], dNew[0], rate0, negligible);
                    worst = smax(smax(smax(momentumScaled, contScaled), changeU), remainingU);
                    if (turbActive) {
                        worst = smax(worst, lastK);
                        worst = smax(worst, ce_remaining(n[1], dNew[1],
                                                         ce_rate(n[1], dNew[1], dOld[1]), negligible));
                        worst = smax(worst, lastOm);
                        worst = smax(worst, ce_remaining(n[2], dNew[2],
                                                         ce_rate(n[2], dNew[2], dOld[2]), negligible));
                    }
                    if (nobs > 0) {
                        worst = smax(worst, lastQ);
                        worst = smax(worst, ce_remaining(n[3], dNew[3],
                                                         ce_rate(n[3], dNew[3], dOld[3]), negligible));
                    }
                    if (!isfinite(worst)) worst = INFINITY;
                    if (nonfinite) worst = INFINITY;
                }
                const bool turbScheduled = S.turb_start_iter < S.max_iterations;
                if (lane0) {
                    bk->normUx0 = normUx0;
                    bk->normUy0 = normUy0;
                    bk->normP0 = normP0;
                    bk->pfRef = pfRef;
                    bk->worst = worst;
                    bk->rate = rate0;
                    if (worst < tol && iter > 0 && (!turbScheduled || (turbActive && bk->turbEst))) {
                        bk->done = 1;
                        bk->status = ST_CONVERGED;
                        bk->finalIter = iter;
                    } else if (diverging) {
                        bk->done = 1;
                        bk->status = ST_DIVERGED;
                        bk->finalIter = iter;
                    }
                }
            }
        }
        __syncthreads();
    }
};

// ----------------------------------------------------------------------------
// One draw, start to finish, by one workgroup: the arena starts as a copy of
// the mesh's pre-filled image (geometry and int tables), the solve runs, the
// observations, fields and result record are written. Shared by the batch
// kernel and the serving kernel.
// ----------------------------------------------------------------------------
template <int BS>
struct BlockShared {
    Book book;
    double redbuf[2 * (BS / 64) * Reducer<BS>::kStride];
    double coeffs[11];
    DevSettings sset;
    Krylov kslot;
    GridLocal gl;
    double coarse_b[kMaxAmgCoarse];
};

// rank/nblk/grid/aux describe a multi-workgroup draw (kGrid): its workgroups
// share the arena and meet at grid barriers; aux is this workgroup's LDS for
// its bookkeeping replica (contraction rings and monitor values).
template <int BS, bool kLDS, bool kGrid = false, bool kF32 = false, bool kServe = false>
DEV void solve_draw(BlockShared<BS> &sh, const DevMesh &M, const DevSettings &settings,
                    const double *coeffSrc, double *arena, double *amg, DevResult *res,
                    double *pred, double *fields, GridState *grid = nullptr, int rank = 0,
                    int nblk = 1, double *aux = nullptr, const double *warm = nullptr,
                    double *save = nullptr, double tol = 0.0, int max_iter = 0, int flags = 0) {
    if (threadIdx.x < 11) sh.coeffs[threadIdx.x] = coeffSrc[threadIdx.x];
    if (threadIdx.x == 0) {
        sh.sset = settings;
        if (tol > 0.0) sh.sset.convergence_tol = tol;
        if (max_iter > 0) sh.sset.max_iterations = max_iter;
    }
    const int g0 = rank * BS + threadIdx.x, gstride = nblk * BS;
    {
        const int64_t n16 = M.lay.image_bytes / 16;
        const uint4 *gsrc = reinterpret_cast<const uint4 *>(kLDS ? M.img : M.img_soa);
        uint4 *dst = reinterpret_cast<uint4 *>(arena);
        for (int64_t i = g0; i < n16; i += gstride) dst[i] = gsrc[i];
    }
    if constexpr (kGrid) {
        // the first barrier is agent-scope; it also publishes each
        // workgroup's XCD, and later barriers stay within the L2 when the
        // draw's workgroups all share one XCD
        if (threadIdx.x == 0) {
            sh.gl.target = 0;
            sh.gl.local = 0;
            sh.gl.wait = 0;
            sh.gl.count = 0;
            grid->xcc[rank] = (int)__builtin_amdgcn_s_getreg(kHwRegXccId);
        }
        __syncthreads();
        grid_barrier(grid, nblk, &sh.gl);
        if (threadIdx.x == 0) {
            int same = 1;
            for (int r = 1; r < nblk; ++r) same &= grid->xcc[r] == grid->xcc[0];
            sh.gl.local = same;
        }
        __syncthreads();
    } else {
        __syncthreads();
    }
    Draw<BS, kLDS, kGrid, kF32, kServe> d(M, sh.sset, sh.coeffs, arena, &sh.book, sh.redbuf, &sh.kslot, amg,
                            aux);
    d.cbuf = sh.coarse_b;
    d.warm = warm;
    d.warm_bcs = (flags & kSlotWarmBCs) != 0;
    if (flags & kSlotNoMonitor) d.nobs = 0;  // the final observation still runs
    if constexpr (kGrid) {
        d.rank = rank;
        d.nblk = nblk;
        d.grid = grid;
        d.gl = &sh.gl;
        d.gt = g0;
        d.gs = gstride;
        d.red.grid = grid;
        d.red.nblk = nblk;
        d.red.rank = rank;
        d.red.gl = &sh.gl;
    }
    d.run();
    // every workgroup's last writes are complete before rank 0 observes
    if constexpr (kGrid) grid_barrier(grid, nblk, &sh.gl);
    const Book &book = sh.book;
    const bool observe = book.status == ST_CONVERGED || book.status == ST_UNCONVERGED;
    if (rank == 0)
        for (int i = threadIdx.x; i < M.nobs; i += BS) pred[i] = observe ? d.observe(i) : NAN;
    if (fields) {
        const int nc = M.nc;
        const int vars[10] = {V_Ux, V_Uy, V_p, V_k, V_w, V_nuT, V_F1, V_F2, V_Pk, V_CDkw};
        auto ld = [&](int64_t k) {
            if constexpr (kGrid)
                return __hip_atomic_load(arena + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            else
                return arena[k];
        };
        for (int a = 0; a < 10; ++a)
            for (int ci = g0; ci < nc; ci += gstride)
                fields[a * nc + ci] = ld(kLDS ? (int64_t)ci * kNA + vars[a]
                                              : (int64_t)vars[a] * nc + ci);
        for (int f = g0; f < M.nf; f += gstride)
            fields[10 * nc + f] = ld(M.lay.face + (kLDS ? (int64_t)f * kNF + W_phi
                                                        : (int64_t)W_phi * M.nf + f));
    }
    if (save) {  // the final flow state (its layout: state_doubles)
        const int nc = M.nc, nb = M.nb, nbs = nb > 1 ? nb : 1;
        const int64_t nbo = (int64_t)kStateInterior * nc;
        const int vars[kStateInterior] = {V_Ux, V_Uy, V_p, V_k, V_w, V_nuT, V_F1, V_F2, V_Pk, V_CDkw};
        const int bvars[kStateBoundary] = {B_Ux, B_Uy, B_p, B_k, B_w};
        for (int a = 0; a < kStateInterior; ++a)
            for (int ci = g0; ci < nc; ci += gstride) save[(int64_t)a * nc + ci] = d.C(ci, vars[a]);
        for (int a = 0; a < kStateBoundary; ++a)
            for (int b = g0; b < nb; b += gstride) save[nbo + (int64_t)a * nbs + b] = d.Bd(b, bvars[a]);
        if (rank == 0 && threadIdx.x == 0) {
            save[nbo + (int64_t)kStateBoundary * nbs] = sh.book.turbEst ? 1.0 : 0.0;
            save[nbo + (int64_t)kStateBoundary * nbs + 1] = 0.0;
        }
    }
    if (rank == 0 && threadIdx.x == 0) {
        DevResult &r = *res;
        r.status = book.status;
        r.error = book.error;
        r.iters = book.status == ST_UNKNOWN ? 0 : book.finalIter;
        r.worst = book.status == ST_UNKNOWN ? 0.0 : book.worst;
        r.rate = book.status == ST_UNKNOWN ? 0.0 : book.rate;
        r.stalled = book.stalled;
        r.converged = book.status == ST_CONVERGED;
        r.diverged = book.status == ST_DIVERGED;
        for (int e = 0; e < 5; ++e) r.krylov[e] = book.kry[e];
        for (int e = 0; e < kPhases; ++e) r.cycles[e] = book.cyc[e];
        if (kGrid && settings.profile) {  // grid barriers: clocks waited, count
            r.cycles[kPhases - 2] = sh.gl.wait;
            r.cycles[kPhases - 1] = sh.gl.count;
        }
    }
}

// The batch kernel: blockIdx.x -> one slot (draw); in LDS mode the arena is
// the workgroup's dynamic LDS.
template <int BS, bool kLDS, bool kF32>
__global__ void __launch_bounds__(BS) simple_solve_kernel(const DevSlot *__restrict__ slots,
                                                          DevSettings settings,
                                                          DevResult *__restrict__ results,
                                                          double *__restrict__ predictions,
                                                          int pred_stride, double *__restrict__ fields,
                                                          int64_t field_stride) {
    extern __shared__ double smem[];
    __shared__ BlockShared<BS> sh;
    const DevSlot &slot = slots[blockIdx.x];
    const int row = slot.row;
    solve_draw<BS, kLDS, false, kF32>(sh, *slot.mesh, settings, slot.coeffs,
                                      kLDS ? smem : slot.arena,
                         slot.amg, results + row, predictions + (int64_t)row * pred_stride,
                         fields ? fields + row * field_stride : nullptr, nullptr, 0, 1, nullptr,
                         slot.warm, slot.save, slot.tol, slot.max_iter, slot.flags);
}

// The multi-workgroup batch kernel: nblk co-resident workgroups per draw
// (launched cooperatively), rank = blockIdx.x % nblk.
template <int BS, bool kF32>
__global__ void __launch_bounds__(BS) grid_solve_kernel(const DevSlot *__restrict__ slots,
                                                        DevSettings settings,
                                                        DevResult *__restrict__ results,
                                                        double *__restrict__ predictions,
                                                        int pred_stride, double *__restrict__ fields,
                                                        int64_t field_stride, GridState *states,
                                                        int nblk, int nslots) {
    extern __shared__ double smem[];
    __shared__ BlockShared<BS> sh;
    // Workgroups are dispatched round-robin over the XCDs (workgroup b on XCD
    // b % kXcds). A draw's workgroups are taken from one XCD, so its barriers
    // and shared arena stay within that XCD's L2: draw x + kXcds * t owns the
    // workgroups x + kXcds * j with t * nblk <= j < (t + 1) * nblk.
    const int x = blockIdx.x % kXcds, j = blockIdx.x / kXcds;
    const int draw = x + kXcds * (j / nblk), rank = j % nblk;
    if (draw >= nslots) return;  // padding workgroup of an XCD with fewer draws
    const DevSlot &slot = slots[draw];
    const int row = slot.row;
    solve_draw<BS, false, true, kF32>(sh, *slot.mesh, settings, slot.coeffs, slot.arena, slot.amg,
                                results + row, predictions + (int64_t)row * pred_stride,
                                fields ? fields + row * field_stride : nullptr, states + draw,
                                rank, nblk, smem, slot.warm, slot.save, slot.tol,
                                slot.max_iter, slot.flags);
}

// ----------------------------------------------------------------------------
// The serving kernel: a persistent grid whose workgroups claim requests one
// at a time from a ring in host memory (published by the host through the
// ring's tail), solve each in their own arena and publish the result with a
// completion sequence number. Every wait is bounded: a workgroup returns when
// the host raises stop, or after spin_limit idle polls, so the grid always
// drains; the host relaunches it when new requests arrive.
// ----------------------------------------------------------------------------
template <int BS>
__global__ void __launch_bounds__(BS) serve_kernel(const DevProblem *__restrict__ problems,
                                                   ServeControl *ctl, const ServeRequest *ring,
                                                   ServeOutput *outs, int ring_size,
                                                   unsigned long long *head,
                                                   double *__restrict__ arenas,
                                                   int64_t arena_stride, double *amgs,
                                                   int64_t amg_stride, long long spin_limit) {
    extern __shared__ double smem[];
    __shared__ BlockShared<BS> sh;
    __shared__ long long claimed;
    __shared__ double coeffs[11];
    for (;;) {
        if (threadIdx.x == 0) {
            // Idle workgroups poll a device-memory mirror of the ring tail
            // (head[1], an L2 hit) and read the host's tail only every
            // host_every polls, publishing what they see to the mirror.
            // Reads of host memory from every CU at full rate saturate the
            // path to the host and slowed the busy workgroups 5-7x. Polls
            // are relaxed; the acquire fences run only when something new is
            // seen: before publishing a host tail (the host's entries
            // happen-before it) and after claiming a ticket.
            unsigned long long *mirror = head + 1;
            const int host_every = gridDim.x > 16 ? (int)(gridDim.x / 16) : 1;
            long long got = -1, idle = 0;
            for (int poll = (int)blockIdx.x;; ++poll) {
                const long long tail = (long long)__hip_atomic_load(mirror, __ATOMIC_RELAXED,
                                                                     __HIP_MEMORY_SCOPE_AGENT);
                unsigned long long h =
                    __hip_atomic_load(head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if ((long long)h < tail) {
                    if (__hip_atomic_compare_exchange_strong(head, &h, h + 1, __ATOMIC_RELAXED,
                                                             __ATOMIC_RELAXED,
                                                             __HIP_MEMORY_SCOPE_AGENT)) {
                        got = (long long)h;
                        break;
                    }
                    continue;
                }
                if (poll % host_every == 0) {
                    const long long ht =
                        __hip_atomic_load(&ctl->tail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                    if (ht > tail) {
                        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");
                        __hip_atomic_fetch_max(mirror, (unsigned long long)ht, __ATOMIC_RELEASE,
                                               __HIP_MEMORY_SCOPE_AGENT);
                        continue;
                    }
                    if (__hip_atomic_load(&ctl->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM))
                        break;
                }
                if (++idle > spin_limit) break;
                __builtin_amdgcn_s_sleep(100);
            }
            // the request's entries happen-before the mirror value it read
            if (got >= 0) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");
            claimed = got;
        }
        __syncthreads();
        const long long ticket = claimed;
        if (ticket < 0) return;
        const int slot = (int)(ticket % ring_size);
        const ServeRequest &rq = ring[slot];
        const DevProblem &P = problems[rq.problem];
        if (threadIdx.x < 11) coeffs[threadIdx.x] = rq.coeffs[threadIdx.x];
        __syncthreads();
        ServeOutput &o = outs[slot];
        double *arena = arenas + (int64_t)blockIdx.x * arena_stride;
        double *amg = amgs ? amgs + (int64_t)blockIdx.x * amg_stride : nullptr;
        if (P.lds)
            solve_draw<BS, true, false, false, true>(sh, P.mesh, P.settings, coeffs, smem, amg,
                                                     &o.res, o.predictions, nullptr);
        else
            solve_draw<BS, false, false, false, true>(sh, P.mesh, P.settings, coeffs, arena, amg,
                                                      &o.res, o.predictions, nullptr);
        // every thread's result stores reach host memory before the
        // completion number is published
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
        __syncthreads();
        if (threadIdx.x == 0)
            __hip_atomic_store(&o.done, ticket + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        __syncthreads();
    }
}

}  // namespace rq

// ----------------------------------------------------------------------------
// launch helpers (host side of this translation unit)
// ----------------------------------------------------------------------------
namespace rq {

template <int BS, bool kLDS, bool kF32>
static hipError_t launch_t(const DevSlot *slots, int nslots, const DevSettings &s,
                           DevResult *results, double *pred, int pred_stride, double *fields,
                           int64_t field_stride, size_t lds_bytes, hipStream_t stream) {
    auto kern = simple_solve_kernel<BS, kLDS, kF32>;
    if (kLDS) {
        hipError_t e = hipFuncSetAttribute((const void *)kern,
                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)lds_bytes);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kern, dim3(nslots), dim3(BS), kLDS ? lds_bytes : 0, stream, slots, s,
                       results, pred, pred_stride, fields, field_stride);
    return hipGetLastError();
}

template <bool kF32>
static hipError_t launch_simple_t(int block, bool lds, const DevSlot *slots, int nslots,
                                  const DevSettings &s, DevResult *results, double *pred,
                                  int pred_stride, double *fields, int64_t field_stride,
                                  size_t lds_bytes, hipStream_t stream) {
#define RQ_CASE(B)                                                                           \
    case B:                                                                                  \
        return lds ? launch_t<B, true, kF32>(slots, nslots, s, results, pred, pred_stride,   \
                                             fields, field_stride, lds_bytes, stream)        \
                   : launch_t<B, false, kF32>(slots, nslots, s, results, pred, pred_stride,  \
                                              fields, field_stride, lds_bytes, stream);
    switch (block) {
        RQ_CASE(64)
        RQ_CASE(128)
        RQ_CASE(256)
        RQ_CASE(512)
    default:
        return hipErrorInvalidValue;
    }
#undef RQ_CASE
}

template <bool kF32>
static int grid_co_resident_t(size_t lds_bytes) {
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 0;
    auto kern = grid_solve_kernel<kGridBlock, kF32>;
    if (hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds_bytes) != hipSuccess)
        return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, kGridBlock, lds_bytes) !=
        hipSuccess)
        return 0;
    return per * cus;
}

template <bool kF32>
static hipError_t launch_grid_t(const DevSlot *slots, int nslots, int nblk, const DevSettings &s,
                                DevResult *results, double *pred, int pred_stride,
                                double *fields, int64_t field_stride, GridState *states,
                                size_t lds_bytes, hipStream_t stream) {
    auto kern = grid_solve_kernel<kGridBlock, kF32>;
    hipError_t e = hipFuncSetAttribute((const void *)kern,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    if (e != hipSuccess) return e;
    DevSettings sc = s;
    void *args[] = {(void *)&slots, (void *)&sc, (void *)&results, (void *)&pred,
                    (void *)&pred_stride, (void *)&fields, (void *)&field_stride,
                    (void *)&states, (void *)&nblk, (void *)&nslots};
    const int per_xcd = (nslots + kXcds - 1) / kXcds;  // draws on the fullest XCD
    if (s.grid_plain_launch) {
        // the same grid as an ordinary dispatch (the caller has checked that
        // it fits the device at once)
        hipLaunchKernelGGL(kern, dim3(kXcds * per_xcd * nblk), dim3(kGridBlock), lds_bytes,
                           stream, slots, sc, results, pred, pred_stride, fields, field_stride,
                           states, nblk, nslots);
        return hipGetLastError();
    }
    return hipLaunchCooperativeKernel((const void *)kern, dim3(kXcds * per_xcd * nblk),
                                      dim3(kGridBlock), args, (unsigned)lds_bytes, stream);
}

template <int BS>
static hipError_t launch_serve_t(int grid, const DevProblem *problems, ServeControl *ctl,
                                 const ServeRequest *ring, ServeOutput *outs, int ring_size,
                                 unsigned long long *head, double *arenas, int64_t arena_stride,
                                 double *amgs, int64_t amg_stride, long long spin_limit,
                                 size_t lds_bytes, hipStream_t stream) {
    auto kern = serve_kernel<BS>;
    hipError_t e = hipFuncSetAttribute((const void *)kern,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(BS), lds_bytes, stream, problems, ctl, ring, outs,
                       ring_size, head, arenas, arena_stride, amgs, amg_stride, spin_limit);
    return hipGetLastError();
}

// Mixed precision runs in kernels of its own, compiled in a translation unit
// of its own (RQ_KERNEL_PART 1), so the FP32 block Krylov code neither shares
// register allocation with the FP64 kernels nor lengthens their build; the
// 512-thread serving kernel is built there too, to balance the two parts.
#if RQ_KERNEL_PART == 1
hipError_t launch_serve_large(int grid, const DevProblem *problems, ServeControl *ctl,
                              const ServeRequest *ring, ServeOutput *outs, int ring_size,
                              unsigned long long *head, double *arenas, int64_t arena_stride,
                              double *amgs, int64_t amg_stride, long long spin_limit,
                              size_t lds_bytes, hipStream_t stream) {
    return launch_serve_t<kServeBlockLarge>(grid, problems, ctl, ring, outs, ring_size, head,
                                            arenas, arena_stride, amgs, amg_stride, spin_limit,
                                            lds_bytes, stream);
}
hipError_t launch_simple_solve_f32(int block, bool lds, const DevSlot *slots, int nslots,
                                   const DevSettings &s, DevResult *results, double *pred,
                                   int pred_stride, double *fields, int64_t field_stride,
                                   size_t lds_bytes, hipStream_t stream) {
    return launch_simple_t<true>(block, lds, slots, nslots, s, results, pred, pred_stride, fields,
                                 field_stride, lds_bytes, stream);
}
int grid_co_resident_f32(size_t lds_bytes) { return grid_co_resident_t<true>(lds_bytes); }
hipError_t launch_grid_solve_f32(const DevSlot *slots, int nslots, int nblk,
                                 const DevSettings &s, DevResult *results, double *pred,
                                 int pred_stride, double *fields, int64_t field_stride,
                                 GridState *states, size_t lds_bytes, hipStream_t stream) {
    return launch_grid_t<true>(slots, nslots, nblk, s, results, pred, pred_stride, fields,
                               field_stride, states, lds_bytes, stream);
}
#else
hipError_t launch_serve_large(int grid, const DevProblem *problems, ServeControl *ctl,
                              const ServeRequest *ring, ServeOutput *outs, int ring_size,
                              unsigned long long *head, double *arenas, int64_t arena_stride,
                              double *amgs, int64_t amg_stride, long long spin_limit,
                              size_t lds_bytes, hipStream_t stream);
hipError_t launch_simple_solve_f32(int block, bool lds, const DevSlot *slots, int nslots,
                                   const DevSettings &s, DevResult *results, double *pred,
                                   int pred_stride, double *fields, int64_t field_stride,
                                   size_t lds_bytes, hipStream_t stream);
int grid_co_resident_f32(size_t lds_bytes);
hipError_t launch_grid_solve_f32(const DevSlot *slots, int nslots, int nblk,
                                 const DevSettings &s, DevResult *results, double *pred,
                                 int pred_stride, double *fields, int64_t field_stride,
                                 GridState *states, size_t lds_bytes, hipStream_t stream);

hipError_t launch_simple_solve(int block, bool lds, const DevSlot *slots, int nslots,
                               const DevSettings &s, DevResult *results, double *pred,
                               int pred_stride, double *fields, int64_t field_stride,
                               size_t lds_bytes, hipStream_t stream) {
    if (s.mixed_precision)
        return launch_simple_solve_f32(block, lds, slots, nslots, s, results, pred, pred_stride,
                                       fields, field_stride, lds_bytes, stream);
    return launch_simple_t<false>(block, lds, slots, nslots, s, results, pred, pred_stride,
                                  fields, field_stride, lds_bytes, stream);
}

int grid_co_resident(size_t lds_bytes, bool mixed) {
    return mixed ? grid_co_resident_f32(lds_bytes) : grid_co_resident_t<false>(lds_bytes);
}

// ----------------------------------------------------------------------------
// The allocation-free evaluation's device memory: a fixed region reserved
// when this code object is loaded on a GPU and held until the process ends
// (each GPU holds its own copy at its own address). Host code never hands the
// array's name to a kernel or a copy -- on the host that name is only a
// stand-in -- but the address hipGetSymbolAddress returns for the current
// device, looked up once per device and kept.
// ----------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

hipError_t noalloc_scratch(void **ptr, size_t *bytes) {
    static std::mutex mu;
    static void *addr[kMaxDevices] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        addr[dev] = p;
    }
    *ptr = addr[dev];
    *bytes = kNoAllocScratchBytes;
    return hipSuccess;
}

// every noalloc call on one GPU shares its region: callers hold this lock
// (one per device) from the first write to the last read of the region
std::mutex &noalloc_lock(int dev) {
    static std::mutex locks[kMaxDevices];
    return locks[dev < 0 || dev >= kMaxDevices ? 0 : dev];
}

// n doubles set to v: clears the region with a kernel (a memset node in a
// captured graph is avoided on purpose)
__global__ void __launch_bounds__(256) fill_kernel(double *__restrict__ p, int64_t n, double v) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

hipError_t launch_fill(double *p, int64_t n, double v, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 1024);
    hipLaunchKernelGGL(fill_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, p, n, v);
    return hipGetLastError();
}

hipError_t launch_grid_solve(const DevSlot *slots, int nslots, int nblk, const DevSettings &s,
                             DevResult *results, double *pred, int pred_stride, double *fields,
                             int64_t field_stride, GridState *states, size_t lds_bytes,
                             hipStream_t stream) {
    if (s.mixed_precision)
        return launch_grid_solve_f32(slots, nslots, nblk, s, results, pred, pred_stride, fields,
                                     field_stride, states, lds_bytes, stream);
    return launch_grid_t<false>(slots, nslots, nblk, s, results, pred, pred_stride, fields,
                                field_stride, states, lds_bytes, stream);
}

hipError_t launch_serve(int block, int grid, const DevProblem *problems, ServeControl *ctl,
                        const ServeRequest *ring, ServeOutput *outs, int ring_size,
                        unsigned long long *head, double *arenas, int64_t arena_stride,
                        double *amgs, int64_t amg_stride, long long spin_limit, size_t lds_bytes,
                        hipStream_t stream) {
    if (block == kServeBlockLarge)
        return launch_serve_large(grid, problems, ctl, ring, outs, ring_size, head, arenas,
                                  arena_stride, amgs, amg_stride, spin_limit, lds_bytes, stream);
    if (block != kServeBlock) return hipErrorInvalidValue;
    return launch_serve_t<kServeBlock>(grid, problems, ctl, ring, outs, ring_size, head, arenas,
                                       arena_stride, amgs, amg_stride, spin_limit, lds_bytes,
                                       stream);
}
#endif

}  // namespace rq
