// Residual and restriction of a 3D f32 level in one pass: r_c = R (r - A z).
//
// A workgroup owns a tile of TXC x TYC coarse cells (one thread per coarse
// column (I, J)) and a chunk of coarse planes [K0, K1). It streams through
// the fine planes p = 2K0-1 .. 2K1:
//  phase 1: the residual of plane p on the fine cells the tile's restriction
//           reads (pairs I0-1 .. I0+TXC, rows 2J0-1 .. 2J0+2TYC) into one of
//           two residual buffers in LDS (steps alternate), from z planes
//           p-1, p, p+1 held in LDS (four slots: planes p-1 .. p+2, with a
//           one-cell ring around the region; zero outside the grid) and b of
//           plane p in registers;
//  phase 2: each thread restricts plane p along x, then y, for its coarse
//           column (the weights and order of restrict_run) into a ring of
//           four plane values; after plane 2K+2 it combines planes
//           2K-1 .. 2K+2 along z into r_c(K, J, I). Phase 2 also fills the
//           slot of plane p+3 from registers loaded PD steps earlier (z and
//           b are loaded PD planes ahead, into PD register sets used in
//           turn, so a step's loads have PD steps to arrive).
// A barrier separates phase 1 of a step from its phase 2; phase 2 of step p
// and phase 1 of step p+1 need none between them (rr_steps). Each residual is residual_run's
// expression of the same six neighbours and each coarse value restrict_run's,
// so the output has the bits of the two separate kernels. Values the
// restriction weights by zero (rows, columns and planes beyond the grid) are
// never computed: r1 selects them away.
//
// Traffic: z and b read once (plus the tile ring and the two extra planes of
// a chunk, mostly from L2), r_c written; the separate kernels also write and
// re-read the fine residual.
#pragma once

#include <utility>

#include "mg_device.hpp"

namespace pano {

template <int TXC, int TYC>
struct RrShape {
    static constexpr int NT = TXC * TYC;    // threads: one per coarse column
    static constexpr int RW = TXC + 2;      // residual region: pairs I0-1 .. I0+TXC
    static constexpr int RH = 2 * TYC + 2;  // rows 2J0-1 .. 2J0+2TYC
    static constexpr int ZW = RW + 2;       // z region: one more pair / row each side
    static constexpr int ZH = RH + 2;
    static constexpr int ZPLANE = ZW * ZH;  // one colour of one z plane
    static constexpr int RPLANE = RW * RH;
    static constexpr int KZ = (ZPLANE + NT - 1) / NT;  // z positions per thread
    static constexpr int KR = (RPLANE + NT - 1) / NT;  // residual positions per thread
    // colour planes padded to whole thread strides: every thread fills and
    // stores all its positions without a branch (the pad is never read)
    static constexpr int ZS = KZ * NT;
    static constexpr int RS = KR * NT;
    static constexpr int ZSLOTS = 4;  // z planes p-1 .. p+2 at step p
    static constexpr int LDS_FLOATS = ZSLOTS * 2 * ZS + 2 * 2 * RS;
    // Position f -> (row, pair) of a region holding the TXC tile pairs at
    // pr = H .. H+TXC-1: the tile pairs of every row first, TXC consecutive
    // ones per row (with TXC = 32 a half-wave touches 32 consecutive LDS
    // words, 32 distinct banks), then the 2H halo pairs of each row.
    template <int H>
    __host__ __device__ static void pos_of(int f, int rows, int& row, int& pr) {
        if (f < rows * TXC) {
            row = f / TXC;
            pr = H + f % TXC;
        } else {
            const int g = f - rows * TXC, j = g % (2 * H);
            row = g / (2 * H);
            pr = j < H ? j : TXC + j;
        }
    }
};

// Tiles and coarse-plane chunks of one launch: chunks of kc coarse planes,
// halved (down to kmin) while there are fewer than `want` workgroups.
struct RrGrid {
    int gx, gy, kc, nch, nb;
};
template <int TXC, int TYC>
__host__ __device__ inline RrGrid rr_grid(const Dims& Dc, int want, int kmin) {
    RrGrid g;
    g.gx = (Dc.w + TXC - 1) / TXC;
    g.gy = (Dc.h + TYC - 1) / TYC;
    const int tiles = g.gx * g.gy;
    int kc = Dc.d;
    while (kc > kmin && (int64_t)tiles * ((Dc.d + kc - 1) / kc) < want) kc = (kc + 1) / 2;
    g.kc = kc < 1 ? 1 : kc;
    g.nch = (Dc.d + g.kc - 1) / g.kc;
    g.nb = tiles * g.nch;
    return g;
}

template <int TXC, int TYC, int LC, int PD>
struct RrThread {
    using S = RrShape<TXC, TYC>;
    static constexpr int NT = S::NT, RW = S::RW, RH = S::RH, ZW = S::ZW, ZPLANE = S::ZPLANE, RPLANE = S::RPLANE,
                         KZ = S::KZ, KR = S::KR, ZS = S::ZS, RS = S::RS;

    float* zl;  // LDS: 4 slots x 2 colours x ZS
    float* rl;  // LDS: 2 residual planes (steps alternate) x 2 colours x RS
    Dims D, Dc;
    const float* z;
    const float* b;
    OutRange rco;  // r_c
    float s;
    int tid, hw, n2, K0, pf0, pf1;
    // z positions of this thread
    int zoff[KZ];
    int zidx[KZ];  // LDS index of the position in a z colour plane
    bool zin[KZ];
    float zpf[PD][2][KZ];
    bool zok[PD];
    // residual positions
    int rzi[KR];   // index in a z colour plane
    int ridx[KR];  // index in a residual colour plane
    int roff[KR];
    int rbits[KR]; // neighbour bits of x0 (bits 0..3) and x1 (bits 8..11), x and y only
    int ryp[KR];   // row parity
    float bpf[PD][2][KR];
    // coarse column
    int I, J, rp, rr0;
    bool hmI, hpI, hmJ, hpJ, own;
    float ty[4];  // x-y restricted planes p-3 .. p

    __host__ __device__ __forceinline__ void init(float* lds, const Dims& D_, const Dims& Dc_, int bx, int by, int bz,
                                                  int kc, const float* z_, const float* b_, float* rc_, float s_,
                                                  int tid_) {
        zl = lds;
        rl = lds + S::ZSLOTS * 2 * ZS;
        D = D_;
        Dc = Dc_;
        z = z_;
        b = b_;
        rco = out_range(rc_, Dc.d * Dc.h * Dc.w);
        s = s_;
        tid = tid_;
        const int wh = D.w >> 1;
        hw = D.h * wh;
        n2 = D.d * hw;
        const int I0 = bx * TXC, J0 = by * TYC;
        K0 = bz * kc;
        const int K1 = K0 + kc < Dc.d ? K0 + kc : Dc.d;
        pf0 = 2 * K0 - 1;
        pf1 = 2 * K1;  // last fine plane (inclusive)
#pragma unroll
        for (int k = 0; k < KZ; ++k) {
            // positions past the region: the pad beyond ZPLANE
            const int e = tid + k * NT;
            int row, pr;
            S::template pos_of<2>(e < ZPLANE ? e : 0, S::ZH, row, pr);
            zidx[k] = e < ZPLANE ? row * ZW + pr : e;
            const int y = 2 * J0 - 2 + row, m = I0 - 2 + pr;
            zin[k] = e < ZPLANE && y >= 0 && y < D.h && m >= 0 && m < wh;
            zoff[k] = zin[k] ? y * wh + m : 0;
        }
#pragma unroll
        for (int k = 0; k < KR; ++k) {
            // positions past the region compute a residual of position 0
            // into the pad; positions outside the grid one from zero z
            // values and b at a valid address, which the restriction never
            // weighs
            const int e = tid + k * NT;
            int row, pr;
            S::template pos_of<1>(e < RPLANE ? e : 0, RH, row, pr);
            ridx[k] = e < RPLANE ? row * RW + pr : e;
            const int y = 2 * J0 - 1 + row, m = I0 - 1 + pr;
            rzi[k] = (row + 1) * ZW + pr + 1;
            roff[k] = y >= 0 && y < D.h && m >= 0 && m < wh ? y * wh + m : 0;
            const int yb = ((y > 0) << 2) | ((y < D.h - 1) << 3);
            rbits[k] = (yb | (m > 0) | 2) | ((yb | 1 | ((m < wh - 1) << 1)) << 8);
            ryp[k] = y & 1;
        }
        const int tx = tid % TXC, tyy = tid / TXC;
        I = I0 + tx;
        J = J0 + tyy;
        rp = tx + 1;
        rr0 = 2 * tyy;
        hmI = I > 0;
        hpI = I < Dc.w - 1;
        hmJ = J > 0;
        hpJ = J < Dc.h - 1;
        own = I < Dc.w && J < Dc.h;
        ty[0] = ty[1] = ty[2] = ty[3] = 0.0f;
    }
    __host__ __device__ __forceinline__ int slot(int p) const { return (p - pf0 + 1) % S::ZSLOTS; }
    // the residual plane buffer of step p
    __host__ __device__ __forceinline__ float* rbuf(int p) const { return rl + ((p - pf0) & 1) * 2 * RS; }
    __host__ __device__ __forceinline__ const float* plane_ptr(const float* f, int p) const {
        return f + (p < 0 ? 0 : (p >= D.d ? D.d - 1 : p)) * hw;
    }
    // z of plane p into register set Q (loads at clamped, valid addresses;
    // the fill zeroes what lies outside the grid)
    template <int Q>
    __host__ __device__ __forceinline__ void load_z(int p) {
        zok[Q] = p >= 0 && p < D.d;
        const float* p0 = plane_ptr(z, p);
        const float* p1 = p0 + n2;
#pragma unroll
        for (int k = 0; k < KZ; ++k) {
            zpf[Q][0][k] = ldx(p0, zoff[k]);
            zpf[Q][1][k] = ldx(p1, zoff[k]);
        }
    }
    template <int Q>
    __host__ __device__ __forceinline__ void fill_z(int p) {
        float* base = zl + slot(p) * 2 * ZS;
#pragma unroll
        for (int k = 0; k < KZ; ++k) {
            const bool ok = zok[Q] && zin[k];
            base[zidx[k]] = ok ? zpf[Q][0][k] : 0.0f;
            base[ZS + zidx[k]] = ok ? zpf[Q][1][k] : 0.0f;
        }
    }
    template <int Q>
    __host__ __device__ __forceinline__ void load_b(int p) {
        const float* p0 = plane_ptr(b, p);
        const float* p1 = p0 + n2;
#pragma unroll
        for (int k = 0; k < KR; ++k) {
            bpf[Q][0][k] = ldx(p0, roff[k]);
            bpf[Q][1][k] = ldx(p1, roff[k]);
        }
    }
    // b of plane pf0+Q and z of plane pf0+3+Q into set Q, for Q < PD: the
    // order of the loop's own loads (b of a plane, then z three planes on), so
    // that the loop's first waits are those of every later step
    template <int Q>
    __host__ __device__ __forceinline__ void load_ahead() {
        if constexpr (Q < PD) {
            load_b<Q>(pf0 + Q);
            load_z<Q>(pf0 + 3 + Q);
            load_ahead<Q + 1>();
        }
    }
    // z planes pf0-1 .. pf0+2 into LDS (all loads issued before the first
    // fill); z of planes pf0+3 .. and b of planes pf0 .. into the PD
    // register sets
    __host__ __device__ __forceinline__ void prologue() {
        float v[S::ZSLOTS][2][KZ];
        bool ok[S::ZSLOTS];
#pragma unroll
        for (int j = 0; j < S::ZSLOTS; ++j) {
            load_z<0>(pf0 - 1 + j);
            ok[j] = zok[0];
#pragma unroll
            for (int k = 0; k < KZ; ++k) {
                v[j][0][k] = zpf[0][0][k];
                v[j][1][k] = zpf[0][1][k];
            }
        }
#pragma unroll
        for (int j = 0; j < S::ZSLOTS; ++j) {
            zok[0] = ok[j];
#pragma unroll
            for (int k = 0; k < KZ; ++k) {
                zpf[0][0][k] = v[j][0][k];
                zpf[0][1][k] = v[j][1][k];
            }
            fill_z<0>(pf0 - 1 + j);
        }
        load_ahead<0>();
    }
    // residual of fine plane p into the residual plane rdst (both colours),
    // from the z planes zc (p), zm (p-1), zp (p+1) in LDS and b of plane p
    __host__ __device__ __forceinline__ void residual_plane(float* rdst, const float* zc, const float* zm,
                                                            const float* zp, int p, const float (&bb)[2][KR]) const {
        const int zb = ((p > 0) << 4) | ((p < D.d - 1) << 5);
        const int cz = p & 1;
#pragma unroll
        for (int k = 0; k < KR; ++k) {
            const int c0 = cz ^ ryp[k];  // colour of the pair's even cell x0
            const int c1 = 1 - c0;
            const int e = rzi[k];
            const float* o0 = zc + c1 * ZS;  // x0's neighbours in the plane, and x1
            const float* o1 = zc + c0 * ZS;  // x1's neighbours in the plane, and x0
            float n0[6], n1[6];
            n0[0] = o0[e - 1];
            n0[1] = o0[e];
            n0[2] = o0[e - ZW];
            n0[3] = o0[e + ZW];
            n0[4] = zm[c1 * ZS + e];
            n0[5] = zp[c1 * ZS + e];
            n1[0] = o1[e];
            n1[1] = o1[e + 1];
            n1[2] = o1[e - ZW];
            n1[3] = o1[e + ZW];
            n1[4] = zm[c0 * ZS + e];
            n1[5] = zp[c0 * ZS + e];
            const float z0 = o1[e], z1 = o0[e];
            const float b0 = c0 ? bb[1][k] : bb[0][k], b1 = c0 ? bb[0][k] : bb[1][k];
            const int h0 = (rbits[k] & 0xF) | zb, h1 = ((rbits[k] >> 8) & 0xF) | zb;
            const float a0 = s * ((float)ncount_bits<3>(h0, D.bc) * z0 - nsum<float, 3>(n0));
            const float a1 = s * ((float)ncount_bits<3>(h1, D.bc) * z1 - nsum<float, 3>(n1));
            rdst[c0 * RS + ridx[k]] = b0 - a0;
            rdst[c1 * RS + ridx[k]] = b1 - a1;
        }
    }
    // fine plane p of the residual plane rsrc restricted along x, then y, for
    // the thread's coarse column
    __host__ __device__ __forceinline__ float restrict_xy(const float* rsrc, int p) const {
        float t[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int row = rr0 + q;                 // fine row 2J-1+q
            const int ce = (2 * J - 1 + q + p) & 1;  // colour of the fine x = 2I
            const float* pe = rsrc + ce * RS + row * RW + rp;
            const float* po = rsrc + (1 - ce) * RS + row * RW + rp;
            t[q] = r1(po[-1], pe[0], po[0], pe[1], hmI, hpI);
        }
        return r1(t[0], t[1], t[2], t[3], hmJ, hpJ);
    }
    // residual of plane p (b in register set Q) into LDS (nothing for planes
    // outside the grid); then b of plane p+PD into set Q
    template <int Q>
    __host__ __device__ __forceinline__ void phase1(int p) {
        if (p >= 0 && p < D.d)
            residual_plane(rbuf(p), zl + slot(p) * 2 * ZS, zl + slot(p - 1) * 2 * ZS, zl + slot(p + 1) * 2 * ZS,
                           p, bpf[Q]);
        load_b<Q>(p + PD);
    }
    // plane p restricted along x and y into the ring; r_c(K) once plane
    // 2K+2 is in; then the slot of plane p+3 (that of plane p-1, last read by
    // phase 1 of this step) from set Q and z of plane p+3+PD into set Q
    template <int Q>
    __host__ __device__ __forceinline__ void phase2(int p) {
        ty[0] = ty[1];
        ty[1] = ty[2];
        ty[2] = ty[3];
        ty[3] = p >= 0 && p < D.d ? restrict_xy(rbuf(p), p) : 0.0f;
        // r_c(K) for p = 2K+2 (an always-issued store, dropped on other steps)
        const bool st = p >= 2 * K0 + 2 && !(p & 1);
        const int K = st ? p / 2 - 1 : K0;
        st_if(rco, own ? cidx<LC>(Dc, K, J, I) : 0, r1(ty[0], ty[1], ty[2], ty[3], K > 0, K < Dc.d - 1), st && own);
        fill_z<Q>(p + 3);
        load_z<Q>(p + 3 + PD);
    }
};

// The 2D form: one plane, so a workgroup loads its z tile with the ring and
// b once, computes the residual region in LDS and restricts it (x, then y:
// restrict_run in 2D). Same positions as the 3D form; the residual is
// residual_run's 2D expression (four neighbours).
template <int TXC, int TYC, int LC>
struct Rr2Thread : RrThread<TXC, TYC, LC, 1> {
    using B = RrThread<TXC, TYC, LC, 1>;
    using B::KR;
    using B::KZ;
    using B::RS;
    using B::ZS;
    using B::ZW;
    static constexpr int LDS_FLOATS = 2 * ZS + 2 * RS;

    __host__ __device__ __forceinline__ void init(float* lds, const Dims& D_, const Dims& Dc_, int bx, int by,
                                                  const float* z_, const float* b_, float* rc_, float s_, int tid_) {
        B::init(lds, D_, Dc_, bx, by, 0, 1, z_, b_, rc_, s_, tid_);
        B::rl = lds + 2 * ZS;
    }
    // z (both colours) into LDS from its base, zero outside the grid; b into registers
    __host__ __device__ __forceinline__ void load_fill() {
        B::template load_z<0>(0);
        B::template load_b<0>(0);
#pragma unroll
        for (int k = 0; k < KZ; ++k) {
            const bool ok = B::zin[k];
            B::zl[B::zidx[k]] = ok ? B::zpf[0][0][k] : 0.0f;
            B::zl[ZS + B::zidx[k]] = ok ? B::zpf[0][1][k] : 0.0f;
        }
    }
    __host__ __device__ __forceinline__ void residual() {
        const float* zc = B::zl;
#pragma unroll
        for (int k = 0; k < KR; ++k) {
            const int c0 = B::ryp[k];  // colour of the pair's even cell x0 (plane 0)
            const int c1 = 1 - c0;
            const int e = B::rzi[k];
            const float* o0 = zc + c1 * ZS;
            const float* o1 = zc + c0 * ZS;
            float n0[4], n1[4];
            n0[0] = o0[e - 1];
            n0[1] = o0[e];
            n0[2] = o0[e - ZW];
            n0[3] = o0[e + ZW];
            n1[0] = o1[e];
            n1[1] = o1[e + 1];
            n1[2] = o1[e - ZW];
            n1[3] = o1[e + ZW];
            const float z0 = o1[e], z1 = o0[e];
            const float b0 = c0 ? B::bpf[0][1][k] : B::bpf[0][0][k], b1 = c0 ? B::bpf[0][0][k] : B::bpf[0][1][k];
            const int h0 = B::rbits[k] & 0xF, h1 = (B::rbits[k] >> 8) & 0xF;
            const float a0 = B::s * ((float)ncount_bits<2>(h0, B::D.bc) * z0 - nsum<float, 2>(n0));
            const float a1 = B::s * ((float)ncount_bits<2>(h1, B::D.bc) * z1 - nsum<float, 2>(n1));
            B::rl[c0 * RS + B::ridx[k]] = b0 - a0;
            B::rl[c1 * RS + B::ridx[k]] = b1 - a1;
        }
    }
    __host__ __device__ __forceinline__ void restrict_store() {
        const float v = B::restrict_xy(B::rl, 0);
        st_if(B::rco, B::own ? cidx<LC>(B::Dc, 0, B::J, B::I) : 0, v, B::own);
    }
};

// Coarse tile of the 2D form (32 x 8 coarse cells, 256 threads).
constexpr int kRr2TX = 32, kRr2TY = 8;

// One workgroup of the 2D form: tile (bx, by) of coarse cells.
template <int TXC, int TYC, int LC>
__device__ __forceinline__ void rr2_workgroup(float* lds, const Dims& D, const Dims& Dc, int bx, int by,
                                              const float* z, const float* b, float* rc, float s) {
    Rr2Thread<TXC, TYC, LC> th;
    th.init(lds, D, Dc, bx, by, z, b, rc, s, threadIdx.x);
    th.load_fill();
    __syncthreads();
    th.residual();
    __syncthreads();
    th.restrict_store();
}

// Steps p .. p+PD-1 (up to p1): a barrier, phase 2 of step p (reads the
// residual phase 1 of step p wrote), then phase 1 of step p+1. Phase 1 of step
// p+1 reads z planes p .. p+2, filled by phase 2 of step p-1 or earlier (a
// barrier behind), and writes the other residual buffer, last read by phase
// 2 of step p-1; phase 2 of step p fills the slot of plane p-1, last read by
// phase 1 of step p. So one barrier per step.
template <int PD, typename TH, int... Q>
__device__ __forceinline__ void rr_steps(TH& th, int& p, int p1, std::integer_sequence<int, Q...>) {
    ((p <= p1 ? (__syncthreads(), th.template phase2<Q>(p),
                 (p + 1 <= p1 ? th.template phase1<(Q + 1) % PD>(p + 1) : (void)0), ++p)
              : p),
     ...);
}

// One workgroup: tile (bx, by) of coarse columns, coarse planes bz*kc ..
template <int TXC, int TYC, int LC, int PD>
__device__ __forceinline__ void rr_workgroup(float* lds, const Dims& D, const Dims& Dc, int bx, int by, int bz,
                                             int kc, const float* z, const float* b, float* rc, float s) {
    RrThread<TXC, TYC, LC, PD> th;
    th.init(lds, D, Dc, bx, by, bz, kc, z, b, rc, s, threadIdx.x);
    th.prologue();
    __syncthreads();
    int p = th.pf0;
    th.template phase1<0>(p);
    while (p <= th.pf1) rr_steps<PD>(th, p, th.pf1, std::make_integer_sequence<int, PD>{});
}

}  // namespace pano
