// Temporally blocked red-black Gauss-Seidel on 3D f32 levels: NS consecutive
// half-sweeps (colours CF, CF^1, CF, ...) in one pass over the field.
//
// A workgroup owns a tile of TXP x-pairs by TY rows and a chunk of planes
// [zs, ze). It streams through the planes: at step t half-sweep h updates
// plane t-h, so half-sweep h reads plane t-h-1 after half-sweep h-1 finished
// it (one step earlier) and plane t-h+1 before half-sweep h+1 reaches it
// (half-sweeps run in order h = 0..NS-1 within a step). In-plane, half-sweep h
// covers the tile grown by NS-1-h cells (a halo of NS cells is loaded), and in
// z the chunk grown by NS-1-h planes, so the tile's cells end with the values
// of NS full-grid half-sweeps. Halo cells are recomputed by each neighbour
// with the same arithmetic (gs_value of the same six neighbour values in the
// same order), so every cell gets the bits of the unfused half-sweeps.
//
// LDS holds SLOTS = NS + 2 planes of the tile + halo in the red-black
// compacted layout of the field (plane slot, colour, row, pair); cells outside
// the grid hold 0, which is the value a missing neighbour contributes to the
// sum. Plane z sits in slot (z - t0 + 1) mod SLOTS; steps run in rounds of
// SLOTS, so every LDS offset is a compile-time constant. At step t the ring
// holds planes t-NS .. t+1: plane t+1 is filled at the start of step t, into
// the slot of plane t-NS-1. Half-sweep 0 of step t reads plane t+1 only as
// the z+1 neighbour of its own positions, which the same thread has just
// filled (a thread's z neighbours are always at its own positions), so no
// barrier separates the fill from that read; other threads read the plane
// in-plane from step t+1 on. b stays in registers: a ring of the NS planes in
// flight, both colours.
//
// Thread positions are split by row parity (rows 2i+q, q = 0, 1), so the x
// parity of the cell a half-sweep updates is the same across a position slot:
// the activity test and the reciprocal 1/(s*nc) come from per-position
// tables selected by that parity. With an even number of slots the parity of
// step t - t0 = J is that of t0 + J, so the workgroup runs one of two copies
// of the loop (t0 even or odd) in which every such choice is a constant.
//
// Within a step a thread reads other threads' LDS values only as written in
// earlier steps (tb_step), so the kernel puts one barrier after each step:
// [coarse prologue |] prologue | begin + half-sweeps 0 .. NS-1 + end | begin
// of the next step ...
#pragma once

#include <utility>

#include "mg_device.hpp"

namespace pano {

template <int NS, int TXP, int TY, int NTHREADS>
struct TbShape {
    static_assert(NS % 2 == 0 && TY % 2 == 0, "even half-sweep count and tile height");
    static constexpr int HP = NS / 2;                // halo pairs in x (2*HP = NS cells)
    static constexpr int PW = TXP + 2 * HP;          // pairs per LDS row
    static constexpr int PH = TY + 2 * NS;           // LDS rows
    static constexpr int PLANE = PW * PH;            // cells of one colour of one plane
    static constexpr int SLOTS = NS + 2;             // resident planes t-NS .. t+1 at step t
    static constexpr int NT = NTHREADS;              // threads per workgroup
    static constexpr int CPOS = (PH / 2) * PW;       // positions of one row parity
    static constexpr int KC = (CPOS + NT - 1) / NT;  // position slots per thread and parity
    static constexpr int KP = 2 * KC;
    static constexpr int LDS_FLOATS = SLOTS * 2 * PLANE;
    // With the prolongation in the fill: the coarse cells under the tile +
    // halo and one more on each side, CW x CH of one coarse plane (the coarse
    // cell of pair m, row y is (m, y >> 1)).
    static constexpr int CW = PW + 2, CH = PH / 2 + 2, CPL = CW * CH;
    static constexpr int CKN = (CPL + NT - 1) / NT;  // coarse cells per thread
    // Position f of a row parity -> (row pair rp, pair pr). The first
    // (PH/2)*TXP positions are the tile's own pairs, TXP consecutive ones per
    // row pair, so with TXP = 32 every half-wave reads and writes 32
    // consecutive LDS words (32 distinct banks) and global addresses; the
    // 2*HP halo pairs of each row pair follow.
    static constexpr int MAINP = (PH / 2) * TXP;
    __host__ __device__ static void pos_of(int f, int& rp, int& pr) {
        if (f < MAINP) {
            rp = f / TXP;
            pr = HP + f % TXP;
        } else {
            const int g = f - MAINP, j = g % (2 * HP);
            rp = g / (2 * HP);
            pr = j < HP ? j : TXP + j;
        }
    }
};

// Tiles and plane chunks of one launch: chunks of zc planes (zc >= zmin,
// halved while the launch has fewer than `want` workgroups).
struct TbGrid {
    int gx, gy, zc, nch, nb;
};
template <int TXP, int TY>
__host__ __device__ inline TbGrid tb_grid(const Dims& D, int want, int zmin) {
    TbGrid g;
    g.gx = ((D.w >> 1) + TXP - 1) / TXP;
    g.gy = (D.h + TY - 1) / TY;
    int tiles = g.gx * g.gy, zc = D.d;
    while (zc > zmin && (int64_t)tiles * ((D.d + zc - 1) / zc) < want) zc = (zc + 1) / 2;
    if (zc < zmin) zc = zmin < D.d ? zmin : D.d;
    g.zc = zc;
    g.nch = (D.d + zc - 1) / zc;
    g.nb = tiles * g.nch;
    return g;
}

// A register value the optimiser may not trace back to its memory (keeps a
// select between two table entries a select of values, not a load from a
// selected address).
__device__ __forceinline__ float tb_opaque(float v) {
    asm volatile("" : "+v"(v));
    return v;
}
__host__ inline float tb_opaque(float v) { return v; }
// An int the optimiser may not treat as loop-invariant: what is derived from
// it is recomputed where it is used instead of being held in registers.
__device__ __forceinline__ int tb_opaque(int v) {
    asm volatile("" : "+v"(v));
    return v;
}
__host__ inline int tb_opaque(int v) { return v; }

// f(std::integral_constant<int, i>) for i = 0 .. N-1: a loop whose index is a
// compile-time constant in the body from the start (table entries it selects
// are constant indices, never a select between two entries).
template <typename F, int... I>
__host__ __device__ __forceinline__ void tb_static_for(F&& f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, I>{}), ...);
}

// The coarse level of the prolongation fused into the one-pass smoothers
// (layout PL, dims Dc; a 2D level has Dc.d = 1): a coarse plane's region of
// CW x CH cells around the tile + halo (region origin Ib, Jb in coarse x,
// y), loaded with clamped coordinates, and Q of a position's two cells from
// a region in LDS. Position bits in act (8..10 the edge flags y, x-, x+;
// 11..22 the LDS offset of the position's coarse cell in a region).
template <int NS, int TXP, int TY, int NTHREADS, int PL>
struct TbPro {
    using S = TbShape<NS, TXP, TY, NTHREADS>;
    static constexpr int CW = S::CW, CPL = S::CPL, CKN = S::CKN, KC = S::KC, KP = S::KP, NT = S::NT, HP = S::HP,
                         CPOS = S::CPOS;
    const float* zco;
    Dims Dc;
    int chw, cn2, Ib, Jb;  // plane size, colour offset, region origin

    __host__ __device__ __forceinline__ void init(const float* zco_, const Dims& Dc_, int bx, int by, int tid,
                                                  int (&act)[KP]) {
        zco = zco_;
        Dc = Dc_;
        chw = PL == REDBLACK ? Dc.h * (Dc.w >> 1) : Dc.h * Dc.w;
        cn2 = Dc.d * chw;
        const int m0 = bx * TXP, y0 = by * TY;
        Ib = m0 - HP - 1;
        Jb = ((y0 - NS) >> 1) - 1;  // y0 - NS is even
#pragma unroll
        for (int k = 0; k < KP; ++k) {
            const int q = k / KC, f = tid + (k % KC) * NT;
            int rp, pr;
            S::pos_of(f < CPOS ? f : 0, rp, pr);
            const int y = y0 - NS + 2 * rp + q, m = m0 - HP + pr;
            const int I = y >> 1, J = (y & 1) ? I + 1 : I - 1;  // p1 along y
            const int ey = J < 0 || J >= Dc.h, el = m == 0, er = m + 1 >= Dc.w;
            act[k] |= ((ey | (el << 1) | (er << 2)) << 8) | (((rp + 1) * CW + pr + 1) << 11);
        }
    }
    // the thread's region cells of coarse plane K (clamped to the level)
    __host__ __device__ __forceinline__ void load(int K, float (&v)[CKN], int tid) const {
        K = K < 0 ? 0 : (K >= Dc.d ? Dc.d - 1 : K);
        const int tv = tb_opaque(tid);  // offsets recomputed per load (no registers held)
#pragma unroll
        for (int i = 0; i < CKN; ++i) {
            const int e = tv + i * NT < CPL ? tv + i * NT : 0;
            int J = Jb + e / CW, I = Ib + e % CW;
            J = J < 0 ? 0 : (J >= Dc.h ? Dc.h - 1 : J);
            I = I < 0 ? 0 : (I >= Dc.w ? Dc.w - 1 : I);
            if constexpr (PL == REDBLACK)
                v[i] = zco[((I + J + K) & 1) * cn2 + K * chw + J * (Dc.w >> 1) + (I >> 1)];
            else
                v[i] = zco[K * chw + J * Dc.w + I];
        }
    }
    __host__ __device__ __forceinline__ void store(float* dst, const float (&v)[CKN], int tid) const {
#pragma unroll
        for (int i = 0; i < CKN; ++i)
            if (tid + i * NT < CPL) dst[tid + i * NT] = v[i];
    }
    // Q of the coarse plane whose region is at C, both cells of every position.
    // Positions k (row 2rp) and k + KC (row 2rp + 1) share pair and row pair,
    // so their coarse row J0 = y >> 1 (x-interpolated once) and x weights;
    // J1 is the row above (even y) / below (odd y).
    __host__ __device__ __forceinline__ void q_of(const int (&act)[KP], const float* C, float (&qv)[2][KP]) const {
        const float w34 = 0.75f, w14 = 0.25f, one = 1.0f, zero = 0.0f;
#pragma unroll
        for (int k = 0; k < KC; ++k) {
            const int ae = tb_opaque(act[k]), ao = tb_opaque(act[k + KC]);  // weights recomputed per call
            const int fe = (ae >> 8) & 7, fo = (ao >> 8) & 7;
            const float* r0 = C + (ae >> 11);
            const float wl0 = (fe & 2) ? one : w34, wl1 = (fe & 2) ? zero : w14;
            const float wr0 = (fe & 4) ? one : w34, wr1 = (fe & 4) ? zero : w14;
            const float am = r0[-CW - 1], bm = r0[-CW], cm = r0[-CW + 1];  // row J0 - 1
            const float a0 = r0[-1], b0 = r0[0], c0 = r0[1];                // row J0
            const float ap = r0[CW - 1], bp = r0[CW], cp = r0[CW + 1];      // row J0 + 1
            const float pl0 = wl0 * b0 + wl1 * a0, pr0 = wr0 * b0 + wr1 * c0;
            const float plm = wl0 * bm + wl1 * am, prm = wr0 * bm + wr1 * cm;
            const float plp = wl0 * bp + wl1 * ap, prp = wr0 * bp + wr1 * cp;
            const float ye0 = (fe & 1) ? one : w34, ye1 = (fe & 1) ? zero : w14;
            const float yo0 = (fo & 1) ? one : w34, yo1 = (fo & 1) ? zero : w14;
            qv[0][k] = ye0 * pl0 + ye1 * plm;
            qv[1][k] = ye0 * pr0 + ye1 * prm;
            qv[0][k + KC] = yo0 * pl0 + yo1 * plp;
            qv[1][k + KC] = yo0 * pr0 + yo1 * prp;
        }
    }
};

// Per-thread state and phases of one workgroup. xin: the field before the
// half-sweeps (unused with ZERO_X: the sweeps start from 0); xout: the result
// (may equal xin only with ZERO_X: other workgroups still read xin's halo
// while this one stores). r: right-hand side. With DOT, acc[0] and acc[1]
// collect sum(v*r) over the tile's cells of the last two half-sweeps.
// Needs D.d >= 2 (every cell then has a z neighbour, so nc >= 1).
//
// PL >= 0: the half-sweeps start from xin + P zco instead of xin, P the
// trilinear prolongation from the coarse level zco (layout PL, dims Dc, set
// by init_pro), with the arithmetic of prolong_run / k_prolong3_lds, so the
// result has the bits of the prolongation kernel followed by the smoother.
// The prolongation of a fine cell (z, y, x = 2m + side) is
//   wz0 * Q_side(K0) + wz1 * Q_side(K1),
//   Q_side(K) = wy0 * px_side(K, J0) + wy1 * px_side(K, J1),
//   px_side(K, J) = w0 * c(K, J, m) + w1 * c(K, J, m -/+ 1),
// with (K0, K1) = (z >> 1, the other coarse plane next to z) and the edge
// weights of p1. Q depends on the fine plane only through K, so the thread
// keeps Q of the two coarse planes M, M+1 that fine plane z needs (M =
// floor((z - 1) / 2)) and computes Q of a new coarse plane every second step,
// from a copy of that plane's CW x CH region in LDS. The region is loaded
// with clamped coarse coordinates: at a grid edge the neighbour the weights
// ignore (weight 0) then holds the cell's own value, as in the kernel.
template <int NS, int TXP, int TY, int NTHREADS, int CF, bool ZERO_X, bool DOT, int PL = -1>
struct TbThread {
    using S = TbShape<NS, TXP, TY, NTHREADS>;
    static constexpr int PW = S::PW, PH = S::PH, PLANE = S::PLANE, SLOTS = S::SLOTS, NT = S::NT, KC = S::KC,
                         KP = S::KP, HP = S::HP, CPOS = S::CPOS;
    static constexpr bool PRO = PL >= 0;
    static_assert(!(PRO && ZERO_X), "the prolongation adds to a given field");
    static_assert(!PRO || 3 * S::CPL <= 2 * PLANE, "three coarse planes fit one plane slot");
    static constexpr int CW = S::CW, CPL = S::CPL, CKN = S::CKN;
    // x ring, then (PRO) one coarse plane region, written at the start of odd
    // steps and read after their first barrier
    static constexpr int LDS_FLOATS = S::LDS_FLOATS + (PRO ? CPL : 0);

    float* lds;
    Dims D;
    const float* xin;
    float* xout;
    const float* r;
    float s;
    int tid, zs, ze, t0, hw, n2;
    int e_[KP];              // LDS index of the position (row * PW + pair)
    unsigned goff[KP];       // in-plane element offset y*wh + m (0 outside the grid)
    // bits 0..7: h + 4p, half-sweep h updates the cell when its x parity is p;
    // PRO: bits 8..10 the edge flags (y, x-, x+) of its prolongation, bits
    // 11..22 the LDS offset of its coarse cell in a region
    int act[KP];
    bool ind[KP], tile[KP];  // inside the grid; inside the tile
    float rin[KP][2], rbd[KP][2];  // 1/(s*nc) for x parity p: interior plane / one z side missing
    float xpf[2][KP], bpf[2][KP], bq[NS][2][KP];
    bool xok, bok;  // whether the plane held in xpf / bpf is used (else it reads as 0)
    double acc[2];
    TbPro<NS, TXP, TY, NTHREADS, PL> pro;  // PRO: the coarse level
    float cpf[CKN];  // the thread's region cells of the next coarse plane to store
    // Q of coarse planes K0, K1 of the next fine plane to fill, for the cells
    // x = 2m, 2m+1 (K0 = z >> 1, K1 the other coarse plane next to z)
    float q0[2][KP], q1[2][KP];

    static constexpr int mod(int a) { return ((a % SLOTS) + SLOTS) % SLOTS; }

    // rinv: the level's reciprocals 1/(s*nc), nc = 0..7 (read here only)
    __host__ __device__ __forceinline__ void init(float* lds_, const Dims& D_, int bx, int by, int bz, int zc,
                                                  const float* xin_, float* xout_, const float* r_, float s_,
                                                  const float* rinv, int tid_) {
        lds = lds_;
        D = D_;
        xin = xin_;
        xout = xout_;
        r = r_;
        s = s_;
        tid = tid_;
        const int wh = D.w >> 1;
        hw = D.h * wh;
        n2 = D.d * hw;
        const int m0 = bx * TXP, y0 = by * TY;
        zs = bz * zc;
        ze = zs + zc < D.d ? zs + zc : D.d;
        t0 = zs - NS + 1 > 0 ? zs - NS + 1 : 0;
        acc[0] = acc[1] = 0.0;
#pragma unroll
        for (int k = 0; k < KP; ++k) {
            const int q = k / KC, f = tid + (k % KC) * NT;
            const bool valid = f < CPOS;
            const int ff = valid ? f : 0;
            int rp, pr;
            S::pos_of(ff, rp, pr);
            const int row = 2 * rp + q;
            const int y = y0 - NS + row, m = m0 - HP + pr;
            e_[k] = row * PW + pr;
            ind[k] = valid && y >= 0 && y < D.h && m >= 0 && m < wh;
            tile[k] = ind[k] && row >= NS && row < NS + TY && pr >= HP && pr < HP + TXP;
            goff[k] = ind[k] ? (unsigned)(y * wh + m) : 0u;
            int a = 0;
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                const int xl = 2 * pr + p, x = 2 * m + p;
                for (int h = 0; h < NS; ++h)
                    if (ind[k] && row >= h + 1 && row < PH - 1 - h && xl >= h + 1 && xl < 2 * PW - 1 - h)
                        a |= 1 << (h + 4 * p);
                const int hxy = (x > 0) | ((x < D.w - 1) << 1) | ((y > 0) << 2) | ((y < D.h - 1) << 3);
                const int nxy = __builtin_popcount((unsigned)((hxy | D.bc) & 0xF));
                rin[k][p] = rinv[nxy + 2];
                rbd[k][p] = rinv[nxy + 1];
            }
            act[k] = a;
        }
    }
    // PRO: the coarse level (after init).
    __host__ __device__ __forceinline__ void init_pro(const float* zco, const Dims& Dc, int bx, int by) {
        pro.init(zco, Dc, bx, by, tid, act);
    }
    __host__ __device__ __forceinline__ void load_c(int K, float (&v)[CKN]) const { pro.load(K, v, tid); }
    __host__ __device__ __forceinline__ void store_c(float* dst, const float (&v)[CKN]) const {
        pro.store(dst, v, tid);
    }
    __host__ __device__ __forceinline__ void q_of(const float* C, float (&qv)[2][KP]) const { pro.q_of(act, C, qv); }
    __host__ __device__ __forceinline__ void q_swap() {
#pragma unroll
        for (int k = 0; k < KP; ++k) {
            const float a = q0[0][k], b = q0[1][k];
            q0[0][k] = q1[0][k];
            q0[1][k] = q1[1][k];
            q1[0][k] = a;
            q1[1][k] = b;
        }
    }

    // x of plane z into xpf, b of plane z into v. The loads are unconditional
    // (positions outside the grid and unused planes read a valid clamped
    // address) and the values that must read as 0 are zeroed where they are
    // consumed (fill, end_step): straight-line loads let each consumer wait
    // for its own loads only, not for prefetches issued after them.
    __host__ __device__ __forceinline__ const float* plane_ptr(const float* f, int z) const {
        return f + (z < 0 ? 0 : (z >= D.d ? D.d - 1 : z)) * hw;
    }
    __host__ __device__ __forceinline__ void load_x(int z) {
        xok = !ZERO_X && z >= 0 && z < D.d && z >= zs - NS && z < ze + NS;
        if (ZERO_X) return;
        const float* p0 = plane_ptr(xin, z);
        const float* p1 = p0 + n2;
#pragma unroll
        for (int k = 0; k < KP; ++k) {
            xpf[0][k] = p0[goff[k]];
            xpf[1][k] = p1[goff[k]];
        }
    }
    __host__ __device__ __forceinline__ bool load_b(int z, float (&v)[2][KP]) const {
        const float* p0 = plane_ptr(r, z);
        const float* p1 = p0 + n2;
#pragma unroll
        for (int k = 0; k < KP; ++k) {
            v[0][k] = p0[goff[k]];
            v[1][k] = p1[goff[k]];
        }
        return z >= 0 && z < D.d && z >= zs - NS + 1 && z < ze + NS - 1;
    }
    // xpf (plane z, of parity ZP) into slot SL; PRO: plus its prolongation,
    // from Q of its coarse planes K0, K1 in q0, q1
    template <int SL, int ZP>
    __host__ __device__ __forceinline__ void fill(int z) {
        float* base = lds + SL * 2 * PLANE;
        float wz0 = 0.0f, wz1 = 0.0f;
        if constexpr (PRO) {
            const int I = z >> 1, J = (z & 1) ? I + 1 : I - 1;  // p1 along z
            const bool ez = J < 0 || J >= pro.Dc.d;
            wz0 = ez ? 1.0f : 0.75f;
            wz1 = ez ? 0.0f : 0.25f;
        }
#pragma unroll
        for (int k = 0; k < KP; ++k) {
            if ((k % KC) * NT + tid < CPOS) {
                const bool ok = xok && ind[k];
                float v0 = xpf[0][k], v1 = xpf[1][k];
                if constexpr (PRO) {
                    const float pl = wz0 * q0[0][k] + wz1 * q1[0][k];  // cell x = 2m
                    const float pr = wz0 * q0[1][k] + wz1 * q1[1][k];  // cell x = 2m + 1
                    // colour of x = 2m (row parity k / KC): a constant after unrolling
                    const bool c0 = ((k / KC + ZP) & 1) != 0;
                    v0 = v0 + (c0 ? pr : pl);
                    v1 = v1 + (c0 ? pl : pr);
                }
                base[e_[k]] = ok ? v0 : 0.0f;
                base[PLANE + e_[k]] = ok ? v1 : 0.0f;
            }
        }
    }
    __host__ __device__ __forceinline__ int first_step() const { return t0; }
    __host__ __device__ __forceinline__ int last_step() const { return ze + NS - 2; }

    // PRO, before prologue (a barrier between): the regions of coarse planes
    // A .. A+2 (A = floor((t0 - 2) / 2)) into the last x slot, which no step
    // fills before step t0 + SLOTS - 3 (reads of it from the prologue end
    // before that step's barriers); A+3 into cpf (stored into the coarse
    // slot for the first odd step); x of plane t0-1 into xpf.
    __host__ __device__ __forceinline__ void prologue_coarse() {
        const int A = (t0 - 2) >> 1;
        float c0[CKN], c1[CKN], c2[CKN];
        load_c(A, c0);
        load_c(A + 1, c1);
        load_c(A + 2, c2);
        load_c(A + 3, cpf);
        load_x(t0 - 1);
        float* sc = lds + (SLOTS - 1) * 2 * PLANE;
        store_c(sc, c0);
        store_c(sc + CPL, c1);
        store_c(sc + 2 * CPL, c2);
    }
    // planes t0-1, t0 into slots 0, 1; b of plane t0; prefetch x of t0+1, b of
    // t0+1. PAR = t0 & 1: the ring has an even number of slots, so the parity
    // of every step is a compile-time constant within a round.
    template <int PAR>
    __host__ __device__ __forceinline__ void prologue() {
        if constexpr (PRO) {
            // A = t0/2 - 1 (t0 even) or (t0 - 3)/2 (t0 odd); (K0, K1) of plane z
            // is (z >> 1, z >> 1 + 1) for odd z, (z >> 1, z >> 1 - 1) for even z
            const float* sc = lds + (SLOTS - 1) * 2 * PLANE;
            if constexpr (PAR) {
                q_of(sc + CPL, q0);  // plane t0 - 1: (A+1, A)
                q_of(sc, q1);
                fill<0, 0>(t0 - 1);
                q_of(sc + 2 * CPL, q1);  // plane t0: (A+1, A+2)
            } else {
                q_of(sc, q0);  // plane t0 - 1: (A, A+1)
                q_of(sc + CPL, q1);
                fill<0, 1>(t0 - 1);
                q_swap();  // plane t0: (A+1, A)
            }
        } else {
            load_x(t0 - 1);
            fill<0, PAR ^ 1>(t0 - 1);
        }
        load_x(t0);
        fill<1, PAR>(t0);
        if constexpr (PRO && !PAR) {
            // step t0 fills plane t0 + 1 with (A+1, A+2)
            q_of(lds + (SLOTS - 1) * 2 * PLANE + 2 * CPL, q1);
        }
#pragma unroll
        for (int j = 1; j < NS; ++j)
#pragma unroll
            for (int k = 0; k < KP; ++k) bq[j][0][k] = bq[j][1][k] = 0.0f;
        const bool ok0 = load_b(t0, bq[0]);
#pragma unroll
        for (int k = 0; k < KP; ++k) {
            const bool ok = ok0 && ind[k];
            bq[0][0][k] = ok ? bq[0][0][k] : 0.0f;
            bq[0][1][k] = ok ? bq[0][1][k] : 0.0f;
        }
        load_x(t0 + 1);
        bok = load_b(t0 + 1, bpf);
        if constexpr (PRO && PAR) {
            // the first step is odd: its coarse region (A+3) goes in now
            store_c(lds + S::LDS_FLOATS, cpf);
            load_c((t0 + 5) >> 1, cpf);
        }
    }
    // Step t (t - t0 = J mod SLOTS, TP = t & 1): plane t+1 (prefetched) into
    // its slot, which only this thread's own reads of step t touch; prefetch
    // plane t+2. PRO: for odd t plane t+1 has (K0, K1) = ((t+1)/2, (t-1)/2),
    // the two of the step before swapped; coarse plane (t+3)/2 is in the
    // coarse slot (stored by step t-1), and in stage 1 the step computes
    // Q((t+3)/2) into q1, which the fill at step t dropped: the next (even)
    // step fills plane t+2 with ((t+1)/2, (t+3)/2). Even steps store the
    // prefetched region of the next odd step and prefetch the one after.
    template <int J, int TP>
    __host__ __device__ __forceinline__ void begin_step(int t) {
        if constexpr (PRO && !TP) {
            // the region of coarse plane (t+4)/2 for the next (odd) step: the
            // slot's last reads (step t-1) are a barrier behind, its next
            // (step t+1) a barrier ahead
            store_c(lds + S::LDS_FLOATS, cpf);
            load_c((t + 6) >> 1, cpf);
        }
        if constexpr (PRO && TP) q_swap();
        fill<mod(J + 2), TP ^ 1>(t + 1);
        load_x(t + 2);
    }
    // half-sweep H on plane t-H (t - t0 = J mod SLOTS, TP = t & 1)
    template <int H, int J, int TP>
    __host__ __device__ __forceinline__ void stage(int t) {
        if constexpr (PRO && H == 1 && TP) q_of(lds + S::LDS_FLOATS, q1);  // the region step t-1 stored
        const int z = t - H;
        const int lo = zs - (NS - 1 - H) > 0 ? zs - (NS - 1 - H) : 0;
        const int hi = ze + (NS - 1 - H) < D.d ? ze + (NS - 1 - H) : D.d;
        if (z < lo || z >= hi) return;
        constexpr int c = (CF + H) & 1;
        constexpr int sz = mod(J - H + 1), sm = mod(J - H), sp = mod(J - H + 2);
        const float* xo = lds + (sz * 2 + (1 - c)) * PLANE;
        const float* xm = lds + (sm * 2 + (1 - c)) * PLANE;
        const float* xp = lds + (sp * 2 + (1 - c)) * PLANE;
        float* xc = lds + (sz * 2 + c) * PLANE;
        // both z neighbours exist or their sides are open: interior reciprocals
        const bool zin = (z > 0 || (D.bc & 16)) && (z < D.d - 1 || (D.bc & 32));
        const bool own = DOT && H >= NS - 2 && z >= zs && z < ze;
        tb_static_for([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            constexpr int p = (c + k / KC + TP + H) & 1;  // x parity of the colour-c cell (row parity k / KC)
            if (act[k] & (1 << (H + 4 * p))) {
                const int e = e_[k];
                const float rv = bq[H][c][k];
                const float a = zin ? rin[k][p] : rbd[k][p];  // 1/(s*nc) of this cell
                float v;
                if constexpr (ZERO_X && H == 0) {
                    // first half-sweep from z = 0: every neighbour still holds +0, whose
                    // sum is +0 (red_init_run's s*0)
                    v = (rv + s * 0.0f) * a;
                } else {
                    const float* xs = xo + (p - 1);  // x-1 of the cell of parity p
                    float n[6];
                    n[0] = xs[e];       // x-1
                    n[1] = xs[e + 1];   // x+1
                    n[2] = xo[e - PW];  // y-1
                    n[3] = xo[e + PW];  // y+1
                    n[4] = xm[e];       // z-1
                    n[5] = xp[e];       // z+1
                    v = (rv + s * nsum<float, 3>(n)) * a;  // gs_value, nc >= 1
                }
                xc[e] = v;
                if (own && tile[k]) acc[H - (NS - 2)] += (double)v * (double)rv;
            }
        }, std::make_integer_sequence<int, KP>{});
    }
    // store plane t-NS+1 (final after half-sweep NS-1 of step t); advance the
    // b ring; prefetch b of plane t+2
    template <int J>
    __host__ __device__ __forceinline__ void end_step(int t) {
        const int zo = t - NS + 1;
        if (zo >= zs && zo < ze) {
            const float* base = lds + mod(J - NS + 2) * 2 * PLANE;
            float* o0 = xout + zo * hw;
            float* o1 = o0 + n2;
#pragma unroll
            for (int k = 0; k < KP; ++k) {
                if (tile[k]) {
                    o0[goff[k]] = base[e_[k]];
                    o1[goff[k]] = base[PLANE + e_[k]];
                }
            }
        }
#pragma unroll
        for (int j = NS - 1; j > 0; --j)
#pragma unroll
            for (int k = 0; k < KP; ++k) {
                bq[j][0][k] = bq[j - 1][0][k];
                bq[j][1][k] = bq[j - 1][1][k];
            }
#pragma unroll
        for (int k = 0; k < KP; ++k) {
            const bool ok = bok && ind[k];
            bq[0][0][k] = ok ? bpf[0][k] : 0.0f;
            bq[0][1][k] = ok ? bpf[1][k] : 0.0f;
        }
        bok = load_b(t + 2, bpf);
    }
};

// The 2D form: one plane, so a workgroup loads its tile + halo (both colours)
// into LDS once, runs the NS half-sweeps there (a barrier after each) and
// stores the tile. Same positions, activity tables and arithmetic as the 3D
// form with four neighbours (nsum over x-, x+, y-, y+). Every cell of a 2D
// level with an even width has an x neighbour, so nc >= 1.
// PL >= 0: the half-sweeps start from xin + P zco (the prolongation from the
// coarse level zco, layout PL, dims Dc, with the arithmetic of prolong_run):
// init stores the coarse region of the tile into LDS and keeps x; after a
// barrier, fill_pro adds Q (the y and x interpolation) to each cell of x and
// fills the plane.
template <int NS, int TXP, int TY, int NTHREADS, int CF, bool ZERO_X, bool DOT, int PL = -1>
struct Tb2Thread {
    using S = TbShape<NS, TXP, TY, NTHREADS>;
    static constexpr int PW = S::PW, PH = S::PH, PLANE = S::PLANE, NT = S::NT, KC = S::KC, KP = S::KP, HP = S::HP,
                         CPOS = S::CPOS;
    static constexpr bool PRO = PL >= 0;
    static_assert(!(PRO && ZERO_X), "the prolongation adds to a given field");
    static constexpr int LDS_FLOATS = 2 * PLANE + (PRO ? S::CPL : 0);  // both colours, then (PRO) the region

    float* lds;
    Dims D;
    float s;
    int tid, n2;
    int e_[KP];
    unsigned goff[KP];
    int act[KP];
    bool ind[KP], tile[KP];
    float rin[KP][2];
    float bv[2][KP];
    double acc[2];
    TbPro<NS, TXP, TY, NTHREADS, PL> pro;  // PRO: the coarse level
    float xk[2][KP];                       // PRO: x until fill_pro

    __host__ __device__ __forceinline__ void init(float* lds_, const Dims& D_, int bx, int by, const float* xin,
                                                  const float* r, float s_, const float* rinv, int tid_,
                                                  const float* zco = nullptr, const Dims& Dc = Dims{}) {
        lds = lds_;
        D = D_;
        s = s_;
        tid = tid_;
        const int wh = D.w >> 1;
        n2 = D.h * wh;
        const int m0 = bx * TXP, y0 = by * TY;
        acc[0] = acc[1] = 0.0;
#pragma unroll
        for (int k = 0; k < KP; ++k) {
            const int q = k / KC, f = tid + (k % KC) * NT;
            const bool valid = f < CPOS;
            const int ff = valid ? f : 0;
            int rp, pr;
            S::pos_of(ff, rp, pr);
            const int row = 2 * rp + q;
            const int y = y0 - NS + row, m = m0 - HP + pr;
            e_[k] = row * PW + pr;
            ind[k] = valid && y >= 0 && y < D.h && m >= 0 && m < wh;
            tile[k] = ind[k] && row >= NS && row < NS + TY && pr >= HP && pr < HP + TXP;
            goff[k] = ind[k] ? (unsigned)(y * wh + m) : 0u;
            int a = 0;
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                const int xl = 2 * pr + p, x = 2 * m + p;
                for (int h = 0; h < NS; ++h)
                    if (ind[k] && row >= h + 1 && row < PH - 1 - h && xl >= h + 1 && xl < 2 * PW - 1 - h)
                        a |= 1 << (h + 4 * p);
                const int hxy = (x > 0) | ((x < D.w - 1) << 1) | ((y > 0) << 2) | ((y < D.h - 1) << 3);
                rin[k][p] = rinv[__builtin_popcount((unsigned)((hxy | D.bc) & 0xF))];
            }
            act[k] = a;
        }
        if constexpr (PRO) pro.init(zco, Dc, bx, by, tid, act);
        // all loads first (clamped, valid addresses), then the LDS fill and b
        float xv[2][KP], b0[2][KP], cv[S::CKN];
#pragma unroll
        for (int k = 0; k < KP; ++k) {
            xv[0][k] = xv[1][k] = 0.0f;
            if (!ZERO_X) {
                xv[0][k] = xin[goff[k]];
                xv[1][k] = xin[n2 + goff[k]];
            }
            b0[0][k] = r[goff[k]];
            b0[1][k] = r[n2 + goff[k]];
        }
        if constexpr (PRO) pro.load(0, cv, tid);
#pragma unroll
        for (int k = 0; k < KP; ++k) {
            const bool ok = ind[k];
            if constexpr (PRO) {
                xk[0][k] = xv[0][k];
                xk[1][k] = xv[1][k];
            } else if ((k % KC) * NT + tid < CPOS) {
                lds[e_[k]] = ok ? xv[0][k] : 0.0f;
                lds[PLANE + e_[k]] = ok ? xv[1][k] : 0.0f;
            }
            bv[0][k] = ok ? b0[0][k] : 0.0f;
            bv[1][k] = ok ? b0[1][k] : 0.0f;
        }
        if constexpr (PRO) pro.store(lds + 2 * PLANE, cv, tid);
    }
    // PRO, after a barrier behind init: x + P zco into the plane. The cell x =
    // 2m of row y has colour y & 1, the position's row parity k / KC.
    __host__ __device__ __forceinline__ void fill_pro() {
        float qv[2][KP];
        pro.q_of(act, lds + 2 * PLANE, qv);
#pragma unroll
        for (int k = 0; k < KP; ++k) {
            if ((k % KC) * NT + tid < CPOS) {
                const bool ok = ind[k];
                const bool c0 = (k / KC) != 0;
                const float v0 = xk[0][k] + (c0 ? qv[1][k] : qv[0][k]);
                const float v1 = xk[1][k] + (c0 ? qv[0][k] : qv[1][k]);
                lds[e_[k]] = ok ? v0 : 0.0f;
                lds[PLANE + e_[k]] = ok ? v1 : 0.0f;
            }
        }
    }
    // half-sweep H (colour (CF + H) & 1)
    template <int H>
    __host__ __device__ __forceinline__ void stage() {
        constexpr int c = (CF + H) & 1;
        const float* xo = lds + (1 - c) * PLANE;
        float* xc = lds + c * PLANE;
#pragma unroll
        for (int k = 0; k < KP; ++k) {
            const int p = (c + k / KC) & 1;  // x parity of the colour-c cell (row parity k / KC)
            if (act[k] & (1 << (H + 4 * p))) {
                const int e = e_[k];
                const float rv = bv[c][k];
                const float a0 = tb_opaque(rin[k][0]), a1 = tb_opaque(rin[k][1]);
                float v;
                if constexpr (ZERO_X && H == 0) {
                    v = (rv + s * 0.0f) * (p ? a1 : a0);  // all neighbours still +0
                } else {
                    float n[6];
                    n[0] = xo[e - 1 + p];  // x-1
                    n[1] = xo[e + p];      // x+1
                    n[2] = xo[e - PW];     // y-1
                    n[3] = xo[e + PW];     // y+1
                    n[4] = n[5] = 0.0f;
                    v = (rv + s * nsum<float, 2>(n)) * (p ? a1 : a0);  // gs_value, nc >= 1
                }
                xc[e] = v;
                if (DOT && H >= NS - 2 && tile[k]) acc[H - (NS - 2)] += (double)v * (double)rv;
            }
        }
    }
    // the tile's cells, both colours
    __host__ __device__ __forceinline__ void store(float* xout) const {
#pragma unroll
        for (int k = 0; k < KP; ++k)
            if (tile[k]) {
                xout[goff[k]] = lds[e_[k]];
                xout[n2 + goff[k]] = lds[PLANE + e_[k]];
            }
    }
};

// Configurations by level size (measured on 256^3, 128^3, 64^3): tall tiles
// on large levels (less halo work; two 768-thread workgroups per CU with 79
// KiB of LDS and at most 80 VGPRs each), small tiles and short plane chunks on
// small levels, so every CU still gets workgroups. WANT/ZMIN: workgroups
// aimed for and the shortest chunk (tb_grid).
struct TbLarge {  // >= 2^23 cells
    static constexpr int TXP = 32, TY = 32, NT = 768, WPE = 6, WANT = 512, ZMIN = 8;
};
struct TbMedium {  // >= 2^20 cells
    static constexpr int TXP = 32, TY = 32, NT = 768, WPE = 6, WANT = 1024, ZMIN = 4;
};
struct TbSmall {
    static constexpr int TXP = 32, TY = 8, NT = 320, WPE = 4, WANT = 2048, ZMIN = 2;
};
// 2D levels: 32 x-pairs by 32 rows, one plane per workgroup
struct Tb2D {
    static constexpr int TXP = 32, TY = 32, NT = 768, WPE = 8;
};
// The batched engine's other 2D tiles (its autotuner picks one of Tb2D and
// these): 32 x 16, 16 x 16 and 64 x 16 x-pairs by rows, one thread per
// position pair (NT >= (TY/2 + 4) * (TXP + 4)).
struct Tb2D_32x16 {
    static constexpr int TXP = 32, TY = 16, NT = 448, WPE = 8;
};
struct Tb2D_16x16 {
    static constexpr int TXP = 16, TY = 16, NT = 256, WPE = 8;
};
struct Tb2D_64x16 {
    static constexpr int TXP = 64, TY = 16, NT = 832, WPE = 8;
};
constexpr int kTb2Shapes = 4;  // Tb2D, Tb2D_32x16, Tb2D_16x16, Tb2D_64x16
// Calls f with the 2D tile configuration of index k (0 .. kTb2Shapes-1).
template <class F>
inline auto with_tb2(int k, F&& f) {
    switch (k) {
        case 1: return f(Tb2D_32x16{});
        case 2: return f(Tb2D_16x16{});
        case 3: return f(Tb2D_64x16{});
        default: return f(Tb2D{});
    }
}
constexpr int kTbConfigs = 4;  // index: 0 large, 1 medium, 2 small (3D), 3 2D

template <class C>
inline TbGrid tb_grid_of(const Dims& D) {
    return tb_grid<C::TXP, C::TY>(D, C::WANT, C::ZMIN);
}
// Calls f with the configuration of a level of dims D.
template <class F>
inline auto with_tb(const Dims& D, F&& f) {
    const int64_t n = (int64_t)D.d * D.h * D.w;
    if (n >= (int64_t(1) << 23)) return f(TbLarge{});
    if (n >= (int64_t(1) << 20)) return f(TbMedium{});
    return f(TbSmall{});
}
inline int tb_config(const Dims& D) {
    if (D.d <= 1) return 3;
    const int64_t n = (int64_t)D.d * D.h * D.w;
    return n >= (int64_t(1) << 23) ? 0 : (n >= (int64_t(1) << 20) ? 1 : 2);
}
// 2D tiles of a level: x-pair tiles by row tiles (TbGrid with one chunk)
inline TbGrid tb2_grid(const Dims& D, int txp = Tb2D::TXP, int ty = Tb2D::TY) {
    TbGrid g;
    g.gx = ((D.w >> 1) + txp - 1) / txp;
    g.gy = (D.h + ty - 1) / ty;
    g.zc = 1;
    g.nch = 1;
    g.nb = g.gx * g.gy;
    return g;
}
inline int tb_blocks(const Dims& D) {
    if (D.d <= 1) return tb2_grid(D).nb;
    return with_tb(D, [&](auto cfg) { return tb_grid_of<decltype(cfg)>(D).nb; });
}
// Whether an f32 level smooths in one pass, both legs (the single-grid
// V-cycle and the batched engine decide alike, so a grid gets the same bits
// in both): when enabled, with nu = 2, and when the post-smoothing's z.r
// partials (2 per workgroup) fit the buffer sized for the red-black tiles.
inline bool tb_on(bool enabled, int nu, const Dims& D) {
    return enabled && nu == 2 && tb_blocks(D) <= grid3_rb(D).nb;
}

// One step (t - t0 = J mod SLOTS, TP = t & 1): begin, half-sweeps 0..NS-1,
// end, then one barrier. Within a step every LDS value a thread reads from
// another thread's positions was written in an earlier step: half-sweep h
// reads the in-plane neighbours of plane t-h, of the colour half-sweep h-1
// wrote at step t-1, and its z neighbours (planes t-h-1, t-h+1) only at its
// own positions; the fill and the store touch its own positions; a slot is
// refilled two steps after its last in-plane read. So the step's phases need
// no barrier between them (the CPU emulation runs each step thread by
// thread, in both thread orders, and gets the same bits). With the
// prolongation, the coarse region that every thread of an odd step reads is
// stored by the even step before it (or the prologue), a barrier ahead.
template <int J, int TP, typename TH, int... H>
__device__ __forceinline__ void tb_step(TH& th, int t, std::integer_sequence<int, H...>) {
    th.template begin_step<J, TP>(t);
    (th.template stage<H, J, TP>(t), ...);
    th.template end_step<J>(t);
    __syncthreads();
}
// Steps t .. t+SLOTS-1 (up to t1), then t advanced past them; PAR = t0 & 1.
template <int NS, int PAR, typename TH, int... J>
__device__ __forceinline__ void tb_round(TH& th, int& t, int t1, std::integer_sequence<int, J...>) {
    ((t <= t1 ? (tb_step<J, (PAR + J) & 1>(th, t, std::make_integer_sequence<int, NS>{}), ++t) : t), ...);
}
template <int PAR, typename TH>
__device__ __forceinline__ void tb_run(TH& th) {
    th.template prologue<PAR>();
    __syncthreads();
    const int t1 = th.last_step();
    int t = th.first_step();
    while (t <= t1) tb_round<4, PAR>(th, t, t1, std::make_integer_sequence<int, TH::SLOTS>{});
}
// One workgroup of the one-pass smoother (four half-sweeps) on tile
// (bx, by, bz) of a level; acc: its tile's z.r sums of the last two
// half-sweeps (with DOT). PL >= 0: the sweeps start from xin + P zco (zco:
// the coarse level, layout PL, dims Dc).
template <int TXP, int TY, int NT, int CF, bool ZERO_X, bool DOT, int PL = -1>
__device__ __forceinline__ void tb_workgroup(float* lds, const Dims& D, int bx, int by, int bz, int zc,
                                             const float* xin, float* xout, const float* r, float s,
                                             const float* rinv, double* acc, const float* zco = nullptr,
                                             const Dims& Dc = Dims{}) {
    using TH = TbThread<4, TXP, TY, NT, CF, ZERO_X, DOT, PL>;
    TH th;
    th.init(lds, D, bx, by, bz, zc, xin, xout, r, s, rinv, threadIdx.x);
    if constexpr (TH::PRO) {
        th.init_pro(zco, Dc, bx, by);
        th.prologue_coarse();
        __syncthreads();
    }
    if (th.first_step() & 1) tb_run<1>(th);
    else tb_run<0>(th);
    acc[0] = th.acc[0];
    acc[1] = th.acc[1];
}

// One workgroup of the 2D one-pass smoother on tile (bx, by) of a level.
// PL >= 0: the sweeps start from xin + P zco (zco: the coarse level, layout
// PL, dims Dc).
template <int TXP, int TY, int NT, int CF, bool ZERO_X, bool DOT, int PL = -1>
__device__ __forceinline__ void tb2_workgroup(float* lds, const Dims& D, int bx, int by, const float* xin, float* xout,
                                              const float* r, float s, const float* rinv, double* acc,
                                              const float* zco = nullptr, const Dims& Dc = Dims{}) {
    using TH = Tb2Thread<4, TXP, TY, NT, CF, ZERO_X, DOT, PL>;
    TH th;
    th.init(lds, D, bx, by, xin, r, s, rinv, threadIdx.x, zco, Dc);
    __syncthreads();
    if constexpr (TH::PRO) {
        th.fill_pro();
        __syncthreads();
    }
    th.template stage<0>();
    __syncthreads();
    th.template stage<1>();
    __syncthreads();
    th.template stage<2>();
    __syncthreads();
    th.template stage<3>();
    __syncthreads();
    th.store(xout);
    acc[0] = th.acc[0];
    acc[1] = th.acc[1];
}

}  // namespace pano
