// Fused advection of a 3D f32 step: density and the three velocity components
// in one pass, every gather served from LDS.
//
// The step advects the density (advect, examples/dec_fluid.rs:173-211, with
// the source reset) and every face component (advect_mac, :213-291, plus the
// buoyancy on z faces) from the same old velocity. The four per-component
// kernels each re-read the velocity they average and gather; here a
// workgroup owns a TX x TY column of cells over a chunk of ZC planes and
// streams through the planes with a ring of four staged planes in LDS. A
// staged plane p holds, for the tile grown by one cell on every side: the
// density as the source leaves it (rho = 1 inside the sphere), the z faces of
// face plane p, and the x and y faces of cell plane p (one extra face at the
// far side). At step z the ring holds planes z-1, z, z+1 and the plane z+2
// being filled.
//
// Every value is computed by the expression of advect_cell_run /
// advect_face_run for one plane (same loads, same sums, same clamps), so the
// results are bit-identical to the per-component kernels. The velocity
// averages read faces at most one cell away, always inside the staged box;
// an interpolation corner lies in the box whenever the departure point moved
// less than one cell along each axis. A corner outside the box (a faster
// flow) is read from global memory instead, lane by lane, so any velocity
// gives the same result.
//
// One barrier per step: the step's gathers read planes z-1 .. z+1, the fill
// writes the slot of plane z+2, which held plane z-2 (last read in step z-1,
// before the previous barrier).
#pragma once

#include "fluid_device.hpp"

namespace pano {

template <int TX, int TY>
struct AdvShape {
    static constexpr int NT = TX * TY;
    static constexpr int RW = TX + 2, RH = TY + 2;  // cell box (tile + 1 on each side)
    static constexpr int XW = TX + 3;                // x faces per box row (x0-1 .. x0+TX+1)
    static constexpr int YH = TY + 3;                // y-face rows (y0-1 .. y0+TY+1)
    static constexpr int O_RHO = 0, O_VZ = RH * RW, O_VX = 2 * RH * RW, O_VY = 2 * RH * RW + RH * XW;
    static constexpr int SLOT = O_VY + YH * RW;
    static constexpr int SLOTS = 4;
    static constexpr int LDS_FLOATS = SLOTS * SLOT;
    // box elements each thread stages per array
    static constexpr int KC = (RH * RW + NT - 1) / NT, KX = (RH * XW + NT - 1) / NT, KY = (YH * RW + NT - 1) / NT;
};

// Workgroup chunks: tiles of TX x TY cells by chunks of zc planes.
struct AdvGrid {
    int gx, gy, zc, nch, nb;
};
template <int TX, int TY>
inline AdvGrid adv_grid(const Dims& D, int zc) {
    AdvGrid g;
    g.gx = (D.w + TX - 1) / TX;
    g.gy = (D.h + TY - 1) / TY;
    g.zc = zc < D.d ? zc : D.d;
    g.nch = (D.d + g.zc - 1) / g.zc;
    g.nb = g.gx * g.gy * g.nch;
    return g;
}

template <int TX, int TY>
struct AdvThread {
    using S = AdvShape<TX, TY>;
    static constexpr int RW = S::RW, RH = S::RH, XW = S::XW, YH = S::YH, NT = S::NT, KC = S::KC, KX = S::KX,
                         KY = S::KY, SLOT = S::SLOT;

    float* lds;
    Dims D;
    Faces<const float> v;   // old velocity (averaged and advected)
    const float* rho;       // old density
    Faces<float> vo;        // advected velocity
    float* rho_o;           // advected density
    Sphere Sph;
    float dt, buoy_c;
    int tid, x0, y0, zs, ze, x, y;
    bool act;
    // staging: per array, the thread's box elements (LDS offset in the slot,
    // in-plane global index, in the domain)
    int ec[KC], gc[KC], ex[KX], gxo[KX], ey[KY], gyo[KY];
    bool vc[KC], vcx[KC], vx_[KX], vy_[KY];
    int cy_[KC], cx_[KC];
    // prefetched plane: rho, vz, vx, vy
    float p_rho[KC], p_vz[KC], p_vx[KX], p_vy[KY];

    __host__ __device__ __forceinline__ void init(float* lds_, const Dims& D_, int bx, int by, int bz, int zc,
                                                  Faces<const float> v_, const float* rho_, Faces<float> vo_,
                                                  float* rho_o_, const Sphere& S_, float dt_, float buoy_c_, int tid_) {
        lds = lds_;
        D = D_;
        v = v_;
        rho = rho_;
        vo = vo_;
        rho_o = rho_o_;
        Sph = S_;
        dt = dt_;
        buoy_c = buoy_c_;
        tid = tid_;
        x0 = bx * TX;
        y0 = by * TY;
        zs = bz * zc;
        ze = zs + zc < D.d ? zs + zc : D.d;
        x = x0 + tid % TX;
        y = y0 + tid / TX;
        act = x < D.w && y < D.h;
#pragma unroll
        for (int k = 0; k < KC; ++k) {
            const int e = tid + k * NT;
            const int r = e / RW, c = e % RW;
            const int yy = y0 - 1 + r, xx = x0 - 1 + c;
            ec[k] = e < RH * RW ? e : 0;
            vc[k] = e < RH * RW && yy >= 0 && yy < D.h && xx >= 0 && xx < D.w;
            vcx[k] = e < RH * RW;
            gc[k] = vc[k] ? yy * D.w + xx : 0;
            cy_[k] = yy;
            cx_[k] = xx;
        }
#pragma unroll
        for (int k = 0; k < KX; ++k) {
            const int e = tid + k * NT;
            const int r = e / XW, c = e % XW;
            const int yy = y0 - 1 + r, xf = x0 - 1 + c;
            ex[k] = e < RH * XW ? e : 0;
            vx_[k] = e < RH * XW && yy >= 0 && yy < D.h && xf >= 0 && xf <= D.w;
            gxo[k] = vx_[k] ? yy * (D.w + 1) + xf : 0;
        }
#pragma unroll
        for (int k = 0; k < KY; ++k) {
            const int e = tid + k * NT;
            const int r = e / RW, c = e % RW;
            const int yf = y0 - 1 + r, xx = x0 - 1 + c;
            ey[k] = e < YH * RW ? e : 0;
            vy_[k] = e < YH * RW && yf >= 0 && yf <= D.h && xx >= 0 && xx < D.w;
            gyo[k] = vy_[k] ? yf * D.w + xx : 0;
        }
    }

    __host__ __device__ __forceinline__ int first_plane() const { return zs; }
    __host__ __device__ __forceinline__ int last_plane() const { return ze; }  // exclusive

    // ---- staging -------------------------------------------------------------------------
    // Loads of plane p into registers. Elements outside the grid load nothing
    // and stage 0 (no gather reads them: every index is clamped into the grid).
    __host__ __device__ __forceinline__ void load_plane(int p) {
        const bool cell_p = p >= 0 && p < D.d, face_p = p >= 0 && p <= D.d;
        const int pcl = cell_p ? p : 0, pf = face_p ? p : 0;
        const float* rp = rho + pcl * D.h * D.w;
        const float* zp = v.z + pf * D.h * D.w;
        const float* xp = v.x + pcl * D.h * (D.w + 1);
        const float* yp = v.y + pcl * (D.h + 1) * D.w;
#pragma unroll
        for (int k = 0; k < KC; ++k) {
            p_rho[k] = (cell_p && vc[k]) ? ldx(rp, gc[k]) : 0.0f;
            p_vz[k] = (face_p && vc[k]) ? ldx(zp, gc[k]) : 0.0f;
        }
#pragma unroll
        for (int k = 0; k < KX; ++k) p_vx[k] = (cell_p && vx_[k]) ? ldx(xp, gxo[k]) : 0.0f;
#pragma unroll
        for (int k = 0; k < KY; ++k) p_vy[k] = (cell_p && vy_[k]) ? ldx(yp, gyo[k]) : 0.0f;
    }
    __host__ __device__ __forceinline__ int slot_of(int p) const { return (p - zs + 1 + 4) & 3; }
    // The prefetched plane p into its slot; the density as the source leaves it.
    __host__ __device__ __forceinline__ void fill_plane(int p) {
        float* s = lds + slot_of(p) * SLOT;
        const bool cell_p = p >= 0 && p < D.d;
#pragma unroll
        for (int k = 0; k < KC; ++k) {
            if (!vcx[k]) continue;
            float r = p_rho[k];
            if (Sph.on && cell_p && vc[k] && in_sphere(Sph, D, p, cy_[k], cx_[k])) r = 1.0f;
            s[S::O_RHO + ec[k]] = r;
            s[S::O_VZ + ec[k]] = p_vz[k];
        }
#pragma unroll
        for (int k = 0; k < KX; ++k)
            if (tid + k * NT < RH * XW) s[S::O_VX + ex[k]] = p_vx[k];
#pragma unroll
        for (int k = 0; k < KY; ++k)
            if (tid + k * NT < YH * RW) s[S::O_VY + ey[k]] = p_vy[k];
    }

    // ---- LDS reads (box-local addressing) -------------------------------------------------
    __host__ __device__ __forceinline__ const float* sl(int p) const { return lds + slot_of(p) * SLOT; }
    __host__ __device__ __forceinline__ float l_rho(int p, int yy, int xx) const {
        return sl(p)[S::O_RHO + (yy - y0 + 1) * RW + (xx - x0 + 1)];
    }
    __host__ __device__ __forceinline__ float l_vz(int p, int yy, int xx) const {
        return sl(p)[S::O_VZ + (yy - y0 + 1) * RW + (xx - x0 + 1)];
    }
    __host__ __device__ __forceinline__ float l_vx(int p, int yy, int xf) const {
        return sl(p)[S::O_VX + (yy - y0 + 1) * XW + (xf - x0 + 1)];
    }
    __host__ __device__ __forceinline__ float l_vy(int p, int yf, int xx) const {
        return sl(p)[S::O_VY + (yf - y0 + 1) * RW + (xx - x0 + 1)];
    }
    // interpolation corners inside the staged box: planes z-1 .. z+1 and the box's rows
    // and columns (i0 <= i1 on every axis)
    __host__ __device__ __forceinline__ bool in_box(int z, const int* gz, const int* gy, const int* gx, int xhi,
                                                    int yhi) const {
        return gz[0] >= z - 1 && gz[1] <= z + 1 && gy[0] >= y0 - 1 && gy[1] <= yhi && gx[0] >= x0 - 1 && gx[1] <= xhi;
    }

    // ---- the four outputs of one position (advect_cell_run / advect_face_run, one plane) ----
    __host__ __device__ __forceinline__ float cell_value(int z) const {
        const float half = 0.5f, two = 2.0f, ndt = -dt;
        const float vx0 = l_vx(z, y, x), vx1 = l_vx(z, y, x + 1);
        const float vy0 = l_vy(z, y, x), vy1 = l_vy(z, y + 1, x);
        const float vz0 = l_vz(z, y, x), vz1 = l_vz(z + 1, y, x);
        int gx[2], gy[2], gz[2];
        float sw, tw, uw;
        const float vcx = (vx0 + vx1) / two;
        const float vcy = (vy0 + vy1) / two;
        const float ppx = ((float)x + half) + ndt * vcx;
        const float ppy = ((float)y + half) + ndt * vcy;
        cell_coord(ppx, D.w, gx[0], gx[1], sw);
        cell_coord(ppy, D.h, gy[0], gy[1], tw);
        const float vcz = (vz0 + vz1) / two;
        const float ppz = ((float)z + half) + ndt * vcz;
        cell_coord(ppz, D.d, gz[0], gz[1], uw);
        float g[8];
        if (in_box(z, gz, gy, gx, x0 + TX, y0 + TY)) {
#pragma unroll
            for (int j = 0; j < 8; ++j) g[j] = l_rho(gz[j >> 2], gy[(j >> 1) & 1], gx[j & 1]);
        } else {
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int zz = gz[j >> 2], yy = gy[(j >> 1) & 1], xx = gx[j & 1];
                g[j] = (Sph.on && in_sphere(Sph, D, zz, yy, xx)) ? 1.0f : ldx(rho, cell_index(D, zz, yy, xx));
            }
        }
        return trilerp(g[0], g[1], g[2], g[3], g[4], g[5], g[6], g[7], sw, tw, uw);
    }

    // x face (xf, y) of cell plane z; xf is x, or w for the thread at x = w-1
    __host__ __device__ __forceinline__ float xface_value(int z, int xf) const {
        const float half = 0.5f, four = 4.0f, zero = 0.0f, ndt = -dt;
        const int xa = imin(xf, D.w - 1), xb = xf > 0 ? xf - 1 : 0;
        const float own = l_vx(z, y, xf);
        const float a0 = l_vy(z, y, xa), a1 = l_vy(z, y + 1, xa), a2 = l_vy(z, y, xb), a3 = l_vy(z, y + 1, xb);
        const float b0 = l_vz(z, y, xa), b1 = l_vz(z + 1, y, xa), b2 = l_vz(z, y, xb), b3 = l_vz(z + 1, y, xb);
        int gx[2], gy[2], gz[2];
        float sw, tw, uw;
        const float vely = (a0 + a1 + a2 + a3) / four;
        const float ppx = ((float)xf + zero) + ndt * own;
        const float ppy = ((float)y + half) + ndt * vely;
        face_coord(ppx, zero, D.w, gx[0], gx[1], sw);
        face_coord(ppy, half, D.h - 1, gy[0], gy[1], tw);
        const float velz = (b0 + b1 + b2 + b3) / four;
        const float ppz = ((float)z + half) + ndt * velz;
        face_coord(ppz, half, D.d - 1, gz[0], gz[1], uw);
        float g[8];
        if (in_box(z, gz, gy, gx, x0 + TX + 1, y0 + TY)) {
#pragma unroll
            for (int j = 0; j < 8; ++j) g[j] = l_vx(gz[j >> 2], gy[(j >> 1) & 1], gx[j & 1]);
        } else {
#pragma unroll
            for (int j = 0; j < 8; ++j) g[j] = ldx(v.x, fx_idx(D, gz[j >> 2], gy[(j >> 1) & 1], gx[j & 1]));
        }
        return trilerp(g[0], g[1], g[2], g[3], g[4], g[5], g[6], g[7], sw, tw, uw);
    }

    // y face (x, yf) of cell plane z; yf is y, or h for the thread at y = h-1
    __host__ __device__ __forceinline__ float yface_value(int z, int yf) const {
        const float half = 0.5f, four = 4.0f, zero = 0.0f, ndt = -dt;
        const int ya = imin(yf, D.h - 1), yb = yf > 0 ? yf - 1 : 0;
        const float a0 = l_vx(z, ya, x), a1 = l_vx(z, ya, x + 1), a2 = l_vx(z, yb, x), a3 = l_vx(z, yb, x + 1);
        const float own = l_vy(z, yf, x);
        const float b0 = l_vz(z, ya, x), b1 = l_vz(z + 1, ya, x), b2 = l_vz(z, yb, x), b3 = l_vz(z + 1, yb, x);
        int gx[2], gy[2], gz[2];
        float sw, tw, uw;
        const float velx = (a0 + a1 + a2 + a3) / four;
        const float ppx = ((float)x + half) + ndt * velx;
        const float ppy = ((float)yf + zero) + ndt * own;
        face_coord(ppx, half, D.w - 1, gx[0], gx[1], sw);
        face_coord(ppy, zero, D.h, gy[0], gy[1], tw);
        const float velz = (b0 + b1 + b2 + b3) / four;
        const float ppz = ((float)z + half) + ndt * velz;
        face_coord(ppz, half, D.d - 1, gz[0], gz[1], uw);
        float g[8];
        if (in_box(z, gz, gy, gx, x0 + TX, y0 + TY + 1)) {
#pragma unroll
            for (int j = 0; j < 8; ++j) g[j] = l_vy(gz[j >> 2], gy[(j >> 1) & 1], gx[j & 1]);
        } else {
#pragma unroll
            for (int j = 0; j < 8; ++j) g[j] = ldx(v.y, fy_idx(D, gz[j >> 2], gy[(j >> 1) & 1], gx[j & 1]));
        }
        return trilerp(g[0], g[1], g[2], g[3], g[4], g[5], g[6], g[7], sw, tw, uw);
    }

    // z face (x, y) of face plane zf (the step's plane z, or d at the last step z = d-1),
    // with the buoyancy of interior and open faces
    __host__ __device__ __forceinline__ float zface_value(int z, int zf) const {
        const float half = 0.5f, four = 4.0f, zero = 0.0f, ndt = -dt;
        const int za = imin(zf, D.d - 1), zb = zf > 0 ? zf - 1 : 0;
        const float a0 = l_vx(za, y, x), a1 = l_vx(za, y, x + 1), a2 = l_vx(zb, y, x), a3 = l_vx(zb, y, x + 1);
        const float b0 = l_vy(za, y, x), b1 = l_vy(za, y + 1, x), b2 = l_vy(zb, y, x), b3 = l_vy(zb, y + 1, x);
        const float own = l_vz(zf, y, x);
        int gx[2], gy[2], gz[2];
        float sw, tw, uw;
        const float velx = (a0 + a1 + a2 + a3) / four;
        const float vely = (b0 + b1 + b2 + b3) / four;
        const float ppx = ((float)x + half) + ndt * velx;
        const float ppy = ((float)y + half) + ndt * vely;
        const float ppz = ((float)zf + zero) + ndt * own;
        face_coord(ppx, half, D.w - 1, gx[0], gx[1], sw);
        face_coord(ppy, half, D.h - 1, gy[0], gy[1], tw);
        face_coord(ppz, zero, D.d, gz[0], gz[1], uw);
        float g[8];
        if (in_box(z, gz, gy, gx, x0 + TX, y0 + TY)) {
#pragma unroll
            for (int j = 0; j < 8; ++j) g[j] = l_vz(gz[j >> 2], gy[(j >> 1) & 1], gx[j & 1]);
        } else {
#pragma unroll
            for (int j = 0; j < 8; ++j) g[j] = ldx(v.z, fz_idx(D, gz[j >> 2], gy[(j >> 1) & 1], gx[j & 1]));
        }
        float val = trilerp(g[0], g[1], g[2], g[3], g[4], g[5], g[6], g[7], sw, tw, uw);
        if (buoy_c != zero) {
            const float rlo = zf > 0 ? l_rho(zf - 1, y, x) : zero;
            const float rhi = zf < D.d ? l_rho(zf, y, x) : zero;
            const bool open_face = (zf > 0 && zf < D.d) || (zf == 0 && (D.bc & 16)) || (zf == D.d && (D.bc & 32));
            if (open_face) val = val + buoy_c * (rlo + rhi);
        }
        return val;
    }

    // ---- one step: outputs of plane z (and the far faces), prefetch/fill of plane z+2 -------
    __host__ __device__ __forceinline__ void prologue() {
        load_plane(zs - 1);
        fill_plane(zs - 1);
        load_plane(zs);
        fill_plane(zs);
        load_plane(zs + 1);
        fill_plane(zs + 1);
    }
    // issue the loads of plane z+2 (consumed by fill_next after the outputs of plane z)
    __host__ __device__ __forceinline__ void prefetch(int z) {
        if (z + 2 <= ze) load_plane(z + 2);
    }
    __host__ __device__ __forceinline__ void outputs(int z) {
        if (!act) return;
        const float c = cell_value(z);
        const float fx = xface_value(z, x);
        const float fy = yface_value(z, y);
        const float fz = zface_value(z, z);
        stx(rho_o, cell_index(D, z, y, x), c);
        stx(vo.x, fx_idx(D, z, y, x), fx);
        stx(vo.y, fy_idx(D, z, y, x), fy);
        stx(vo.z, fz_idx(D, z, y, x), fz);
        if (x == D.w - 1) stx(vo.x, fx_idx(D, z, y, D.w), xface_value(z, D.w));
        if (y == D.h - 1) stx(vo.y, fy_idx(D, z, D.h, x), yface_value(z, D.h));
        if (z == D.d - 1) stx(vo.z, fz_idx(D, D.d, y, x), zface_value(z, D.d));
    }
    __host__ __device__ __forceinline__ void fill_next(int z) {
        if (z + 2 <= ze) fill_plane(z + 2);
    }
};

}  // namespace pano
