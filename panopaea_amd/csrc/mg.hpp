// Geometric multigrid V-cycle used as the PCG preconditioner z = M r.
// No counterpart in the reference (it plugs into the Preconditioner seam of
// panopaea/src/pcg.rs:4-6). Level l has dims >> l (z stays 1 in 2D) and
// operator A_l = s_l * (nc*z - sum(neighbours)), s_l = dt / 4^l.
#pragma once

#include "pano_internal.hpp"

namespace pano {

template <typename T>
struct MG {
    int levels = 0, nu = 2;
    Dims dims[16];
    double s[16];
    T* z[16];   // level >= 1 (level 0 uses the caller's z)
    T* r[16];   // level >= 1
    T* res = nullptr;         // residual scratch, sized for level 0
    double* fdm_tmp[2] = {nullptr, nullptr};
    double* fdm_v[3] = {nullptr, nullptr, nullptr};     // eigenvectors (column k = mode k), x/y/z
    double* fdm_lam[3] = {nullptr, nullptr, nullptr};   // eigenvalues
    int coarse_n[3] = {1, 1, 1};                         // coarsest extents x, y, z
    bool tb = true;  // 3D f32 levels: pre/post-smoothing as one pass each (PANO_MG_FUSED=0 turns it off)
    bool fdm_split = false;  // 3D coarsest level: passes split over many workgroups (PANO_FDM_SPLIT=0: off)
    // f32 one-pass levels above the coarsest: the prolongation runs inside
    // the post-smoothing (PANO_MG_PRO=0: off), which then reads the
    // pre-smoothed z from w[l] (the down leg's z) while it writes z.
    bool pro = false;
    T* w[16] = {};
};

// Device copy of the eigenbasis of the 1D operator along an axis of n cells
// (lo/hi = 1 for an open side), cached in the context.
pano_status axis_basis(Ctx* c, int n, int lo, int hi, double** V, double** lam);

// Validates dims/levels for the V-cycle; returns PANO_ERR_BAD_SHAPE with a
// message when unsupported.
pano_status mg_check(const Dims& D, int levels);

template <typename T>
pano_status mg_setup(Ctx* c, const Dims& D, int levels, int nu, double dt, MG<T>* m);

// z0 = M r0 (one V-cycle from zero), both in red-black layout at level 0.
// Every kernel returns at once when st->done is set.
// When dot_part is not null, the last two level-0 half-sweeps also write the
// partial sums of z.r per workgroup into dot_part (*dot_np entries).
template <typename T>
pano_status mg_vcycle(Ctx* c, MG<T>* m, T* z0, const T* r0, const SolverState* st, double* dot_part,
                      int* dot_np);

}  // namespace pano
